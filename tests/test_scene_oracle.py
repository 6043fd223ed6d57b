"""The scene-helper oracle (oracle/scene_oracle.py) pinned against published values and closed
forms, and the host-side parts of the device generators (the GF(2) jump of the random stream, the
argument checks) against it.  CPU only."""
import ctypes
import math

import numpy as np
import pytest

import scene_oracle as SO

from nuclearmpm_amd import generators, lib, scenes, viz


def test_nc_rand_matches_marsaglia_published_outputs():
    # xor128 with the seeds 123456789, 362436069, 521288629, 88675123 (Marsaglia, J. Stat. Soft. 2003)
    assert list(SO.NcRand().u32(5)) == [3701687786, 458299110, 2500872618, 3633119408, 516391518]


def test_nc_rand_float_conversion():
    r = SO.NcRand()
    u = r.uniform(10000)
    assert u.dtype == np.float32 and (u >= 0).all() and (u <= 1).all()
    # u32 -> float rounds to nearest, then an exact power-of-two scale
    w = SO.NcRand().u32(10000)
    assert np.array_equal(u, (w.astype(np.float64).astype(np.float32) / np.float32(2 ** 32)).astype(np.float32))


@pytest.mark.parametrize("k", [0, 1, 5, 511, 512, 513, 4096, 100003])
def test_library_rand_jump_equals_sequential_stream(k):
    st = (ctypes.c_uint32 * 4)(*SO.SEED)
    assert lib().mpm_rand_advance(st, ctypes.c_uint64(k)) == 0
    r = SO.NcRand()
    for _ in range(k):
        r.next_u32()
    assert tuple(st) == r.state


def test_library_rand_jump_composes():
    a = generators.NcRand((1, 2, 3, 4))
    a.advance(1 << 40)
    a.advance(12345)
    b = generators.NcRand((1, 2, 3, 4))
    b.advance((1 << 40) + 12345)
    assert a.state == b.state
    r = SO.NcRand(a.state)
    a.advance(3)
    for _ in range(3):
        r.next_u32()
    assert a.state == r.state


def test_linspaced_properties():
    for n, lo, hi in [(50, 0.4, 0.6), (25, 0.3, 0.5), (7, -0.9, 0.1), (64, 0.7, 0.05), (1, 0.3, 0.9), (2, 0.1, 0.2)]:
        a = SO.linspaced(n, lo, hi)
        assert a.dtype == np.float32 and len(a) == n
        assert a[-1] == np.float32(hi)
        if n > 1:
            assert a[0] == np.float32(lo)
            d = np.diff(a.astype(np.float64))
            step = (float(np.float32(hi)) - float(np.float32(lo))) / (n - 1)
            assert np.allclose(d, step, rtol=0, atol=4 * np.finfo(np.float32).eps)
    # the product's host generator follows the same rule
    for n, lo, hi in [(50, 0.4, 0.6), (64, 0.7, 0.05), (33, -0.2, -0.8)]:
        assert np.array_equal(SO.linspaced(n, lo, hi), scenes.linspaced(n, lo, hi))


def test_cube_layout():
    c = SO.cube(3, [2, 3, 4], [0.1, 0.2, 0.3], [0.5, 0.6, 0.7])
    assert c.shape == (24, 3)
    assert np.array_equal(c[:4, 2], SO.linspaced(4, 0.3, 0.7))  # last axis fastest
    assert (c[:12, 0] == np.float32(0.1)).all() and (c[12:, 0] == np.float32(0.5)).all()
    assert np.array_equal(SO.cube(2, [50, 50], [0.4] * 2, [0.6] * 2), scenes.cube(2, 50, 0.4, 0.6))


def test_sample_uses_the_stream_in_order():
    r = SO.NcRand()
    pts = SO.sample(2, np.array([[0.5, 0.5], [0.2, 0.7]]), r, per_center=3)
    u = SO.NcRand().uniform(12).reshape(6, 2)
    want = (u * np.float32(2) - np.float32(1)) * np.float32(0.08)
    want[:3] += np.float32([0.5, 0.5])
    want[3:] += np.float32([0.2, 0.7])
    assert np.array_equal(pts, want.astype(np.float32))
    assert (np.abs(pts[:3] - np.float32([0.5, 0.5])) <= 0.08 + 1e-7).all()


def test_gyroid_closed_forms():
    k, t = 0.5, 0.3
    g = SO.gyroid(k, t, np.array([[0, 0, 0], [k / 4, 0, 0], [0, k / 4, 0], [0, 0, k / 4]], np.float32))
    assert g[0] == np.float32(-t)
    for v in g[1:]:
        assert v == pytest.approx(1 - t, abs=1e-6)
    rng = np.random.default_rng(0)
    p = rng.random((50, 3)).astype(np.float32) * np.float32(0.25)
    k2 = 0.25  # period 0.25: shifting by exactly one period (representable) leaves the value
    base = SO.gyroid(k2, 0.0, p)
    shifted = SO.gyroid(k2, 0.0, (p + np.float32(0.25)).astype(np.float32))
    assert np.allclose(base, shifted, atol=2e-6)


def test_projection_oracle_properties():
    c, s, C, S = SO.projection_constants()
    assert abs(float(c) - math.cos(math.radians(28))) < 1e-7 and abs(float(S) - math.sin(math.radians(32))) < 1e-7
    out = SO.pt_3d_to_2d(np.zeros((1, 3), np.float32))
    assert np.array_equal(out, np.float32([[0.5, 0.5]]))
    rng = np.random.default_rng(1)
    p = rng.random((100, 3)).astype(np.float32)
    q = SO.pt_3d_to_2d(p)
    x = q[:, 0].astype(np.float64) - 0.5
    z = p[:, 2].astype(np.float64) * float(c) - p[:, 0].astype(np.float64) * float(s)
    assert np.allclose(x ** 2 + z ** 2, p[:, 0].astype(np.float64) ** 2 + p[:, 2].astype(np.float64) ** 2, atol=1e-6)
    # the centred variant is the plain one applied to p - 0.5
    assert np.array_equal(SO.pt_3d_to_2d(p, centered=True), SO.pt_3d_to_2d((p - np.float32(0.5)).astype(np.float32)))


def test_igl_grid_layout_and_package_helper():
    g = SO.igl_grid([2, 3])
    assert g.tolist() == [[0, 0], [1, 0], [0, 0.5], [1, 0.5], [0, 1], [1, 1]]
    for res in ([2, 3], [4, 3, 2], [5]):
        assert np.array_equal(viz.grid_points(res), SO.igl_grid(res))


def test_generator_argument_checks_without_gpu():
    L = lib()
    cube = generators._cube_struct((4, 0.1, 0.2), 2)
    arr = (type(cube) * 1)(cube)
    assert L.mpm_gen_cubes(4, 1, arr, 0, 16, None, 0) != 0
    assert L.mpm_gen_cubes(2, 1, arr, 0, 15, None, 0) != 0  # n must be the lattice total
    assert L.mpm_gen_cubes(2, 1, arr, 2, 16, None, 0) != 0
    assert L.mpm_gen_cubes(2, 0, None, 0, 0, None, 0) == 0
    st = (ctypes.c_uint32 * 4)(*SO.SEED)
    assert L.mpm_gen_rand(-1, st, None, 0) != 0
    assert L.mpm_gen_sample(2, 1, None, 10, 0.08, st, 0, None, 0) != 0
    bad = generators._cube_struct((0, 0.1, 0.2), 3)
    n = ctypes.c_int64()
    assert L.mpm_gen_gyroid(ctypes.byref(bad), 1.0, 0.0, 0.0, 0, None, 0, ctypes.byref(n), 0) != 0
    assert L.mpm_project_points(-1, None, 0, None, 0) != 0
