"""CPU-side checks of the boundary: the C-ABI library loads and exports every entry point declared
in include/nclr_mpm.h, the options struct matches the header, and the Python mirror and scene
generators behave (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import nuclearmpm_amd as nm
from nuclearmpm_amd import MaterialModel, Particle, Particles, scenes
from nuclearmpm_amd._lib import mpm_options

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nclr_mpm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nm.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_default_options_without_gpu():
    lib = nm.lib()
    assert b"gfx950" in lib.mpm_version()
    o = mpm_options()
    lib.mpm_default_options(ctypes.byref(o))
    assert o.identity_mode == 0 and o.device == 0
    assert ctypes.sizeof(mpm_options) == 16 * 4
    from nuclearmpm_amd._lib import mpm_scene
    assert ctypes.sizeof(mpm_scene) == 32


def test_null_handle_errors_are_reported_without_gpu():
    lib = nm.lib()
    assert lib.mpm_advance(None, 1) == 1
    assert b"NULL" in lib.mpm_last_error()
    assert lib.mpm_num_particles(None) == -1
    assert lib.mpm_destroy(None) == 0
    assert lib.mpm_state_bytes(None) == -1
    for name in ("mpm_save_state", "mpm_load_state", "mpm_get_state_device", "mpm_set_state_device"):
        assert getattr(lib, name)(None, None) == 1
    assert lib.mpm_backward_step(None, None, None, None) == 1
    assert lib.mpm_backward_segment(None, 1, None, None, None) == 1
    assert lib.mpm_set_scene_elasticity(None, 0, 1e4, 0.2) == 1


def test_state_channels_and_packing():
    import torch

    from nuclearmpm_amd import autodiff

    lib = nm.lib()
    assert [lib.mpm_grad_channels(d) for d in (1, 2, 3, 4)] == [-1, 13, 25, -1]
    assert autodiff.channels(2) == 13 and autodiff.channels(3) == 25
    n, d = 5, 3
    g = torch.Generator().manual_seed(0)
    x, v = torch.randn(n, d, generator=g), torch.randn(n, d, generator=g)
    F, C = torch.randn(n, d, d, generator=g), torch.randn(n, d, d, generator=g)
    Jp = torch.rand(n, generator=g)
    s = autodiff.pack(x, v, F, C, Jp)
    assert s.shape == (25, n) and s.is_contiguous()
    # channel layout: x, v, F row-major, C row-major, Jp
    assert torch.equal(s[0:3].t(), x) and torch.equal(s[3:6].t(), v)
    assert torch.equal(s[6 + 1, 2], F[2, 0, 1]) and torch.equal(s[15 + 5, 4], C[4, 1, 2])
    assert torch.equal(s[24], Jp)
    for a, b in zip(autodiff.unpack(s, d), (x, v, F, C, Jp)):
        assert torch.equal(a, b)
    assert torch.equal(autodiff.pack(x, v, F, C)[24], torch.ones(n))  # the Particle ctor's Jp


def test_create_rejects_bad_arguments_before_touching_the_device():
    lib = nm.lib()
    h = ctypes.c_void_p()
    x = np.zeros((1, 2), np.float32)
    args = lambda dim, mat, res, dt: (dim, mat, res, ctypes.c_float(dt), ctypes.c_float(1e4), ctypes.c_float(0.2),
                                      ctypes.c_float(-100), 1, x.ctypes.data, None, None, None, None, None, None,
                                      None, ctypes.byref(h))
    assert lib.mpm_create(*args(4, 0, 64, 1e-4)) == 1
    assert lib.mpm_create(*args(2, 7, 64, 1e-4)) == 1
    assert lib.mpm_create(*args(2, 0, 1, 1e-4)) == 1
    assert lib.mpm_create(*args(2, 0, 64, -1.0)) == 1


def test_particles_defaults_follow_reference_ctor():
    p = Particles.from_positions(np.zeros((3, 3), np.float32))
    assert (p.F[:, 2, 2] == 0).all() and (p.F[:, 0, 0] == 1).all()  # diag<3>(1) = diag(1,1,0)
    p1 = Particles.from_positions(np.zeros((3, 3), np.float32), identity_mode=1)
    assert (p1.F[:, 2, 2] == 1).all()
    q = Particles.from_list([Particle(np.array([0.1, 0.2]), 5), Particle(np.array([0.3, 0.4]), 6, v=np.ones(2))])
    assert q.n == 2 and list(q.c) == [5, 6] and (q.v[1] == 1).all() and (q.Jp == 1).all()
    assert MaterialModel.kSnow == 0 and MaterialModel.kJelly == 1 and MaterialModel.kLiquid == 2


def test_cube_lattice_matches_reference_layout():
    c = nm.cube(2, 50, 0.4, 0.6)
    assert c.shape == (2500, 2) and c.dtype == np.float32
    assert c[0, 0] == np.float32(0.4) and c[-1, 1] == np.float32(0.6)
    assert (c[:50, 0] == c[0, 0]).all(), "last axis fastest (src/nclr_math.h:107-109)"
    c3 = nm.cube(3, 4, 0.1, 0.3)
    assert c3.shape == (64, 3) and (c3[:4, :2] == c3[0, :2]).all()
    import oracle as O
    assert np.array_equal(nm.linspaced(50, 0.4, 0.6), O.linspaced(50, np.float32(0.4), np.float32(0.6)))
    assert np.array_equal(nm.linspaced(9, 0.9, 0.1), O.linspaced(9, np.float32(0.9), np.float32(0.1)))


def test_scene_generators_are_deterministic():
    a = scenes.four_cubes_3d()
    b = scenes.four_cubes_3d()
    assert a.particles.n == 1048576 and np.array_equal(a.particles.x, b.particles.x)
    assert a.particles.x.min() > 0.05 and a.particles.x.max() < 0.95
    c = a.extra["centres"]
    for i in range(4):
        for j in range(i):
            assert np.abs(c[i] - c[j]).max() >= 0.2
    two = scenes.two_cubes_2d()
    assert two.particles.n == 5000 and two.res == 64
    assert np.isclose(two.particles.x[2500:, 1].min(), 0.05, atol=1e-6)
    bat = scenes.batched_2d()
    assert len(bat) == 256
    assert all(1 <= len(s.particles.x) // 400 + 1 for s in bat)
    assert {s.material for s in bat} == set(MaterialModel)
    d4 = scenes.jelly_2d_diff()
    assert d4.particles.n == 19600 and d4.E == 5e3
    d5 = scenes.jelly_3d_batched()
    assert len(d5) == 8 and all(s.particles.n == 64 ** 3 for s in d5)
    assert all(5e2 <= s.E <= 5e4 for s in d5)


def test_xorshift128_reference_sequence():
    # Marsaglia's xorshift128 from the reference's fixed seed words (src/nclr_math.h:22-29)
    r = nm.XorShift128(0)
    r.x, r.y, r.z, r.w = 123456789, 362436069, 521288629, 88675123
    assert [r.next_u32() for _ in range(3)] == [3701687786, 458299110, 2500872618]
