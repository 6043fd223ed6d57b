"""Device scene generators and viewer hooks (nuclearmpm_amd/csrc/mpm_scene.hip) against the
oracle's restatements of src/nclr_math.h:22-39, 100-195, and engines started from device images
against engines started from host particles."""
import numpy as np
import pytest

import scene_oracle as SO

from nuclearmpm_amd import MaterialModel, MPMBatch, MPMSimulation, generators, scenes, viz

pytestmark = pytest.mark.gpu


def host_image(p, dim):
    n = p.x.shape[0]
    return np.concatenate([p.x.T, p.v.T, p.F.reshape(n, -1).T, p.C.reshape(n, -1).T, p.Jp[None]]).astype(np.float32)


def test_cubes_rebuild_the_benchmark_scenes_bit_for_bit():
    print("[gen] batched_2d, four_cubes_3d, jelly_3d_batched", flush=True)
    b2 = scenes.batched_2d()
    img = generators.scene_image(b2).cpu().numpy()
    from nuclearmpm_amd.sim import Particles

    allp = Particles.concatenate([s.particles for s in b2])
    assert np.array_equal(img, host_image(allp, 2))
    sc = scenes.four_cubes_3d()
    img3 = generators.scene_image(sc).cpu().numpy()
    assert np.array_equal(img3, host_image(sc.particles, 3))
    jb = scenes.jelly_3d_batched(scenes=2)
    img4 = generators.scene_image(jb, identity_mode=1).cpu().numpy()
    ref = np.concatenate([host_image(s.particles, 3) for s in jb], axis=1)
    F = ref[6:15].reshape(3, 3, -1)
    F[2, 2] = 1.0  # identity mode 1: F = I
    assert np.array_equal(img4, ref)


def test_cube_matches_oracle_lattice():
    for dim, res, lo, hi in [(2, 25, 0.3, 0.5), (3, 17, 0.7, 0.05), (2, 1, 0.2, 0.4)]:
        img = generators.cube(dim, res, lo, hi).cpu().numpy()
        want = SO.cube(dim, [res] * dim, [lo] * dim, [hi] * dim)
        assert np.array_equal(img[:dim].T, want)


def test_random_stream_matches_the_reference_sequence():
    r = generators.NcRand()
    a = r.uniform(100003).cpu().numpy()
    b = r.uniform(777).cpu().numpy()
    o = SO.NcRand()
    assert np.array_equal(a, o.uniform(100003))
    assert np.array_equal(b, o.uniform(777))
    assert r.state == o.state


def test_sample_matches_oracle():
    centers = np.array([[0.5, 0.5, 0.5], [0.3, 0.6, 0.4], [0.7, 0.2, 0.8]], np.float32)
    r = generators.NcRand()
    img = generators.sample(3, centers, r).cpu().numpy()
    o = SO.NcRand()
    want = SO.sample(3, centers, o)
    assert img.shape == (25, 3000)
    assert np.array_equal(img[:3].T, want)
    assert not img[3:6].any() and not img[15:24].any() and (img[24] == 1).all()
    assert r.state == o.state
    img2 = generators.sample(2, [[0.5, 0.5]], r, per_center=10, half_width=0.05).cpu().numpy()
    assert np.array_equal(img2[:2].T, SO.sample(2, np.float32([[0.5, 0.5]]), o, per_center=10, half=0.05))


def test_gyroid_field_and_solid():
    import torch

    rng = np.random.default_rng(3)
    p = rng.random((20000, 3)).astype(np.float32)
    k, t = 0.37, 0.1
    g = generators.gyroid(torch.as_tensor(p, device="cuda"), k, t).cpu().numpy()
    want = SO.gyroid(k, t, p)
    ulp = np.abs(g.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    print(f"[gen] gyroid: {np.count_nonzero(ulp)} of {len(p)} differ, max {ulp.max()} ulp", flush=True)
    assert ulp.max() <= 1 and np.count_nonzero(ulp) <= len(p) // 1000
    res, lo, hi = [24, 20, 16], [0.2, 0.25, 0.3], [0.8, 0.75, 0.7]
    pts = SO.cube(3, res, lo, hi)
    gv = SO.gyroid(k, t, pts)
    for thickness in (0.0, 0.3):
        img = generators.gyroid_solid(res, lo, hi, k, t, thickness).cpu().numpy()
        keep = np.abs(gv) <= np.float32(thickness) if thickness > 0 else gv <= 0
        got = img[:3].T
        if len(got) == keep.sum():
            assert np.array_equal(got, pts[keep])
        else:  # only points on the boundary may differ (1-ulp gyroid values)
            assert abs(len(got) - keep.sum()) <= 2 and np.abs(gv[keep]).min() < 1e-6
        F = img[6:15].T.reshape(-1, 3, 3)
        assert np.array_equal(F, np.broadcast_to(np.diag([1, 1, 0]).astype(np.float32), F.shape))


def test_projection_matches_oracle():
    rng = np.random.default_rng(4)
    p = rng.random((5000, 3)).astype(np.float32)
    for centered in (False, True):
        assert np.array_equal(viz.pt_3d_to_2d(p, centered=centered), SO.pt_3d_to_2d(p, centered=centered))


def test_engine_from_device_image_equals_host_construction():
    sc = scenes.four_cubes_3d(side_res=24)
    img = generators.scene_image(sc)
    with MPMSimulation.from_image(img, MaterialModel.kSnow, res=128) as a, \
            MPMSimulation(sc.particles, MaterialModel.kSnow, res=128) as b:
        a.advance(3)
        b.advance(3)
        pa, pb = a.particles(), b.particles()
        for k in ("x", "v", "F", "C", "Jp"):
            assert np.array_equal(getattr(pa, k), getattr(pb, k)), k
        ga, gb = viz.export_grid(a), viz.export_grid(b)
        assert np.array_equal(ga["mass"], gb["mass"]) and len(ga["mass"]) > 0
        assert ga["mass"].sum() == pytest.approx(sc.particles.n, rel=1e-5)
        assert np.array_equal(a.state_image().cpu().numpy(), b.state_image().cpu().numpy())


def test_batch_from_device_image_equals_host_batch():
    b2 = scenes.batched_2d(scenes=12)
    img = generators.scene_image(b2)
    with MPMBatch.from_image(img, b2) as a, MPMBatch(b2) as b:
        a.advance(2)
        b.advance(2)
        for k in ("x", "v", "F", "C", "Jp"):
            assert np.array_equal(getattr(a.particles(), k), getattr(b.particles(), k)), k
