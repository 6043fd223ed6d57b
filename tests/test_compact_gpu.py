"""The compact step (mpm_set_compact): every step of an advance in one cooperative launch, no
sort, the grid box of the scene's particles in LDS.

It is the reference's step (src/nclr.h:80-84) with the fixed-point P2G sums grouped by workgroup
instead of by tile, so it is held to the tiled step's bar:
  * one step from identical inputs within tests/parity.py's bound of the oracle, for every
    material, both dimensions and every particles-per-thread shape, from rest and mid-trajectory;
    the exported grid and the active-node count agree with the oracle's grid;
  * long launches (many steps, the box following the particles) end in states from which one more
    step still meets the bound, and stay close to the tiled step's trajectory;
  * a box that outgrows the LDS hands the remaining steps to the tiled step (bitwise the tiled run
    when the box never fits);
  * a particle off its grid stops the simulation as the tiled step does (same step reported, state
    and grid of the last completed step, the step count still advancing).
"""
import numpy as np
import pytest

import oracle as O
from parity import LIGHT_NODE, TOL, assert_accurate_parity, assert_parity, grid_errors

from nuclearmpm_amd import MaterialModel, MPMBatch, MPMSimulation, Particles, SceneSpec, scenes

from test_parity_gpu import MATS, gpu_state, to_oracle

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "C", "Jp")


def compact_scene_3d(identity_mode=0) -> Particles:
    """Two 3D cubes flying at each other and one falling beside them, on a res-24 grid: the box of
    their stencils (about 1000 nodes) fits one workgroup's LDS."""
    a = scenes.lattice([10] * 3, [0.30, 0.35, 0.35], [0.45, 0.50, 0.50])
    b = scenes.lattice([10] * 3, [0.47, 0.36, 0.36], [0.62, 0.51, 0.51])
    c = scenes.lattice([8] * 3, [0.50, 0.12, 0.30], [0.60, 0.22, 0.40])
    x = np.concatenate([a, b, c])
    v = np.concatenate([np.tile([3.0, 0.0, 0.5], (len(a), 1)), np.tile([-3.0, 0.5, 0.0], (len(b), 1)),
                        np.tile([0.0, -4.0, 0.0], (len(c), 1))]).astype(np.float32)
    return Particles.from_positions(x, v=v, identity_mode=identity_mode)


def compact_case(p, mat, res, ppt, label, identity_mode=0, E=1e4, steps=1):
    """`steps` compact steps from p; the last one is checked against the oracle (state and grid)."""
    print(f"[compact] {label} n={p.n} dim={p.dim} mat={int(mat)} res={res} ppt={ppt} steps={steps}", flush=True)
    with MPMSimulation(p, mat, res=res, E=E, identity_mode=identity_mode, compact=ppt) as sim:
        assert sim.tuning()["compact"] == ppt
        if steps > 1:
            sim.advance(steps - 1)
            start = sim.particles().copy()
        else:
            start = p
        sim.advance(1)
        g = sim.particles().copy()
        gv, gm = sim.grid()
        active = sim.active_nodes()
        assert sim.steps_taken == steps
    r, rv, rm = O.run(to_oracle(start), int(mat), res=res, E=E, steps=1, identity_mode=identity_mode,
                      return_grid=True)
    errs = assert_parity(g, r, inv_dx=float(res), label=label)
    gerr = grid_errors(gv, gm, rv, rm)
    print(f"[compact] {label}: {errs} grid {gerr} active={active}", flush=True)
    # velocities of nodes lighter than LIGHT_NODE are coarse in both the tiled and the compact step
    # (same fixed-point sums); mass, momentum and the other nodes' velocities meet the bound
    assert all(gerr[k] <= TOL for k in ("grid_v_massive", "grid_p", "grid_m")), (label, gerr)
    # the same nodes carry mass, except that a contribution below the fixed-point quantum (about
    # 1e-15 of the workgroup's largest) rounds to zero: such a node is lighter than LIGHT_NODE
    rmg = rm.reshape(gm.shape)
    assert (rmg[(gm > 0) != (rmg > 0)] <= LIGHT_NODE * rmg.max()).all(), label
    assert active >= int((rm > LIGHT_NODE * rm.max()).sum())  # post-P2G mass > 0, before the sticky box zeroes nodes
    return g


@pytest.mark.parametrize("ppt", [1, 2, 4])
@pytest.mark.parametrize("mat", MATS, ids=["snow", "jelly", "liquid"])
def test_compact_one_step_2d(mat, ppt):
    sc = scenes.two_cubes_2d(mat)
    compact_case(sc.particles, mat, sc.res, ppt, f"two_cubes_2d rest {mat.name}")
    mid = gpu_state(sc.particles, mat, sc.res, 300)
    compact_case(mid, mat, sc.res, ppt, f"two_cubes_2d after 300 {mat.name}")


@pytest.mark.parametrize("ppt", [1, 2])
@pytest.mark.parametrize("identity_mode", [0, 1], ids=["reference_identity", "physical_identity"])
@pytest.mark.parametrize("mat", MATS, ids=["snow", "jelly", "liquid"])
def test_compact_one_step_3d(mat, identity_mode, ppt):
    p = compact_scene_3d(identity_mode)
    compact_case(p, mat, 24, ppt, f"compact3d rest {mat.name} idm{identity_mode}", identity_mode)
    mid = gpu_state(p, mat, 24, 40, identity_mode)
    compact_case(mid, mat, 24, ppt, f"compact3d after 40 {mat.name} idm{identity_mode}", identity_mode)


def test_compact_whole_grid_3d():
    """3D at res 15: the whole grid fits the LDS, so advance does not wait for the launch."""
    x = scenes.lattice([9] * 3, [0.3, 0.35, 0.3], [0.6, 0.65, 0.6])
    p = Particles.from_positions(x, v=np.tile([1.0, -2.0, 0.5], (len(x), 1)).astype(np.float32))
    compact_case(p, MaterialModel.kJelly, 15, 1, "res15 jelly", steps=30)


def test_compact_accurate_mode_3d():
    p = gpu_state(compact_scene_3d(), MaterialModel.kSnow, 24, 40)
    with MPMSimulation(p, MaterialModel.kSnow, res=24, svd="accurate", compact=2) as sim:
        sim.advance(1)
        g = sim.particles().copy()
    st = to_oracle(p)
    r32 = O.run(st, O.SNOW, res=24, steps=1)
    r64 = O.run(st.astype(np.float64), O.SNOW, res=24, steps=1, fp64=True)
    errs = assert_accurate_parity(g, r32, r64, inv_dx=24.0, label="compact accurate")
    print(f"[compact] accurate: {errs}", flush=True)


@pytest.mark.parametrize("mat", MATS, ids=["snow", "jelly", "liquid"])
def test_compact_long_launch_2d(mat):
    """The 5000-particle 2D workload: 1000 steps in one launch, then one more step against the
    oracle, and the trajectory beside the tiled step's."""
    sc = scenes.two_cubes_2d(mat)
    g = compact_case(sc.particles, mat, sc.res, 1, f"two_cubes_2d 1000 {mat.name}", steps=1000)
    t = gpu_state(sc.particles, mat, sc.res, 1000)
    dx = np.abs(g.x - t.x).max()
    print(f"[compact] 1000 steps {mat.name}: max |x_compact - x_tiled| = {dx:.2e}", flush=True)
    assert np.isfinite(g.x).all() and dx < 0.05  # same trajectory up to the chaos of 1000 steps


def test_compact_differentiable_jelly_2d_workload():
    """The 19,600-particle 2D jelly workload on its 128^2 grid (the box path whose launch is read
    back): 512 steps, then one step against the oracle."""
    sc = scenes.jelly_2d_diff()
    compact_case(sc.particles, sc.material, sc.res, 1, "jelly_2d_diff 512", E=sc.E, steps=512)


def test_compact_box_overflow_hands_over_to_tiled():
    res = 64
    # never fits (the collision scene's box at res 32 has 21 x 17 x 15 > 4266 nodes): every step is
    # tiled (from the launch's Hilbert slot order, so equal to the plain run up to the order of the
    # P2G work items' fixed-point sums), and one more step meets the bound
    from test_parity_gpu import collision_scene_3d
    p = collision_scene_3d()
    with MPMSimulation(p, MaterialModel.kJelly, res=32, compact=1) as sim:
        sim.advance(19)
        a19 = sim.particles().copy()
        sim.advance(1)
        a = sim.particles().copy()
    b = gpu_state(p, MaterialModel.kJelly, 32, 20)
    assert np.abs(a.x - b.x).max() < 1e-5
    assert_parity(a, O.run(to_oracle(a19), O.JELLY, res=32, steps=1), inv_dx=32.0, label="never fits")
    # fits at first, outgrows the LDS within 100 steps (two cubes flying apart diagonally)
    A = scenes.lattice([5] * 3, [0.40, 0.40, 0.45], [0.46, 0.46, 0.51])
    B = scenes.lattice([5] * 3, [0.54, 0.54, 0.45], [0.60, 0.60, 0.51])
    v = np.concatenate([np.tile([-30.0, -30.0, 0.0], (len(A), 1)), np.tile([30.0, 30.0, 0.0], (len(B), 1))])
    q = Particles.from_positions(np.concatenate([A, B]), v=v.astype(np.float32))
    with MPMSimulation(q, MaterialModel.kJelly, res=res, compact=1) as sim:
        sim.advance(99)
        mid = sim.particles().copy()
        sim.advance(1)
        end = sim.particles().copy()
        assert sim.steps_taken == 100
        tiled = gpu_state(q, MaterialModel.kJelly, res, 100)
        assert not np.array_equal(end.x, tiled.x)  # compact steps ran before the box outgrew the LDS
        assert np.abs(end.x - tiled.x).max() < 1e-4
        # the handed-over engine stays tiled: its next step is bitwise the tiled step's
        sim.advance(1)
        after = sim.particles().copy()
    ref_after = gpu_state(end, MaterialModel.kJelly, res, 1)
    for k in FIELDS:
        assert np.array_equal(getattr(after, k), getattr(ref_after, k)), k
    r = O.run(to_oracle(mid), O.JELLY, res=res, steps=1)
    assert_parity(end, r, inv_dx=float(res), label="after the overflow")


def test_compact_out_of_range_stops_like_the_tiled_step():
    p = Particles.from_positions(np.array([[0.5, 0.5], [1.2, 0.5]], np.float32))
    with MPMSimulation(p, MaterialModel.kSnow, res=16, compact=1) as sim:
        sim.advance(1)
        with pytest.raises(IndexError, match="in step 1"):
            sim.particles()
        np.testing.assert_array_equal(sim.particles().x, p.x)
        sim.advance(3)
        with pytest.raises(IndexError, match="in step 1"):
            sim.synchronize()
        np.testing.assert_array_equal(sim.particles().x, p.x)
        assert sim.steps_taken == 4
        gv, gm = sim.grid()
        assert not gm.any()  # no step completed
    nan = Particles.from_positions(np.array([[np.nan, 0.5, 0.5]], np.float32))
    with MPMSimulation(nan, MaterialModel.kJelly, res=16, compact=1) as sim:  # 17^3 nodes: the read-back path
        sim.advance(2)
        with pytest.raises(IndexError, match="in step 1"):
            sim.synchronize()
        assert sim.steps_taken == 2


def test_compact_reflection_stop_keeps_the_last_completed_step():
    """2D snow with one F a scaled reflection: the first step leaves that particle at a NaN
    position (as the reference's does), the second finds it off the grid and stops."""
    c, s = np.float32(0.6), np.float32(0.8)
    p = Particles.from_positions(np.array([[0.5, 0.5], [0.52, 0.5]], np.float32))
    p.F[0] = (np.float32(0.975) * np.array([[c, s], [s, -c]], np.float32)).astype(np.float32)
    with MPMSimulation(p, MaterialModel.kSnow, res=64, compact=1) as sim, \
            MPMSimulation(p, MaterialModel.kSnow, res=64) as ref:
        for e in (sim, ref):
            e.advance(30)
            with pytest.raises(IndexError, match="in step 2"):
                e.synchronize()
            assert e.scene_status() == {0: {"flags": 1, "step": 2}}
            assert e.steps_taken == 30
        a, b = sim.particles(), ref.particles()
        assert not np.isfinite(a.x[0]).all()  # the second particle shares its nodes: NaN as well
        for k in FIELDS:  # same NaN pattern, same finite values
            np.testing.assert_allclose(getattr(a, k), getattr(b, k), rtol=1e-5, atol=1e-6, equal_nan=True)
        (av, am), (bv, bm) = sim.grid(), ref.grid()
        assert ((am > 0) == (bm > 0)).all()
        np.testing.assert_allclose(am, bm, rtol=1e-5, atol=1e-9)
        sim.set_particles(p)  # restarts: the stop is cleared
        assert sim.scene_status() == {}
        sim.advance(1)
        sim.synchronize()
        assert not np.isfinite(sim.particles().x[0]).all()


def test_compact_errors_and_autotune():
    sc = scenes.two_cubes_2d()
    with MPMSimulation(sc.particles, sc.material, res=sc.res) as sim:
        with pytest.raises(ValueError):
            sim.set_compact(3)
        sim.set_compact(4)
        sim.advance(2)
        sim.set_compact(0)
        sim.advance(1)
        assert sim.steps_taken == 3 and np.isfinite(sim.particles().x).all()
        before = sim.particles().copy()
        choice = sim.autotune(steps_per_trial=20)
        print(f"[compact] autotune on the 2D workload: {choice}", flush=True)
        assert choice["compact"] in (0, 1, 2, 4)
        assert sim.steps_taken == 3
        for k in FIELDS:
            assert np.array_equal(getattr(sim.particles(), k), getattr(before, k)), k
    with pytest.raises(ValueError):
        MPMSimulation(compact_scene_3d(), MaterialModel.kJelly, res=24, compact=4)
    big = scenes.four_cubes_3d(side_res=48)  # 442,368 particles: more than 256 workgroups of 256
    with pytest.raises(NotImplementedError):
        MPMSimulation(big.particles, big.material, res=big.res, compact=1)
    specs = [SceneSpec(sc.particles, sc.material, sc.res)] * 2
    with MPMBatch(specs) as b:
        with pytest.raises(NotImplementedError):
            b.set_compact(1)
