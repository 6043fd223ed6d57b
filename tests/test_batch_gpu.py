"""Batched scenes, hipGraph steps and launch-shape tuning.

P2G sums each work item exactly in fixed point and the grid sums the slots in a fixed order, so a
step is deterministic: a scene advanced inside a batch must equal the same scene advanced alone
bit for bit, a graph replay must equal the eager launch sequence bit for bit, and every launch
shape must meet the one-step parity bound against the oracle.
"""
import functools

import numpy as np
import pytest

import oracle as O
from parity import assert_parity, assert_parity_with_floor, reorder_noise

from nuclearmpm_amd import MaterialModel, MPMBatch, MPMSimulation, Particles, SceneSpec, scenes

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "C", "Jp")


def to_oracle(p):
    return O.State(x=p.x, v=p.v, F=p.F, C=p.C, Jp=p.Jp, mass=p.mass, volume=p.volume)


def assert_same(a, b, label):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), f"{label}: field {k} differs"


def mixed_2d_scenes():
    base = scenes.batched_2d(seed=5, scenes=5)
    out = [SceneSpec(s.particles, s.material, s.res) for s in base]
    # different resolution, material, E, nu, gravity and an empty scene in the same batch
    out.append(SceneSpec(scenes.two_cubes_2d().particles, MaterialModel.kSnow, 64, E=5e3, nu=0.3, gravity=-50.0))
    out.append(SceneSpec(Particles.from_positions(np.zeros((0, 2), np.float32)), MaterialModel.kJelly, 32))
    out.append(SceneSpec(Particles.from_positions(scenes.cube(2, 30, 0.2, 0.4), v=[1.0, 2.0]), MaterialModel.kLiquid,
                         100, dt=5e-5))
    return out


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_batch_equals_sequential_2d(graph):
    specs = mixed_2d_scenes()
    assert len({int(s.material) for s in specs}) == 3
    steps = 30
    print(f"[batch] 2D {len(specs)} scenes, {steps} steps, graph={graph}", flush=True)
    with MPMBatch(specs, graph=graph) as b:
        b.advance(steps)
        outs = [b.particles(s) for s in range(len(specs))]
        grids = [b.grid(s) for s in range(len(specs))]
    for s, sp in enumerate(specs):
        with MPMSimulation(sp.particles, sp.material, res=sp.res, dt=sp.dt, E=sp.E, nu=sp.nu,
                           gravity=sp.gravity) as sim:
            sim.advance(steps)
            assert_same(outs[s], sim.particles(), f"scene {s}")
            gv, gm = sim.grid()
            assert np.array_equal(grids[s][0], gv) and np.array_equal(grids[s][1], gm), f"scene {s} grid"


def test_batch_equals_sequential_3d_and_parity():
    specs = [SceneSpec(s.particles, s.material, 32, E=s.E) for s in scenes.jelly_3d_batched(seed=3, scenes=3,
                                                                                             side_res=12)]
    specs.append(SceneSpec(scenes.four_cubes_3d(side_res=10).particles, MaterialModel.kSnow, 48))
    print(f"[batch] 3D {len(specs)} scenes", flush=True)
    with MPMBatch(specs) as b:
        b.advance(12)
        mid = b.particles().copy()
        b.advance(1)
        after = [b.particles(s) for s in range(len(specs))]
    for s, sp in enumerate(specs):
        with MPMSimulation(sp.particles, sp.material, res=sp.res, E=sp.E) as sim:
            sim.advance(13)
            assert_same(after[s], sim.particles(), f"3D scene {s}")
        a, bnd = np.cumsum([0] + [x.particles.n for x in specs])[s:s + 2]
        ref = O.run(to_oracle(mid.slice(a, bnd)), int(sp.material), res=sp.res, E=sp.E, steps=1)
        assert_parity(after[s], ref, inv_dx=float(sp.res), label=f"batched 3D scene {s}")


def test_graph_replays_equal_eager_with_interleaved_work():
    sc = scenes.two_cubes_2d()
    with MPMSimulation(sc.particles, sc.material, res=sc.res) as eager, \
            MPMSimulation(sc.particles, sc.material, res=sc.res, graph=True) as g:
        for rnd in range(4):
            eager.advance(7)
            g.advance(7)
            e_state = eager.particles()
            g_state = g.particles()  # synchronises between replays
            assert_same(g_state, e_state, f"replay round {rnd}")
            np.testing.assert_array_equal(g.grid()[1], eager.grid()[1])
        g.set_particles(e_state)  # eager host work between replays
        eager.set_particles(e_state)
        g.advance(3)
        eager.advance(3)
        assert_same(g.particles(), eager.particles(), "after set_particles")


@functools.lru_cache(maxsize=None)
def _launch_shape_case(identity_mode):
    """Four 24^3 snow cubes after 15 steps at res 64, the oracle's step from there and the oracle's
    own reorder noise on that state.  In the reference identity mode (F0 = diag(1, 1, 0)) the snow
    sits on its clamp and the step is ill-conditioned in F; the physical identity mode is not."""
    sc = scenes.four_cubes_3d(side_res=24)
    p = sc.particles if identity_mode == 0 else Particles.from_positions(sc.particles.x, v=sc.particles.v,
                                                                          identity_mode=1)
    with MPMSimulation(p, sc.material, res=64, identity_mode=identity_mode) as sim:
        sim.advance(15)
        mid = sim.particles().copy()
    ref, floor = reorder_noise(to_oracle(mid), lambda st: O.run(st, int(sc.material), res=64, steps=1,
                                                                identity_mode=identity_mode), 64.0)
    return sc, mid, ref, floor


@pytest.mark.parametrize("identity_mode", [0, 1], ids=["reference_identity", "physical_identity"])
@pytest.mark.parametrize("chunk", [256, 512, 1024])
@pytest.mark.parametrize("threads", [64, 256])
def test_every_launch_shape_meets_parity(chunk, threads, identity_mode):
    """Every launch shape takes the same step: strictly within 1e-5 of the oracle in the physical
    identity mode; in the reference identity mode, whose F the oracle itself reproduces only to its
    reorder noise (the same step with the particles in another order), within max(1e-5, 2 x that
    noise) per field."""
    sc, mid, ref, floor = _launch_shape_case(identity_mode)
    with MPMSimulation(mid, sc.material, res=64, p2g_chunk=chunk, g2p_threads=threads,
                       identity_mode=identity_mode) as sim:
        assert sim.tuning() == {"p2g_chunk": chunk, "g2p_threads": threads, "compact": 0}
        sim.advance(1)
        out = sim.particles()
    label = f"chunk {chunk} threads {threads} idm{identity_mode}"
    if identity_mode == 1:
        errs = assert_parity(out, ref, inv_dx=64.0, label=label)
    else:
        errs = assert_parity_with_floor(out, ref, floor, inv_dx=64.0, label=label)
    print(f"[tuning] {label} {errs} floor {floor}", flush=True)


def test_autotune_keeps_state_and_picks_a_valid_shape():
    sc = scenes.two_cubes_2d()
    with MPMSimulation(sc.particles, sc.material, res=sc.res) as sim:
        sim.advance(5)
        before = sim.particles().copy()
        choice = sim.autotune(steps_per_trial=2)
        assert choice["p2g_chunk"] in (256, 512, 1024) and choice["g2p_threads"] in (64, 128, 256)
        assert sim.steps_taken == 5
        assert_same(sim.particles(), before, "autotune restores the state")
        sim.advance(1)
        assert np.isfinite(sim.particles().x).all()
    with MPMSimulation(sc.particles, sc.material, res=sc.res, autotune=True) as sim:
        assert sim.steps_taken == 0
        assert_same(sim.particles(), sc.particles, "autotune at creation")


def test_batch_errors_and_scene_params():
    specs = [SceneSpec(scenes.two_cubes_2d().particles, MaterialModel.kSnow, 64),
             SceneSpec(Particles.from_positions(np.array([[0.5, 0.5], [1.5, 0.5]], np.float32)), MaterialModel.kJelly,
                       32, E=2e3)]
    with MPMBatch(specs) as b:
        p = b.params(1)
        assert np.isclose(p["mu_0"], np.float32(2e3) / (np.float32(2) * np.float32(1.2)))
        assert p["dx"] == np.float32(1 / 32)
        b.advance(1)
        with pytest.raises(IndexError, match="scene 1 "):
            b.particles()


def reflection_scene():
    """Two 2D snow particles, one with F a scaled reflection: the reference's closed-form polar
    divides by zero there, so its first step leaves the particle at a NaN position and the second
    step's grid index of that position throws (src/nclr.h:115, 163-164)."""
    c, s = np.float32(0.6), np.float32(0.8)
    p = Particles.from_positions(np.array([[0.5, 0.5], [0.52, 0.5]], np.float32))
    p.F[0] = (np.float32(0.975) * np.array([[c, s], [s, -c]], np.float32)).astype(np.float32)
    return p


def assert_same_nan(a, b, label):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k), equal_nan=True), f"{label}: field {k} differs"


def test_failed_scenes_stop_alone_and_the_others_match_their_plain_runs():
    """Each scene of a batch fails on its own, as each MPMSimulation of the reference throws on its
    own: the step that finds a particle of scene s off its grid leaves scene s as it was before that
    step, reports s (and the step) at the next synchronising call, and every later step reports it
    again; the other scenes step on bit for bit like their plain runs.  Setting the state restarts
    every scene."""
    outside = Particles.from_positions(np.array([[0.5, 0.5], [1.2, 0.5], [0.3, 0.6]], np.float32))
    specs = [SceneSpec(scenes.two_cubes_2d().particles, MaterialModel.kSnow, 64),
             SceneSpec(reflection_scene(), MaterialModel.kSnow, 64),
             SceneSpec(scenes.batched_2d(seed=5, scenes=3)[2].particles, MaterialModel.kJelly, 128),
             SceneSpec(outside, MaterialModel.kJelly, 32)]
    steps = 30
    print(f"[batch] failure isolation: {len(specs)} scenes, {steps} steps", flush=True)
    with MPMBatch(specs) as b:
        b.advance(steps)
        with pytest.raises(IndexError) as exc:
            b.synchronize()
        msg = str(exc.value)
        print(f"[batch] reported: {msg}", flush=True)
        assert "scene 1 (step 2), 3 (step 1)" in msg, msg
        assert b.scene_status() == {1: {"flags": 1, "step": 2}, 3: {"flags": 1, "step": 1}}
        b.synchronize()  # reported once per finding
        outs = [b.particles(s).copy() for s in range(len(specs))]
        grids = [b.grid(s) for s in range(len(specs))]
        b.advance(5)
        with pytest.raises(IndexError, match=r"scene 1 \(step 2\), 3 \(step 1\)"):
            b.synchronize()
        later = [b.particles(s).copy() for s in range(len(specs))]
        b.set_particles(Particles.concatenate([sp.particles for sp in specs]))
        assert b.scene_status() == {}
        b.advance(1)
        assert b.scene_status() == {3: {"flags": 1, "step": 36}}
    for s, sp in enumerate(specs):
        with MPMSimulation(sp.particles, sp.material, res=sp.res) as sim:
            for total, got in ((steps, outs[s]), (steps + 5, later[s])):
                sim.advance(total - sim.steps_taken)
                try:
                    sim.synchronize()
                except IndexError:
                    assert s in (1, 3)
                assert_same_nan(got, sim.particles(), f"scene {s} after {total} steps")
                if total == steps:
                    gv, gm = sim.grid()
                    assert np.array_equal(grids[s][0], gv, equal_nan=True) and np.array_equal(grids[s][1], gm), s
    assert not np.isfinite(outs[1].x[0]).all()
    assert_same_nan(outs[3], specs[3].particles, "scene 3 stays at its initial state")
    assert not grids[3][1].any()  # stopped before completing a step: empty grid
