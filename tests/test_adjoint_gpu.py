"""GPU adjoint of the jelly step (hand-written HIP kernels behind mpm_backward_step /
mpm_backward_segment and the torch.autograd.Function of nuclearmpm_amd.autodiff).

Pinned against
  * the FP64 oracle adjoint on the same FP32 inputs (one step, random cotangents),
  * the FP64 oracle's chained multi-step adjoint, and central finite differences (eps = 1e-3)
    of the FP64 oracle rollout (an FP32 forward is too noisy for 1e-3 finite differences),
  * the closed-form gradient of the centre-of-mass loss (momentum-conserving transfers),
  * checkpoint-interval invariance (k = 1 and k = steps give identical gradients), and
  * batching (each scene's gradient in a batch equals the scene differentiated alone).
"""
import numpy as np
import pytest
import torch

import oracle as O

from nuclearmpm_amd import MaterialModel, MPMBatch, MPMSimulation, Particles, SceneSpec, autodiff, scenes
from nuclearmpm_amd._lib import check, lib

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "C")


def jelly_particles(d, side, seed, lo=0.4, hi=0.6, F_noise=0.05, v_scale=0.5):
    rng = np.random.default_rng(seed)
    g = np.linspace(lo, hi, side)
    x = np.stack(np.meshgrid(*([g] * d), indexing="ij"), -1).reshape(-1, d)
    x = (x + rng.uniform(-0.01, 0.01, x.shape)).astype(np.float32)
    n = x.shape[0]
    p = Particles.from_positions(x)
    p.v = rng.normal(0, v_scale, (n, d)).astype(np.float32)
    p.F = (np.eye(d)[None] + rng.normal(0, F_noise, (n, d, d))).astype(np.float32)
    p.C = rng.normal(0, 2.0, (n, d, d)).astype(np.float32)
    return p


def to_state(p):
    return autodiff.pack(*(torch.from_numpy(getattr(p, k)).cuda() for k in FIELDS))


def to_oracle(p):
    return O.State(x=p.x, v=p.v, F=p.F, C=p.C, Jp=p.Jp, mass=p.mass, volume=p.volume)


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("d,side,res,svd", [(2, 24, 64, "exact"), (3, 10, 32, "exact"), (3, 10, 32, "accurate")])
def test_one_step_vjp_matches_fp64_oracle_adjoint(d, side, res, svd):
    """The accurate SVD mode's adjoint takes its polar factor from the FP64 Newton iteration of its
    forward step; both modes are held to the FP64 oracle adjoint."""
    p = jelly_particles(d, side, seed=d)
    n = p.n
    rng = np.random.default_rng(11)
    cot = {k: rng.normal(size=getattr(p, k).shape) for k in FIELDS}
    print(f"[adjoint] one-step VJP d={d} n={n} res={res} svd={svd}", flush=True)
    with MPMSimulation(p, MaterialModel.kJelly, res=res, E=1e4, svd=svd) as sim:
        s0 = to_state(p).requires_grad_(True)
        E = torch.tensor([1e4], dtype=torch.float32)
        E.requires_grad_(True)
        out = autodiff.rollout(sim, s0, E, steps=1, checkpoint_every=1)
        go = autodiff.pack(*(torch.from_numpy(cot[k]).float().cuda() for k in FIELDS),
                           Jp=torch.zeros(n, device="cuda"))
        (out * go).sum().backward()
        g = autodiff.unpack(s0.grad.detach(), d)
        gE = float(E.grad)
    # the cotangent the GPU saw is the FP32 rounding of `cot`
    G = O.Grads(*(cot[k].astype(np.float32).astype(np.float64) for k in FIELDS))
    ref = O.adjoint_jelly(to_oracle(p), G, res=res, E=1e4)
    for k, gk in zip(FIELDS, g):
        e = rel_err(gk.cpu().numpy().reshape(getattr(ref, k).shape), getattr(ref, k))
        print(f"[adjoint] d={d} d{k}: normwise rel err vs FP64 adjoint {e:.2e}", flush=True)
        assert e < 1e-4, (k, e)
    assert abs(gE - ref.E) <= 1e-4 * abs(ref.E), (gE, ref.E)


def test_backward_step_c_abi_and_permutations():
    """mpm_backward_step on a state whose slot order is not the particle order."""
    d = 2
    p = jelly_particles(d, 20, seed=5)
    n, ng = p.n, autodiff.channels(d)
    rng = np.random.default_rng(3)
    with MPMSimulation(p, MaterialModel.kJelly, res=64) as sim:
        sim.advance(3)  # slots are now sorted by tile, not in particle order
        mid = sim.particles().copy()
        L, h = lib(), sim._handle()
        snap = torch.empty(int(L.mpm_state_bytes(h)), dtype=torch.uint8, device="cuda")
        check(L.mpm_save_state(h, snap.data_ptr()))
        sim.advance(1)
        gu = torch.from_numpy(rng.normal(size=(ng, n)).astype(np.float32)).cuda()
        gslot = torch.empty_like(gu)
        check(L.mpm_grad_from_particle_order(h, gu.data_ptr(), gslot.data_ptr()))
        check(L.mpm_load_state(h, snap.data_ptr()))
        gin = torch.empty_like(gu)
        gE = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()  # filled on torch's stream; the engine reads it on its own
        check(L.mpm_backward_step(h, gslot.data_ptr(), gin.data_ptr(), gE.data_ptr()))
        g0 = torch.empty_like(gu)
        check(L.mpm_grad_to_particle_order(h, gin.data_ptr(), g0.data_ptr()))
        check(L.mpm_synchronize(h))
        assert sim.steps_taken == 4
        # the state was left at the snapshot
        np.testing.assert_array_equal(sim.particles().x, mid.x)
    gun = gu.cpu().numpy().astype(np.float64)
    G = O.Grads(*(np.ascontiguousarray(gun[s].T).reshape(getattr(mid, k).shape) for k, s in
                  zip(FIELDS, (slice(0, 2), slice(2, 4), slice(4, 8), slice(8, 12)))), Jp=gun[12])
    ref = O.adjoint_jelly(to_oracle(mid), G, res=64)
    got = g0.cpu().numpy()
    for k, s in zip(FIELDS + ("Jp",), (slice(0, 2), slice(2, 4), slice(4, 8), slice(8, 12), slice(12, 13))):
        e = rel_err(got[s].T.reshape(getattr(ref, k).shape), getattr(ref, k))
        assert e < 1e-4, (k, e)
    assert abs(float(gE) - ref.E) <= 1e-4 * abs(ref.E)


def test_backward_step_leaves_the_following_steps_unchanged():
    """mpm_backward_step sorts the current state itself; the next advance must then sort from
    scratch, so stepping on after it equals stepping on without it, bit for bit."""
    d = 2
    p = jelly_particles(d, 20, seed=9)
    n, ng = p.n, autodiff.channels(d)
    rng = np.random.default_rng(4)
    with MPMSimulation(p, MaterialModel.kJelly, res=64) as a, MPMSimulation(p, MaterialModel.kJelly, res=64) as b:
        a.advance(5)
        b.advance(5)
        L, h = lib(), a._handle()
        gout = torch.from_numpy(rng.normal(size=(ng, n)).astype(np.float32)).cuda()
        gin = torch.empty_like(gout)
        gE = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()  # filled on torch's stream; the engine reads it on its own
        check(L.mpm_backward_step(h, gout.data_ptr(), gin.data_ptr(), gE.data_ptr()))
        check(L.mpm_synchronize(h))
        for k in range(3):
            a.advance(4)
            b.advance(4)
            pa, pb = a.particles(), b.particles()
            for f in FIELDS + ("Jp",):
                np.testing.assert_array_equal(getattr(pa, f), getattr(pb, f), err_msg=f"round {k} field {f}")


def test_two_rollouts_with_different_E_then_one_backward():
    """Two rollouts of one simulation at different Young's moduli, differentiated by one backward
    pass: each rollout's adjoint recomputes its segments with its own E, so the gradients equal
    those of the two rollouts run and differentiated separately."""
    d, steps, every = 2, 16, 8
    p = affine_jelly(d, 16, 0.35, 0.55, seed=12)
    n = p.n
    rng = np.random.default_rng(8)
    w = [autodiff.pack(*(torch.from_numpy(rng.normal(size=getattr(p, k).shape).astype(np.float32)).cuda()
                         for k in FIELDS), Jp=torch.zeros(n, device="cuda")) for _ in range(2)]
    Es = (3e3, 8e3)
    with MPMSimulation(p, MaterialModel.kJelly, res=64, E=5e3) as sim:
        s0 = to_state(p).requires_grad_(True)
        E = [torch.tensor([e], dtype=torch.float64, requires_grad=True) for e in Es]
        outs = [autodiff.rollout(sim, s0, E[i], steps, every) for i in range(2)]
        sum((outs[i] * w[i]).sum() for i in range(2)).backward()
        g_both = s0.grad.cpu().numpy().astype(np.float64)
        gE_both = [float(e.grad) for e in E]
    g_sep, gE_sep = [], []
    for i in range(2):
        with MPMSimulation(p, MaterialModel.kJelly, res=64, E=5e3) as sim:
            s1 = to_state(p).requires_grad_(True)
            e1 = torch.tensor([Es[i]], dtype=torch.float64, requires_grad=True)
            (autodiff.rollout(sim, s1, e1, steps, every) * w[i]).sum().backward()
            g_sep.append(s1.grad.cpu().numpy().astype(np.float64))
            gE_sep.append(float(e1.grad))
    assert gE_both == gE_sep, (gE_both, gE_sep)
    np.testing.assert_allclose(g_both, g_sep[0] + g_sep[1], rtol=1e-5, atol=1e-6 * np.abs(g_both).max())


def com_loss(out, d, target, n_per_scene):
    x = autodiff.unpack(out, d)[0]
    loss = 0.0
    start = 0
    for s, m in enumerate(n_per_scene):
        com = x[start:start + m].double().mean(0)
        loss = loss + ((com - target[s]) ** 2).sum()
        start += m
    return loss


def affine_jelly(d, side, lo, hi, seed):
    """A cube with an affine initial velocity field plus noise, so that it shears, rotates and
    deforms and the elastic forces (and with them dL/dE) matter."""
    x = scenes.cube(d, side, lo, hi)
    rng = np.random.default_rng(seed)
    p = Particles.from_positions(x)
    p.v = ((x - x.mean(0)) @ rng.normal(0, 2.0, (d, d)).T + rng.normal(0, 0.1, x.shape)).astype(np.float32)
    return p


@pytest.mark.parametrize("d", [2, 3])
def test_multi_step_gradient_matches_fp64_adjoint_and_finite_differences(d):
    """GPU gradients of L = <c, x_T> + <c', v_T> + <c'', F_T> after `steps` steps (checkpointed)
    against the FP64 oracle's chained adjoint on every field, and against central differences of
    the FP64 oracle rollout with eps = 1e-3 for the initial velocities and E (relative)."""
    side, res, steps, every = (20, 128, 64, 16) if d == 2 else (8, 64, 32, 8)
    E0 = 5e3
    p = affine_jelly(d, side, 0.3, 0.5, seed=d)
    n = p.n
    rng = np.random.default_rng(7)
    G = O.Grads(x=rng.normal(size=(n, d)), v=rng.normal(size=(n, d)), F=rng.normal(size=(n, d, d)),
                C=np.zeros((n, d, d)))
    print(f"[adjoint] multi-step d={d} n={n} steps={steps} k={every}", flush=True)
    with MPMSimulation(p, MaterialModel.kJelly, res=res, E=E0) as sim:
        s0 = to_state(p).requires_grad_(True)
        E = torch.tensor([E0], dtype=torch.float64, requires_grad=True)
        out = autodiff.rollout(sim, s0, E, steps, every)
        w = autodiff.pack(*(torch.from_numpy(getattr(G, k)).float().cuda() for k in FIELDS),
                          Jp=torch.zeros(n, device="cuda"))
        (out * w).sum().backward()
        g = [t.detach().cpu().numpy().astype(np.float64) for t in autodiff.unpack(s0.grad, d)]
        gE = float(E.grad)
    G32 = O.Grads(*(getattr(G, k).astype(np.float32).astype(np.float64) for k in FIELDS))
    ref = O.rollout_vjp_jelly(to_oracle(p), G32, steps, res=res, E=E0)
    for k, gk in zip(FIELDS, g):
        e = rel_err(gk.reshape(getattr(ref, k).shape), getattr(ref, k))
        print(f"[adjoint] d={d} d{k}: normwise rel err vs FP64 chained adjoint {e:.2e}", flush=True)
        assert e < 1e-3, (k, e)
    s64 = to_oracle(p).astype(np.float64)

    def L(st, E=E0):
        o = O.run(st, O.JELLY, res=res, E=E, steps=steps, fp64=True)
        return sum(float(np.sum(getattr(G32, k) * getattr(o, k))) for k in ("x", "v", "F"))

    eps = 1e-3
    for trial in range(2):
        u = rng.normal(size=(n, d))
        sp, sm = s64.copy(), s64.copy()
        sp.v = s64.v + eps * u
        sm.v = s64.v - eps * u
        fd = (L(sp) - L(sm)) / (2 * eps)
        an = float(np.sum(g[1] * u))
        print(f"[adjoint] d={d} v0 direction {trial}: fd={fd:.6e} gpu={an:.6e}", flush=True)
        assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)
    hE = eps * E0
    fdE = (L(s64, E0 + hE) - L(s64, E0 - hE)) / (2 * hE)
    print(f"[adjoint] d={d} E: fd={fdE:.6e} gpu={gE:.6e}", flush=True)
    assert abs(fdE - gE) <= 1e-3 * abs(fdE), (fdE, gE)


def test_centre_of_mass_loss_gradient():
    """The differentiable-jelly loss |com(x_T) - target|^2 on a falling cube: APIC transfers
    conserve momentum and no particle reaches the boundary band, so com(x_T) = com(x_0) +
    T dt (mean v_0 + g), every initial velocity gets the gradient 2 (com - target) T dt / n and
    E none; checked on the GPU gradient."""
    d, steps, every = 2, 128, 32
    p = Particles.from_positions(scenes.cube(d, 60, 0.3, 0.5))
    target = torch.tensor([[0.7, 0.3]], dtype=torch.float64, device="cuda")
    with MPMSimulation(p, MaterialModel.kJelly, res=128, E=5e3) as sim:
        s0 = to_state(p).requires_grad_(True)
        E = torch.tensor([5e3], dtype=torch.float64, requires_grad=True)
        out = autodiff.rollout(sim, s0, E, steps, every)
        loss = com_loss(out, d, target, [p.n])
        loss.backward()
        com = autodiff.unpack(out.detach(), d)[0].double().mean(0)
        expect = (2 * (com - target[0]) * steps * 1e-4 / p.n).cpu().numpy()
        gv = autodiff.unpack(s0.grad, d)[1].cpu().numpy()
        gx = autodiff.unpack(s0.grad, d)[0].cpu().numpy()
    assert np.abs(gv - expect).max() <= 1e-3 * np.abs(expect).max(), (gv[:3], expect)
    np.testing.assert_allclose(gx, np.broadcast_to(2 * (com - target[0]).cpu().numpy() / p.n, gx.shape),
                               rtol=1e-3, atol=1e-3 * float(np.abs(gx).max()))
    # the cube falls without deforming (F = I, no stress), so E only enters through rounding
    assert abs(float(E.grad)) * 5e3 <= 1e-3 * float(loss), (float(E.grad), float(loss))


def test_checkpoint_interval_does_not_change_gradients():
    """Every checkpoint interval, and graph replay on or off (segments recompute in place into
    the tape with graphs off and restore the engine's own buffers), gives the same gradients."""
    p = Particles.from_positions(scenes.cube(2, 30, 0.3, 0.5))
    target = torch.tensor([[0.7, 0.3]], dtype=torch.float64, device="cuda")
    grads, after, grids = [], [], []
    for every, graph in ((1, False), (7, False), (40, False), (7, True), (40, True)):
        with MPMSimulation(p, MaterialModel.kJelly, res=128, E=5e3, graph=graph) as sim:
            s0 = to_state(p).requires_grad_(True)
            E = torch.tensor([5e3], dtype=torch.float32, requires_grad=True)
            com_loss(autodiff.rollout(sim, s0, E, 40, every), 2, target, [p.n]).backward()
            grads.append((s0.grad.cpu().numpy(), float(E.grad)))
            # the backward pass leaves the engine on its own buffers at the first checkpoint, and its
            # grid export shows the last adjoint step's grid: that of a step from the checkpoint
            np.testing.assert_array_equal(sim.particles().x, p.x)
            grids.append(sim.grid())
            sim.advance(3)
            after.append(sim.particles().x.copy())
    for g, e in grads[1:]:
        np.testing.assert_array_equal(g, grads[0][0])
        assert e == grads[0][1]
    with MPMSimulation(p, MaterialModel.kJelly, res=128, E=5e3) as fresh:
        fresh.advance(1)
        gv, gm = fresh.grid()
        for v, m in grids:
            np.testing.assert_array_equal(v, gv)
            np.testing.assert_array_equal(m, gm)
        fresh.advance(2)
        for x in after:
            np.testing.assert_array_equal(x, fresh.particles().x)


@pytest.mark.parametrize("d,material,svd", [(2, "jelly", "exact"), (3, "jelly", "exact"), (3, "snow", "accurate")])
def test_recorded_segments_give_the_same_gradients(d, material, svd):
    """Segments recorded in the forward pass (with grids, without, or as many as fit) give the
    recomputing checkpoint schedule's final state and gradients bit for bit, leave the engine where
    it leaves it, and a recorded tape is refused for other nsteps or after the elasticity changed.
    Also for snow (the general G2P adjoint) in the accurate SVD mode."""
    mat = MaterialModel.kJelly if material == "jelly" else MaterialModel.kSnow
    ident = 0 if material == "jelly" else 1  # snow's SVD backward needs the physical identity
    if d == 2:
        specs = [SceneSpec(Particles.from_positions(scenes.cube(2, 30, 0.3, 0.5)), mat, 128, E=5e3)]
        steps, every = 36, 7  # the last segment is one step long
    else:
        specs = [SceneSpec(s.particles, mat, 48, E=s.E)
                 for s in scenes.jelly_3d_batched(seed=4, scenes=2, side_res=10)]
        steps, every = 20, 8
    nps = [s.particles.n for s in specs]
    targets = torch.from_numpy(np.random.default_rng(2).uniform(0.2, 0.8, (len(specs), d))).cuda()
    results = []
    for record, graph in (("none", False), ("grid", False), ("nogrid", False), ("auto", True), ("grid", True)):
        print(f"[adjoint] recorded segments d={d} {material} {svd} record={record} graph={graph}", flush=True)
        with MPMBatch(specs, graph=graph, svd=svd, identity_mode=ident) as b:
            s0 = torch.cat([to_state(s.particles) for s in specs], dim=1).requires_grad_(True)
            E = torch.tensor([s.E for s in specs], dtype=torch.float64, requires_grad=True)
            out = autodiff.rollout(b, s0, E, steps, every, record=record)
            fin = out.detach().cpu().numpy()
            com_loss(out, d, targets, nps).backward()
            x_after = np.concatenate([b.particles(s).x for s in range(len(specs))])
            b.advance(3)
            x3 = np.concatenate([b.particles(s).x for s in range(len(specs))])
            results.append((fin, s0.grad.cpu().numpy(), E.grad.numpy(), x_after, x3))
    for r in results[1:]:
        for a, ref in zip(r, results[0]):
            np.testing.assert_array_equal(a, ref)
    np.testing.assert_array_equal(results[0][3], np.concatenate([s.particles.x for s in specs]))
    # a tape is checked against its recording
    L = lib()
    with MPMBatch(specs, svd=svd, identity_mode=ident) as b:
        h = b._handle()
        k = 4
        tape = torch.empty(int(L.mpm_record_bytes(h, k, 1)), dtype=torch.uint8, device="cuda")
        snap = torch.empty(int(L.mpm_state_bytes(h)), dtype=torch.uint8, device="cuda")
        check(L.mpm_save_state(h, snap.data_ptr()))
        check(L.mpm_advance_recorded(h, k, tape.data_ptr(), 1))
        check(L.mpm_load_state(h, snap.data_ptr()))
        gch = autodiff.channels(d)
        g = torch.zeros(gch, b.n, device="cuda")
        gE = torch.zeros(len(specs), dtype=torch.float64, device="cuda")
        assert L.mpm_backward_recorded(h, k - 1, tape.data_ptr(), 1, g.data_ptr(), g.data_ptr(), gE.data_ptr()) != 0
        assert L.mpm_backward_recorded(h, k, tape.data_ptr(), 0, g.data_ptr(), g.data_ptr(), gE.data_ptr()) != 0
        b.set_elasticity(0, specs[0].E * 2)
        assert L.mpm_backward_recorded(h, k, tape.data_ptr(), 1, g.data_ptr(), g.data_ptr(), gE.data_ptr()) != 0
        b.set_elasticity(0, specs[0].E)
        check(L.mpm_backward_recorded(h, k, tape.data_ptr(), 1, g.data_ptr(), g.data_ptr(), gE.data_ptr()))


def test_batched_gradients_equal_single_scene_gradients():
    specs = [SceneSpec(s.particles, MaterialModel.kJelly, 48, E=s.E)
             for s in scenes.jelly_3d_batched(seed=4, scenes=3, side_res=10)]
    nps = [s.particles.n for s in specs]
    rng = np.random.default_rng(1)
    targets = torch.from_numpy(rng.uniform(0.2, 0.8, (len(specs), 3))).cuda()
    steps, every = 24, 8
    with MPMBatch(specs) as b:
        s0 = torch.cat([to_state(s.particles) for s in specs], dim=1).requires_grad_(True)
        E = torch.tensor([s.E for s in specs], dtype=torch.float64, requires_grad=True)
        com_loss(autodiff.rollout(b, s0, E, steps, every), 3, targets, nps).backward()
        g_batch = s0.grad.cpu().numpy()
        gE_batch = E.grad.numpy()
    start = 0
    for s, sp in enumerate(specs):
        with MPMSimulation(sp.particles, MaterialModel.kJelly, res=48, E=sp.E) as sim:
            s1 = to_state(sp.particles).requires_grad_(True)
            E1 = torch.tensor([sp.E], dtype=torch.float64, requires_grad=True)
            com_loss(autodiff.rollout(sim, s1, E1, steps, every), 3, targets[s:s + 1], [sp.particles.n]).backward()
            np.testing.assert_array_equal(g_batch[:, start:start + nps[s]], s1.grad.cpu().numpy())
            assert gE_batch[s] == float(E1.grad), (s, gE_batch[s], float(E1.grad))
        start += nps[s]


def test_large_2d_batch_gradients_equal_single_scene_gradients():
    """A 2D batch past 256 full work items takes the two-particles-per-thread P2G and G2P adjoint
    and the 256-thread grid kernel; each scene alone (a small launch) takes one particle per thread
    and the 1024-thread grid kernel.  Both give the same gradients bit for bit."""
    base = affine_jelly(2, 150, 0.3, 0.6, seed=21)  # 22,500 particles
    specs = [SceneSpec(base, MaterialModel.kJelly, 128, E=e) for e in (3e3, 4e3, 5e3, 6e3, 7e3, 8e3)]
    nps = [s.particles.n for s in specs]
    assert sum(nps) >= 256 * 512
    targets = torch.from_numpy(np.random.default_rng(5).uniform(0.2, 0.8, (len(specs), 2))).cuda()
    steps, every = 6, 3
    print(f"[adjoint] large 2D batch of {len(specs)} x {nps[0]} particles", flush=True)
    with MPMBatch(specs, p2g_chunk=512) as b:
        s0 = torch.cat([to_state(s.particles) for s in specs], dim=1).requires_grad_(True)
        E = torch.tensor([s.E for s in specs], dtype=torch.float64, requires_grad=True)
        com_loss(autodiff.rollout(b, s0, E, steps, every), 2, targets, nps).backward()
        g_batch = s0.grad.cpu().numpy()
        gE_batch = E.grad.numpy()
    start = 0
    for s, sp in enumerate(specs):
        with MPMSimulation(sp.particles, MaterialModel.kJelly, res=128, E=sp.E, p2g_chunk=512) as sim:
            s1 = to_state(sp.particles).requires_grad_(True)
            E1 = torch.tensor([sp.E], dtype=torch.float64, requires_grad=True)
            com_loss(autodiff.rollout(sim, s1, E1, steps, every), 2, targets[s:s + 1], [sp.particles.n]).backward()
            np.testing.assert_array_equal(g_batch[:, start:start + nps[s]], s1.grad.cpu().numpy())
            assert gE_batch[s] == float(E1.grad), (s, gE_batch[s], float(E1.grad))
        start += nps[s]


# ---- snow and liquid -----------------------------------------------------------------------

def plastic_particles(d, side, seed):
    """Particles with F = R diag(s) Q, s drawn on both sides of the snow clamp [0.975, 1.0045] but
    away from it, Jp in [0.9, 1.1], random v and C (physical identity mode)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.4, 0.6, side)
    x = np.stack(np.meshgrid(*([g] * d), indexing="ij"), -1).reshape(-1, d)
    x = (x + rng.uniform(-0.01, 0.01, x.shape)).astype(np.float32)
    n = x.shape[0]
    p = Particles.from_positions(x, identity_mode=1)
    p.v = rng.normal(0, 0.5, (n, d)).astype(np.float32)
    F = np.empty((n, d, d))
    for i in range(n):
        sig = rng.choice([0.96, 0.985, 0.995, 1.0, 1.002, 1.02], size=d) + rng.uniform(-1e-3, 1e-3, d)
        Rm, _ = np.linalg.qr(rng.normal(size=(d, d)))
        Qm, _ = np.linalg.qr(rng.normal(size=(d, d)))
        Rm[:, 0] *= np.sign(np.linalg.det(Rm))
        Qm[:, 0] *= np.sign(np.linalg.det(Qm))
        F[i] = Rm @ np.diag(sig) @ Qm
    p.F = F.astype(np.float32)
    p.C = rng.normal(0, 0.5, (n, d, d)).astype(np.float32)
    p.Jp = rng.uniform(0.9, 1.1, n).astype(np.float32)
    return p


MATS = {"snow": (MaterialModel.kSnow, O.SNOW), "liquid": (MaterialModel.kLiquid, O.LIQUID)}


@pytest.mark.parametrize("mat", ["snow", "liquid"])
@pytest.mark.parametrize("d,side,res,svd", [(2, 20, 64, "exact"), (3, 8, 32, "exact"), (3, 8, 32, "accurate")])
def test_one_step_vjp_snow_liquid_matches_fp64_oracle_adjoint(mat, d, side, res, svd):
    """The G2P adjoint through snow's SVD clamp and Jp update and liquid's determinant reset, and
    the P2G adjoint through snow's hardening, against the FP64 oracle adjoint on the same state."""
    em, om = MATS[mat]
    p = plastic_particles(d, side, seed=3 + d)
    n = p.n
    rng = np.random.default_rng(17)
    cot = {k: rng.normal(size=getattr(p, k).shape) for k in FIELDS + ("Jp",)}
    print(f"[adjoint] one-step VJP {mat} d={d} n={n} res={res} svd={svd}", flush=True)
    with MPMSimulation(p, em, res=res, E=1e4, identity_mode=1, svd=svd) as sim:
        s0 = autodiff.pack(*(torch.from_numpy(getattr(p, k)).cuda() for k in FIELDS),
                           Jp=torch.from_numpy(p.Jp).cuda()).requires_grad_(True)
        E = torch.tensor([1e4], dtype=torch.float64, requires_grad=True)
        out = autodiff.rollout(sim, s0, E, steps=1, checkpoint_every=1)
        go = autodiff.pack(*(torch.from_numpy(cot[k]).float().cuda() for k in FIELDS),
                           Jp=torch.from_numpy(cot["Jp"]).float().cuda())
        (out * go).sum().backward()
        g = autodiff.unpack(s0.grad.detach(), d)
        gE = float(E.grad)
    G = O.Grads(*(cot[k].astype(np.float32).astype(np.float64) for k in FIELDS),
                Jp=cot["Jp"].astype(np.float32).astype(np.float64))
    ref = O.adjoint(to_oracle(p), G, om, res=res, E=1e4, identity_mode=1)
    for k, gk in zip(FIELDS + ("Jp",), g):
        e = rel_err(gk.cpu().numpy().reshape(getattr(ref, k).shape), getattr(ref, k))
        print(f"[adjoint] {mat} d={d} d{k}: normwise rel err vs FP64 adjoint {e:.2e}", flush=True)
        assert e < 1e-4, (k, e)
    print(f"[adjoint] {mat} d={d} dE gpu={gE:.6e} fp64={ref.E:.6e}", flush=True)
    assert abs(gE - ref.E) <= 1e-4 * abs(ref.E) + 1e-6 * ref.E_abs, (gE, ref.E)


@pytest.mark.parametrize("d", [2, 3])
def test_multi_step_snow_gradient_matches_fp64_adjoint(d):
    """A snow rollout's gradients (checkpointed) against the FP64 oracle's chained adjoint: a
    deforming cube whose F leaves the elastic range, so the clamp, Jp and the hardening all act."""
    side, res, steps, every = (16, 64, 24, 8) if d == 2 else (6, 32, 12, 4)
    p = affine_jelly(d, side, 0.35, 0.55, seed=20 + d)
    p.v *= 4.0
    # the physical identity: the reference's 3D F = diag(1, 1, 0) makes every particle singular,
    # where the signed SVD's clamp is ill-conditioned (its FP32 and FP64 forwards already disagree)
    p.F = np.broadcast_to(np.eye(d, dtype=np.float32), p.F.shape).copy()
    n = p.n
    rng = np.random.default_rng(27)
    G = O.Grads(x=rng.normal(size=(n, d)), v=rng.normal(size=(n, d)), F=rng.normal(size=(n, d, d)),
                C=np.zeros((n, d, d)), Jp=rng.normal(size=n))
    print(f"[adjoint] multi-step snow d={d} n={n} steps={steps} k={every}", flush=True)
    with MPMSimulation(p, MaterialModel.kSnow, res=res, E=1e4, identity_mode=1) as sim:
        s0 = to_state(p).requires_grad_(True)
        E = torch.tensor([1e4], dtype=torch.float64, requires_grad=True)
        out = autodiff.rollout(sim, s0, E, steps, every)
        w = autodiff.pack(*(torch.from_numpy(getattr(G, k)).float().cuda() for k in FIELDS),
                          Jp=torch.from_numpy(G.Jp).float().cuda())
        (out * w).sum().backward()
        g = [t.detach().cpu().numpy().astype(np.float64) for t in autodiff.unpack(s0.grad, d)]
        gE = float(E.grad)
        jp_end = autodiff.unpack(out.detach(), d)[4].cpu().numpy()
    print(f"[adjoint] snow d={d}: Jp after the rollout in [{jp_end.min():.4f}, {jp_end.max():.4f}]", flush=True)
    assert jp_end.min() < 0.999 or jp_end.max() > 1.001, "the scene never left the elastic range"
    G32 = O.Grads(*(getattr(G, k).astype(np.float32).astype(np.float64) for k in FIELDS),
                  Jp=G.Jp.astype(np.float32).astype(np.float64))
    ref = O.rollout_vjp(to_oracle(p), G32, steps, O.SNOW, res=res, E=1e4, identity_mode=1)
    for k, gk in zip(FIELDS + ("Jp",), g):
        e = rel_err(gk.reshape(getattr(ref, k).shape), getattr(ref, k))
        print(f"[adjoint] snow d={d} d{k}: normwise rel err vs FP64 chained adjoint {e:.2e}", flush=True)
        assert e < 1e-3, (k, e)
    print(f"[adjoint] snow d={d} dE gpu={gE:.6e} fp64={ref.E:.6e}", flush=True)
    assert abs(gE - ref.E) <= 1e-3 * abs(ref.E) + 1e-6 * ref.E_abs, (gE, ref.E)


@pytest.mark.parametrize("d", [2, 3])
def test_segment_with_the_grid_tape_equals_chained_single_steps(d):
    """mpm_backward_segment keeps each recomputed step's forward grid and pre-update sums on its
    tape and runs only the adjoint kernels per step; mpm_backward_step recomputes the grid from the
    P2G slots.  Both sum the same terms in the same order: the gradients agree bit for bit."""
    p = jelly_particles(d, 20 if d == 2 else 9, seed=11)
    res = 64 if d == 2 else 32
    n, ng, k = p.n, autodiff.channels(d), 5
    rng = np.random.default_rng(6)
    with MPMSimulation(p, MaterialModel.kJelly, res=res) as sim:
        sim.advance(2)  # slots in tile order
        L, h = lib(), sim._handle()
        sb = int(L.mpm_state_bytes(h))
        snaps = []
        for t in range(k):
            buf = torch.empty(sb, dtype=torch.uint8, device="cuda")
            check(L.mpm_save_state(h, buf.data_ptr()))
            snaps.append(buf)
            sim.advance(1)
        gout = torch.from_numpy(rng.normal(size=(ng, n)).astype(np.float32)).cuda()
        # chained single steps, from the last state back to the first (torch fills its tensors on
        # its own stream: synchronise before the engine's stream reads them)
        g = gout.clone()
        gE_chain = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        # every buffer stays referenced until the engine's stream is done with it: torch's
        # allocator would otherwise hand a released one to its own stream while the engine reads it
        held = [g]
        for t in reversed(range(k)):
            check(L.mpm_load_state(h, snaps[t].data_ptr()))
            nxt = torch.empty_like(g)
            held.append(nxt)
            check(L.mpm_backward_step(h, g.data_ptr(), nxt.data_ptr(), gE_chain.data_ptr()))
            g = nxt
        check(L.mpm_synchronize(h))
        # one segment from the first state
        check(L.mpm_load_state(h, snaps[0].data_ptr()))
        gseg = gout.clone()
        gE_seg = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        check(L.mpm_backward_segment(h, k, gseg.data_ptr(), gseg.data_ptr(), gE_seg.data_ptr()))
        check(L.mpm_synchronize(h))
        print(f"[adjoint] {d}D tape bytes for {k} steps: {int(L.mpm_tape_bytes(h, k))}", flush=True)
    diff = float((gseg - g).abs().max())
    assert torch.equal(gseg, g), (diff, float(g.abs().max()), float(gE_seg), float(gE_chain))
    assert float(gE_seg) == float(gE_chain)
