"""Differentiable rollouts on the HIP engine (jelly, snow, liquid), as a torch.autograd.Function.

The forward pass advances the engine `steps` steps from a given state, saving a device snapshot
of the state every `checkpoint_every` steps.  The backward pass walks the segments in reverse: it
loads a snapshot and calls the engine's segment adjoint (mpm_backward_segment), which recomputes
the states inside the segment and runs the hand-written adjoint kernels (G2P with the snow
clamp's singular-value backward, grid update, P2G with the polar-decomposition backward and the
snow hardening) step by step in reverse.  Gradients flow to the initial x, v, F, C, Jp of every
particle and to Young's modulus of every scene.

With `record`, segments are recorded in the forward pass instead (mpm_advance_recorded: the
states, sorts and optionally grids the recomputation would produce, on a device tape per segment)
and their backward skips the recomputation (mpm_backward_recorded); the gradients are bitwise the
same, the device memory grows by one tape per recorded segment.

State tensors are (channels, n) float32 on the engine's device: rows x[d], v[d], F[d*d], C[d*d]
(matrices row-major), Jp, particles in the order the simulation was created with.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from ._lib import check, lib
from .sim import _Engine


def channels(dim: int) -> int:
    """Rows of a state tensor: x, v, F, C, Jp."""
    return 2 * dim + 2 * dim * dim + 1


def pack(x: torch.Tensor, v: torch.Tensor, F: torch.Tensor, C: torch.Tensor,
         Jp: Optional[torch.Tensor] = None) -> torch.Tensor:
    """(n,d), (n,d), (n,d,d), (n,d,d)[, (n,)] -> (channels, n) contiguous float32 (Jp = 1 if omitted,
    the Particle ctor's value)."""
    n = x.shape[0]
    if Jp is None:
        Jp = torch.ones(n, dtype=x.dtype, device=x.device)
    rows = torch.cat([x.reshape(n, -1), v.reshape(n, -1), F.reshape(n, -1), C.reshape(n, -1), Jp.reshape(n, 1)],
                     dim=1)
    return rows.t().contiguous().float()


def unpack(state: torch.Tensor, dim: int) -> Tuple[torch.Tensor, ...]:
    """(channels, n) -> views x (n,d), v (n,d), F (n,d,d), C (n,d,d), Jp (n,)."""
    d = dim
    t = state.t()
    n = t.shape[0]
    c0 = 2 * d + d * d
    return (t[:, :d], t[:, d:2 * d], t[:, 2 * d:c0].reshape(n, d, d), t[:, c0:c0 + d * d].reshape(n, d, d),
            t[:, c0 + d * d])


def _ready(t: torch.Tensor) -> None:
    # the engine runs on its own stream: inputs produced on torch's stream must be complete
    torch.cuda.current_stream(t.device).synchronize()


RECORD_MODES = ("none", "auto", "grid", "nogrid")


def segment_sizes(steps: int, every: int) -> list:
    """Steps of each checkpoint segment: `every` each, the last one the remainder."""
    return [min(every, steps - d) for d in range(0, steps, every)]


def plan_records(sizes, record: str, record_bytes, budget: float = 0.0) -> list:
    """Per segment: None (checkpoint only, recomputed in the backward pass), 0 (recorded without
    grids) or 1 (recorded with grids).  record_bytes(k, with_grid) is a segment's tape size.
    "auto" records as many segments as fit in `budget` bytes, first without grids, then adds
    grids while they fit."""
    if record == "none":
        return [None] * len(sizes)
    if record in ("grid", "nogrid"):
        return [1 if record == "grid" else 0] * len(sizes)
    plan, used = [None] * len(sizes), 0
    for i, k in enumerate(sizes):
        b = record_bytes(k, 0)
        if used + b <= budget:
            plan[i], used = 0, used + b
    for i, k in enumerate(sizes):
        if plan[i] == 0:
            extra = record_bytes(k, 1) - record_bytes(k, 0)
            if used + extra <= budget:
                plan[i], used = 1, used + extra
    return plan


def _record_plan(L, h, steps: int, every: int, record: str, device) -> list:
    """plan_records for an engine; "auto" takes half the device memory that is free or cached by
    torch as its budget."""
    budget = 0.0
    if record == "auto":
        free, _ = torch.cuda.mem_get_info(device)
        budget = 0.5 * (free + torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device))
    return plan_records(segment_sizes(steps, every), record, lambda k, g: int(L.mpm_record_bytes(h, k, g)), budget)


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0: torch.Tensor, E: torch.Tensor, sim: _Engine, steps: int, every: int, record: str):
        L, h = lib(), sim._handle()
        if state0.shape != (channels(sim.dim), sim.n) or state0.dtype != torch.float32 or not state0.is_cuda:
            raise ValueError("state0 must be a (channels, n) float32 tensor on the engine's device")
        if E.numel() != sim.scenes:
            raise ValueError("E needs one value per scene")
        if steps < 1 or every < 1:
            raise ValueError("steps and checkpoint_every must be >= 1")
        state0 = state0.contiguous()
        _ready(state0)
        E_vals = E.detach().double().cpu().tolist()
        for s, e in enumerate(E_vals):
            sim.set_elasticity(s, e)
        check(L.mpm_set_state_device(h, state0.data_ptr()))
        sb = int(L.mpm_state_bytes(h))
        plan = _record_plan(L, h, steps, every, record, state0.device)
        ckpts = []
        done = 0
        for mode in plan:
            k = min(every, steps - done)
            buf = torch.empty(sb, dtype=torch.uint8, device=state0.device)
            check(L.mpm_save_state(h, buf.data_ptr()))
            rec = None
            if mode is None:
                check(L.mpm_advance(h, k))
            else:
                rec = torch.empty(int(L.mpm_record_bytes(h, k, mode)), dtype=torch.uint8, device=state0.device)
                check(L.mpm_advance_recorded(h, k, rec.data_ptr(), mode))
            ckpts.append((buf, k, rec, mode))
            done += k
        final = torch.empty(sb, dtype=torch.uint8, device=state0.device)
        check(L.mpm_save_state(h, final.data_ptr()))
        out = torch.empty_like(state0)
        check(L.mpm_get_state_device(h, out.data_ptr()))
        check(L.mpm_synchronize(h))
        sim._cache = None
        ctx.sim = sim
        ctx.ckpts = ckpts
        ctx.final = final
        ctx.E_meta = (E.dtype, E.device, E.shape)
        ctx.E_vals = E_vals
        ctx.nu_vals = list(sim._nus)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        sim = ctx.sim
        L, h = lib(), sim._handle()
        g_user = grad_out.contiguous().float()
        gE = torch.zeros(sim.scenes, dtype=torch.float64, device=g_user.device)
        _ready(g_user)  # also completes gE's fill: the engine accumulates into it on its own stream
        # the segment adjoint recomputes the forward with the engine's current elasticity: restore
        # this rollout's E (another rollout or set_elasticity may have changed it since)
        for s, (e, nu) in enumerate(zip(ctx.E_vals, ctx.nu_vals)):
            sim.set_elasticity(s, e, nu)
        g = torch.empty_like(g_user)
        check(L.mpm_load_state(h, ctx.final.data_ptr()))
        check(L.mpm_grad_from_particle_order(h, g_user.data_ptr(), g.data_ptr()))
        for buf, k, rec, mode in reversed(ctx.ckpts):
            check(L.mpm_load_state(h, buf.data_ptr()))
            if rec is None:
                check(L.mpm_backward_segment(h, int(k), g.data_ptr(), g.data_ptr(), gE.data_ptr()))
            else:
                check(L.mpm_backward_recorded(h, int(k), rec.data_ptr(), mode, g.data_ptr(), g.data_ptr(),
                                              gE.data_ptr()))
        g0 = torch.empty_like(g)
        check(L.mpm_grad_to_particle_order(h, g.data_ptr(), g0.data_ptr()))
        check(L.mpm_synchronize(h))
        sim._cache = None
        dtype, device, shape = ctx.E_meta
        return g0, gE.to(dtype=dtype, device=device).reshape(shape), None, None, None, None


def rollout(sim: _Engine, state0: torch.Tensor, E: torch.Tensor, steps: int, checkpoint_every: int = 32,
            record: str = "none"):
    """Differentiable `steps`-step rollout of a simulation or batch (any materials) from `state0`
    ((channels, n) float32 on the engine's device) with Young's modulus E (one per scene).
    Returns the final state in the same layout.  The simulation is left at the rollout's final
    state after the forward pass and at its initial state after the backward pass.

    record: "none" saves a checkpoint every `checkpoint_every` steps and recomputes each segment in
    the backward pass; "grid" / "nogrid" record every segment in the forward pass (with / without
    its grids) so that the backward pass does not recompute it; "auto" records as many segments as
    fit in half the free device memory.  Nothing is recorded when no gradient is needed."""
    if record not in RECORD_MODES:
        raise ValueError(f"record must be one of {RECORD_MODES}")
    if not (torch.is_grad_enabled() and (state0.requires_grad or E.requires_grad)):
        record = "none"
    return _Rollout.apply(state0, E, sim, int(steps), int(checkpoint_every), record)
