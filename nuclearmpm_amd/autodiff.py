"""Differentiable rollouts on the HIP engine (jelly, snow, liquid), as a torch.autograd.Function.

The forward pass advances the engine `steps` steps from a given state, saving a device snapshot
of the state every `checkpoint_every` steps.  The backward pass walks the segments in reverse: it
loads a snapshot and calls the engine's segment adjoint (mpm_backward_segment), which recomputes
the states inside the segment and runs the hand-written adjoint kernels (G2P with the snow
clamp's singular-value backward, grid update, P2G with the polar-decomposition backward and the
snow hardening) step by step in reverse.  Gradients flow to the initial x, v, F, C, Jp of every
particle and to Young's modulus of every scene.

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


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0: torch.Tensor, E: torch.Tensor, sim: _Engine, steps: int, every: int):
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
        ckpts = []
        done = 0
        while done < steps:
            k = min(every, steps - done)
            buf = torch.empty(sb, dtype=torch.uint8, device=state0.device)
            check(L.mpm_save_state(h, buf.data_ptr()))
            check(L.mpm_advance(h, k))
            ckpts.append((buf, k))
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
        for buf, k in reversed(ctx.ckpts):
            check(L.mpm_load_state(h, buf.data_ptr()))
            check(L.mpm_backward_segment(h, int(k), g.data_ptr(), g.data_ptr(), gE.data_ptr()))
        g0 = torch.empty_like(g)
        check(L.mpm_grad_to_particle_order(h, g.data_ptr(), g0.data_ptr()))
        check(L.mpm_synchronize(h))
        sim._cache = None
        dtype, device, shape = ctx.E_meta
        return g0, gE.to(dtype=dtype, device=device).reshape(shape), None, None, None


def rollout(sim: _Engine, state0: torch.Tensor, E: torch.Tensor, steps: int, checkpoint_every: int = 32):
    """Differentiable `steps`-step rollout of a simulation or batch (any materials) from `state0`
    ((channels, n) float32 on the engine's device) with Young's modulus E (one per scene).
    Returns the final state in the same layout.  The simulation is left at the rollout's final
    state after the forward pass and at its initial state after the backward pass."""
    return _Rollout.apply(state0, E, sim, int(steps), int(checkpoint_every))
