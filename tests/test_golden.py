"""Golden step fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
oracle): small seeded two-cube scenes for every dimension and material, at a mid-trajectory
state, with the one-step result in FP32 and FP64, a cotangent and its FP64 vector-Jacobian
product.

CPU: the oracle still reproduces every fixture bit for bit.
GPU: the engine's step from each fixture's state meets the parity bound against the FP32 result,
and the GPU adjoint against the FP64 vector-Jacobian product.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from parity import assert_parity, field_errors

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
FIELDS = ("x", "v", "F", "C", "Jp", "mass", "volume")


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def state(g, prefix):
    return O.State(*(g[f"{prefix}_{k}"] for k in FIELDS))


def kwargs(g):
    return dict(res=int(g["res"]), dt=float(g["dt"]), E=float(g["E"]), nu=float(g["nu"]),
                gravity=float(g["gravity"]), identity_mode=int(g["identity_mode"]))


def test_fixtures_present():
    names = {os.path.basename(f)[:-4] for f in FILES}
    assert {"snow_2d", "jelly_2d", "liquid_2d", "snow_3d", "jelly_3d"} <= names


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(path):
    g = load(path)
    mid = state(g, "in")
    out32 = O.run(mid, int(g["material"]), steps=1, **kwargs(g))
    out64 = O.run(mid, int(g["material"]), steps=1, fp64=True, **kwargs(g))
    for k in FIELDS:
        assert np.array_equal(getattr(out32, k), g[f"out32_{k}"]), k
        assert np.array_equal(getattr(out64, k), g[f"out64_{k}"]), k
    cot = O.Grads(*(g[f"cot_{k}"].astype(np.float64) for k in ("x", "v", "F", "C")),
                  Jp=g["cot_Jp"].astype(np.float64))
    vjp = O.adjoint(mid, cot, int(g["material"]), **kwargs(g))
    for k in ("x", "v", "F", "C", "Jp"):
        np.testing.assert_array_equal(getattr(vjp, k), g[f"vjp_{k}"])
    assert vjp.E == float(g["vjp_E"]) and vjp.E_abs == float(g["vjp_E_abs"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_engine_step_matches_fixture(path):
    from nuclearmpm_amd import MaterialModel, MPMSimulation, Particles

    g = load(path)
    kw = kwargs(g)
    p = Particles(*(g[f"in_{k}"] for k in FIELDS))
    name = os.path.basename(path)[:-4]
    print(f"[golden] {name}: n={p.n} res={kw['res']}", flush=True)
    with MPMSimulation(p, MaterialModel(int(g["material"])), res=kw["res"], dt=kw["dt"], E=kw["E"], nu=kw["nu"],
                       gravity=kw["gravity"], identity_mode=kw["identity_mode"]) as sim:
        sim.advance(1)
        out = sim.particles()
    ref32 = state(g, "out32")
    errs = assert_parity(out, ref32, inv_dx=float(kw["res"]), label=name)
    e64 = field_errors(out, state(g, "out64"), inv_dx=float(kw["res"]))
    o64 = field_errors(ref32, state(g, "out64"), inv_dx=float(kw["res"]))
    print(f"[golden] {name}: vs FP32 {errs}; GPU vs FP64 {e64}; oracle FP32 vs FP64 {o64}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_engine_vjp_matches_fixture(path):
    import torch

    from nuclearmpm_amd import MaterialModel, MPMSimulation, Particles, autodiff

    g = load(path)
    kw = kwargs(g)
    d = int(g["dim"])
    p = Particles(*(g[f"in_{k}"] for k in FIELDS))
    with MPMSimulation(p, MaterialModel(int(g["material"])), res=kw["res"], E=kw["E"],
                       identity_mode=kw["identity_mode"]) as sim:
        s0 = autodiff.pack(*(torch.from_numpy(g[f"in_{k}"]).cuda() for k in ("x", "v", "F", "C", "Jp")))
        s0.requires_grad_(True)
        E = torch.tensor([kw["E"]], dtype=torch.float64, requires_grad=True)
        out = autodiff.rollout(sim, s0, E, steps=1, checkpoint_every=1)
        cot = autodiff.pack(*(torch.from_numpy(g[f"cot_{k}"]).cuda() for k in ("x", "v", "F", "C", "Jp")))
        (out * cot).sum().backward()
        got = [t.detach().cpu().numpy().astype(np.float64) for t in autodiff.unpack(s0.grad, d)]
        gE = float(E.grad)
    # x's cotangent sums, over every stencil node, weight-derivative terms of both signs through the
    # stress: bound 2e-4 (measured at most 1.05e-4 here).  Snow's Jp cotangent goes through the
    # hardening, d(mu, lambda)/dJp, i.e. through the stress's sensitivity to mu, 2 (F - R) F^T, whose
    # FP32 rounding is eps / strain relative at small strain, the conditioning of dL/dE below:
    # bound 1e-3 (measured 5.5e-4 on the 3D snow state).  The other fields: 1e-4.
    bounds = {"x": 2e-4, "Jp": 1e-3}
    for k, gk in zip(("x", "v", "F", "C", "Jp"), got):
        ref = g[f"vjp_{k}"]
        e = float(np.abs(gk.reshape(ref.shape) - ref).max() / max(np.abs(ref).max(), 1e-30))
        print(f"[golden] {os.path.basename(path)[:-4]}: VJP d{k} normwise rel err {e:.2e}", flush=True)
        assert e < bounds.get(k, 1e-4), (k, e)
    # dL/dE sums per-particle terms that cancel (here 50x), and each term carries the FP32 rounding
    # of F - R, relative eps / strain at small strain (the forward stress's own conditioning): the
    # bound is 1e-3 of the sum of the terms' magnitudes
    assert abs(gE - float(g["vjp_E"])) <= 1e-4 * abs(float(g["vjp_E"])) + 1e-3 * float(g["vjp_E_abs"]), \
        (gE, float(g["vjp_E"]), float(g["vjp_E_abs"]))
