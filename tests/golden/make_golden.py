"""Writes the golden step fixtures in this folder from the oracle (oracle/).

Each fixture is a small seeded scene advanced a few steps by the FP32 oracle to a
mid-trajectory state (so that C, F and Jp are non-trivial), followed by one more step in FP32 and
in FP64 from that state.  Every fixture also carries a random cotangent of the state after the
step (x, v, F, C, Jp) and the FP64 vector-Jacobian product of the step.

    python tests/golden/make_golden.py        # rewrites tests/golden/*.npz

The reference has no tests or fixtures of its own; these are the oracle's, committed so that a
change of the oracle or of the engine shows up against fixed vectors.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

FIELDS = ("x", "v", "F", "C", "Jp", "mass", "volume")
MATERIALS = {"snow": O.SNOW, "jelly": O.JELLY, "liquid": O.LIQUID}


def lattice(d, side, lo, hi):
    g = np.linspace(lo, hi, side, dtype=np.float32)
    return np.stack(np.meshgrid(*([g] * d), indexing="ij"), -1).reshape(-1, d)


def scene(d, material, identity_mode, seed):
    """Two cubes with seeded per-cube velocities (2D: 2 x 10^2, 3D: 2 x 5^3 particles)."""
    rng = np.random.default_rng(seed)
    side = 10 if d == 2 else 5
    a = lattice(d, side, 0.35, 0.5)
    b = lattice(d, side, 0.5, 0.65) + np.float32(0.02)
    x = np.concatenate([a, b]).astype(np.float32)
    st = O.default_state(x, identity_mode=identity_mode)
    v = np.concatenate([np.repeat(rng.uniform(-1, 1, (1, d)), len(a), 0),
                        np.repeat(rng.uniform(-1, 1, (1, d)), len(b), 0)]).astype(np.float32)
    st.v = v
    return st


CASES = [
    # name, dim, material, res, identity_mode, warm-up steps, E
    ("snow_2d", 2, "snow", 32, 0, 8, 1e4),
    ("jelly_2d", 2, "jelly", 32, 0, 8, 1e4),
    ("liquid_2d", 2, "liquid", 32, 0, 8, 1e4),
    ("snow_3d", 3, "snow", 16, 0, 6, 1e4),
    ("jelly_3d", 3, "jelly", 16, 0, 6, 1e4),
    ("liquid_3d_physical", 3, "liquid", 16, 1, 6, 1e4),
    ("snow_3d_physical", 3, "snow", 16, 1, 6, 1e4),
]


def make(name, d, material, res, idm, warm, E, seed):
    mat = MATERIALS[material]
    st = scene(d, mat, idm, seed)
    kw = dict(res=res, dt=1e-4, E=E, nu=0.2, gravity=-100.0, identity_mode=idm)
    mid = O.run(st, mat, steps=warm, **kw)
    out32 = O.run(mid, mat, steps=1, **kw)
    out64 = O.run(mid, mat, steps=1, fp64=True, **kw)
    rec = {"dim": d, "material": mat, "res": res, "dt": 1e-4, "E": E, "nu": 0.2, "gravity": -100.0,
           "identity_mode": idm, "warmup": warm}
    for k in FIELDS:
        rec[f"in_{k}"] = getattr(mid, k).astype(np.float32)
        rec[f"out32_{k}"] = getattr(out32, k).astype(np.float32)
        rec[f"out64_{k}"] = getattr(out64, k).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    n = mid.n
    g = O.Grads(x=rng.normal(size=(n, d)), v=rng.normal(size=(n, d)), F=rng.normal(size=(n, d, d)),
                C=rng.normal(size=(n, d, d)), Jp=rng.normal(size=n))
    g = O.Grads(*(getattr(g, k).astype(np.float32).astype(np.float64) for k in ("x", "v", "F", "C")),
                Jp=g.Jp.astype(np.float32).astype(np.float64))
    vjp = O.adjoint(mid, g, mat, **kw)
    for k in ("x", "v", "F", "C", "Jp"):
        rec[f"cot_{k}"] = getattr(g, k).astype(np.float32)
        rec[f"vjp_{k}"] = getattr(vjp, k)
    rec["vjp_E"] = np.float64(vjp.E)
    rec["vjp_E_abs"] = np.float64(vjp.E_abs)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **{k: np.asarray(v) for k, v in rec.items()})
    return mid.n


def main():
    O.build()
    for i, case in enumerate(CASES):
        n = make(*case, seed=10 + i)
        print(f"{case[0]}: {n} particles")


if __name__ == "__main__":
    main()
