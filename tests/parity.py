"""Parity metrics shared by the GPU tests (not a test module).

One-step agreement is measured per field as a normwise relative error
    err(f) = max|f_gpu - f_ref| / max|f_ref|
for x, v, F and Jp, each bounded by TOL = 1e-5.

C is bounded by TOL on
    max|C_gpu - C_ref| / max(max|C_ref|, C_FLOOR * 4 inv_dx max|v_ref|)
with C_FLOOR = 0.1: the plain normwise error wherever C is at least a tenth of its velocity scale
4 inv_dx max|v|, as in every mid-trajectory state.  C is a weighted difference of grid velocities,
C = 4 inv_dx sum_i w_i v_i (i - fx)^T, so its FP32 rounding scales with 4 inv_dx max|v|, not with
max|C|: in a cube falling at rest (C is rounding noise, max|C| ~ 1e-7 of that scale) no two FP32
evaluations agree on C to 1e-5 of max|C|, and two runs of the reference's own step that differ
only in P2G's summation order disagree by up to 5e-5 of max|C| in free fall.

Reported beside them, not bounded:
  * `C_plain`: max|dC| / max|C_ref|, and `C_scaled`: max|dC| / max(max|C_ref|, 4 inv_dx max|v_ref|);
  * `<field>_elem`: the elementwise relative error max |f_gpu - f_ref| / |f_ref| over the elements
    with |f_ref| > 1e-3 max|f_ref| (small elements of a field carry the rounding of its large ones).
"""
from __future__ import annotations

import numpy as np

TOL = 1e-5  # one-step relative agreement required for x, v, C, F (and Jp)
C_FLOOR = 0.1  # C is normalised by max(max|C|, C_FLOOR * 4 inv_dx max|v|)
ELEM_FLOOR = 1e-3  # elementwise errors are taken over |f_ref| > ELEM_FLOOR * max|f_ref|


def elementwise_error(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    if not b.size:
        return 0.0
    big = np.abs(b) > ELEM_FLOOR * np.abs(b).max()
    if not big.any():
        return 0.0
    return float((np.abs(a[big] - b[big]) / np.abs(b[big])).max())


def field_errors(gpu, ref, inv_dx: float) -> dict:
    out = {}
    for k in ("x", "v", "F", "Jp"):
        a = np.asarray(getattr(gpu, k), np.float64)
        b = np.asarray(getattr(ref, k), np.float64)
        den = np.abs(b).max() if b.size else 0.0
        out[k] = float(np.abs(a - b).max() / den) if b.size and den > 0 else float(np.abs(a - b).max(initial=0.0))
    a = np.asarray(gpu.C, np.float64)
    b = np.asarray(ref.C, np.float64)
    vmax = np.abs(np.asarray(ref.v, np.float64)).max(initial=0.0)
    den = max(np.abs(b).max(initial=0.0), 4.0 * inv_dx * vmax)
    dC = np.abs(a - b).max(initial=0.0)
    cmax = np.abs(b).max(initial=0.0)
    cden = max(cmax, C_FLOOR * 4.0 * inv_dx * vmax)
    out["C"] = float(dC / cden) if cden > 0 else float(dC)
    out["C_plain"] = float(dC / cmax) if cmax > 0 else float(dC)
    out["C_scaled"] = float(dC / den) if den > 0 else float(dC)
    for k in ("x", "v", "F", "C", "Jp"):
        out[k + "_elem"] = elementwise_error(getattr(gpu, k), getattr(ref, k))
    return out


def bounded(errs: dict, scaled_C: bool = False) -> dict:
    """The normwise errors the parity bound applies to (C on the plain normalisation, or on the
    scaled one for two FP32 float-summing paths compared with each other)."""
    out = {k: v for k, v in errs.items() if k not in ("C", "C_plain", "C_scaled") and not k.endswith("_elem")}
    out["C"] = errs["C_scaled"] if scaled_C else errs["C"]
    return out


def assert_parity(gpu, ref, inv_dx: float, tol: float = TOL, label: str = "", scaled_C: bool = False) -> dict:
    errs = field_errors(gpu, ref, inv_dx)
    bad = {k: v for k, v in bounded(errs, scaled_C).items() if not (v <= tol)}
    assert not bad, f"{label}: parity above {tol}: {bad} (all: {errs})"
    return errs


def reorder_noise(state, run_fn, inv_dx: float, seed: int = 1) -> dict:
    """The reference step's own FP32 reproducibility on `state`: field errors between two oracle
    runs that differ only in particle order (hence in P2G's summation order).  In the exploding
    3D reference-mode snow workload the signed-SVD plasticity clamp (src/nclr_math.h:63-71 with
    src/nclr.h:241-249) is ill-conditioned for det(F^) < 0, and this floor reaches 3e-4 in F."""
    import copy

    n = state.x.shape[0]
    perm = np.random.default_rng(seed).permutation(n)
    inv = np.argsort(perm)
    sp = copy.copy(state)
    for k in ("x", "v", "F", "C", "Jp", "mass", "volume"):
        setattr(sp, k, np.ascontiguousarray(getattr(state, k)[perm]))
    a = run_fn(state)
    b = run_fn(sp)
    for k in ("x", "v", "F", "C", "Jp"):
        setattr(b, k, getattr(b, k)[inv])
    return a, field_errors(b, a, inv_dx)


def assert_parity_with_floor(gpu, ref, noise: dict, inv_dx: float, tol: float = TOL, factor: float = 2.0,
                             label: str = "") -> dict:
    """Per-field bound max(tol, factor * floor), where floor is reorder_noise()'s measurement."""
    errs = field_errors(gpu, ref, inv_dx)
    bad = {k: (v, noise[k]) for k, v in bounded(errs).items() if not (v <= max(tol, factor * noise[k]))}
    assert not bad, f"{label}: above max({tol}, {factor} x reorder floor): {bad} (all: {errs}, floor: {noise})"
    return errs


# Nodes lighter than this fraction of the heaviest node hold a few fixed-point quanta of mass and
# momentum after P2G (the step sums exactly in 64-bit fixed point scaled to the largest
# contribution), so their velocity p / m is coarse; their weight in any particle's G2P is as small.
LIGHT_NODE = 1e-9


def grid_errors(gv, gm, ref_v, ref_m) -> dict:
    """Normwise grid errors: velocity over every node (grid_v) and over nodes of mass >= LIGHT_NODE
    of the heaviest (grid_v_massive), momentum m v (grid_p) and mass (grid_m)."""
    ref_v = np.asarray(ref_v, np.float64)
    ref_m = np.asarray(ref_m, np.float64)
    gv = np.asarray(gv, np.float64).reshape(ref_v.shape)
    gm = np.asarray(gm, np.float64).reshape(ref_m.shape)

    def rel(a, b):
        return float(np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-30))

    heavy = ref_m >= LIGHT_NODE * ref_m.max(initial=0.0)
    vm = ref_v.reshape(ref_m.size, -1)
    gvm = gv.reshape(ref_m.size, -1)
    hm = heavy.reshape(-1)
    mm, gmm = ref_m.reshape(-1, 1), gm.reshape(-1, 1)
    return {"grid_v": rel(gv, ref_v), "grid_m": rel(gm, ref_m), "grid_p": rel(gmm * gvm, mm * vm),
            "grid_v_massive": rel(gvm[hm], vm[hm]) if hm.any() else 0.0}


def assert_accurate_parity(gpu, ref32, ref64, inv_dx: float, tol: float = TOL, label: str = "") -> dict:
    """One-step agreement of a step whose polar factor and SVD are computed more accurately than
    the reference's FP32 two-sided Jacobi (the engine's accurate SVD mode).  Per field f:
        err(gpu, FP32 reference) <= tol,   or, where the reference's FP32 arithmetic itself is
        farther than that from the FP64 step on f:  err(gpu, FP64) <= err(FP32 reference, FP64),
    i.e. the engine is within tol of the reference, or at least as close to the exact step as the
    reference's own FP32 result is.  Returns the three error dictionaries."""
    e32 = field_errors(gpu, ref32, inv_dx)
    e64 = field_errors(gpu, ref64, inv_dx)
    r64 = field_errors(ref32, ref64, inv_dx)
    b32, b64, br = bounded(e32), bounded(e64), bounded(r64)
    bad = {k: (b32[k], b64[k], br[k]) for k in b32 if not (b32[k] <= tol or b64[k] <= br[k])}
    assert not bad, (f"{label}: (vs FP32, vs FP64, FP32 reference vs FP64) above {tol} and farther from FP64 "
                     f"than the reference: {bad}")
    return {"vs_fp32": b32, "vs_fp64": b64, "fp32_vs_fp64": br}
