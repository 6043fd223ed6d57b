"""The engine's accurate SVD mode (mpm_options.fast_svd = 2): the polar factor of P2G by Newton's
iteration in FP64 and the snow / liquid SVD of G2P by an FP64 Jacobi eigen-decomposition.

Pinned by
  * mpm_svd(..., 2): the decomposition's defining properties (A = U diag(s) V^T, U and V proper
    rotations, |s| descending, the sign of det A on the last value) and numpy's FP64 SVD, to which
    it is closer than nclr_svd's FP32 two-sided Jacobi is; rank-deficient and non-finite inputs;
  * one step against the FP32 and FP64 oracles (tests/parity.py assert_accurate_parity: within
    1e-5 of the reference's FP32 step, or at least as close to the FP64 step as the reference's own
    FP32 result is, field by field), on the golden fixtures, the 3D collision scenes of every
    material and identity mode, and the 1,048,576-particle 3D snow workload in both identity modes.
The adjoint kernels keep nclr_svd's factors in this mode.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from parity import assert_accurate_parity

from nuclearmpm_amd import MaterialModel, MPMSimulation, Particles, scenes

from test_parity_gpu import MATS, collision_scene_3d, to_oracle
from test_svd_gpu import device_svd, matrices, numpy_svd_so3

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_accurate_svd_properties_and_fp64_agreement():
    A = matrices(3, 3000, seed=7)
    U, S, V = device_svd(A, exact=2)
    Ue, Se, Ve = device_svd(A, exact=1)
    A64 = A.astype(np.float64)
    rec = np.einsum("nij,nj,nkj->nik", U.astype(np.float64), S.astype(np.float64), V.astype(np.float64))
    scale = np.abs(A64).max(axis=(1, 2))[:, None, None]
    assert (np.abs(rec - A64) / scale).max() < 2e-6
    for Q in (U, V):
        Q = Q.astype(np.float64)
        assert np.abs(np.einsum("nji,njk->nik", Q, Q) - np.eye(3)).max() < 2e-6
        assert np.abs(np.linalg.det(Q) - 1.0).max() < 2e-6
    assert (np.abs(S[:, 0]) >= np.abs(S[:, 1]) * (1 - 1e-6)).all()
    assert (np.abs(S[:, 1]) >= np.abs(S[:, 2]) * (1 - 1e-6)).all()
    Un, Sn, Vn = numpy_svd_so3(A)
    err_acc = np.abs(S - Sn).max(axis=1) / np.abs(Sn).max(axis=1)
    err_exact = np.abs(Se - Sn).max(axis=1) / np.abs(Sn).max(axis=1)
    print(f"[accurate] singular values vs numpy FP64: accurate {err_acc.max():.2e}, exact {err_exact.max():.2e}",
          flush=True)
    assert err_acc.max() <= max(err_exact.max(), 2e-7)
    # the products the step uses (independent of the SVD's sign and order freedom): U V^T, U clamp(s) V^T
    def prods(U_, S_, V_):
        c = np.clip(S_, 0.975, 1.0045)
        return (np.einsum("nij,nkj->nik", U_, V_), np.einsum("nij,nj,nkj->nik", U_, c, V_))
    full = np.abs(Sn[:, 1]) > 1e-3 * np.abs(Sn[:, 0])  # rank >= 2: both products unique
    sep = full & (np.abs(np.abs(Sn[:, 1]) - np.abs(Sn[:, 2])) > 1e-3)
    for pa, pn, pe in zip(prods(U.astype(np.float64), S.astype(np.float64), V.astype(np.float64)),
                          prods(Un, Sn, Vn), prods(Ue.astype(np.float64), Se.astype(np.float64), Ve.astype(np.float64))):
        da = np.abs(pa - pn).max(axis=(1, 2))[sep]
        de = np.abs(pe - pn).max(axis=(1, 2))[sep]
        print(f"[accurate] product vs numpy FP64: accurate {da.max():.2e}, exact {de.max():.2e}", flush=True)
        assert da.max() <= max(de.max(), 1e-6)


def test_accurate_svd_degenerate_inputs():
    A = np.zeros((6, 3, 3), np.float32)
    A[1] = np.diag([1.0, 1.0, 0.0])
    A[2] = np.outer([1.0, 2.0, 3.0], [0.5, -1.0, 2.0])
    A[3] = np.eye(3)
    A[4] = np.diag([2.0, 2.0, -2.0])
    A[5, 0, 0] = np.nan
    U, S, V = device_svd(A, exact=2)
    for i in range(5):
        assert np.isfinite(U[i]).all() and np.isfinite(V[i]).all() and np.isfinite(S[i]).all(), i
        rec = U[i].astype(np.float64) @ np.diag(S[i]) @ V[i].T.astype(np.float64)
        assert np.abs(rec - A[i]).max() < 1e-5 * max(1.0, np.abs(A[i]).max()), i
        assert abs(np.linalg.det(U[i]) - 1) < 1e-5 and abs(np.linalg.det(V[i]) - 1) < 1e-5, i
    assert S[4, 2] < 0  # det < 0: the sign sits on the last singular value
    assert not np.isfinite(S[5]).any()


def accurate_case(p, mat, res, identity_mode=0, label="", E=1e4):
    print(f"[accurate] {label} n={p.n} mat={int(mat)} res={res} idm={identity_mode}", flush=True)
    with MPMSimulation(p, mat, res=res, E=E, identity_mode=identity_mode, svd="accurate") as sim:
        sim.advance(1)
        g = sim.particles().copy()
    st = to_oracle(p)
    r32 = O.run(st, int(mat), res=res, E=E, steps=1, identity_mode=identity_mode)
    r64 = O.run_threads(st.astype(np.float64), int(mat), res=res, E=E, steps=1, identity_mode=identity_mode,
                        threads=min(16, os.cpu_count() or 1), fp64=True)
    errs = assert_accurate_parity(g, r32, r64, inv_dx=float(res), label=label)
    print(f"[accurate] {label}: {errs}", flush=True)
    return errs


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_accurate_step_matches_golden(path):
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    p = Particles(*(g[f"in_{k}"] for k in ("x", "v", "F", "C", "Jp", "mass", "volume")))
    accurate_case(p, MaterialModel(int(g["material"])), int(g["res"]), int(g["identity_mode"]),
                  label=os.path.basename(path)[:-4], E=float(g["E"]))


@pytest.mark.parametrize("identity_mode", [0, 1], ids=["reference_identity", "physical_identity"])
@pytest.mark.parametrize("mat", MATS, ids=["snow", "jelly", "liquid"])
def test_accurate_step_3d_collision(mat, identity_mode):
    p = collision_scene_3d(identity_mode=identity_mode)
    with MPMSimulation(p, mat, res=32, identity_mode=identity_mode, svd="accurate") as sim:
        sim.advance(40)
        s = sim.particles().copy()
    assert np.isfinite(s.x).all()
    accurate_case(s, mat, 32, identity_mode, label=f"collision3d after 40 {mat.name} idm{identity_mode}")


@pytest.mark.parametrize("identity_mode", [0, 1], ids=["reference_identity", "physical_identity"])
def test_accurate_full_size_four_cubes_3d(identity_mode):
    """The 1,048,576-particle 3D snow workload from its state after 30 engine steps (accurate mode)."""
    sc = scenes.four_cubes_3d()
    p = Particles.from_positions(sc.particles.x, v=sc.particles.v, identity_mode=identity_mode)
    with MPMSimulation(p, sc.material, res=sc.res, identity_mode=identity_mode, svd="accurate") as sim:
        sim.advance(30)
        s = sim.particles().copy()
    accurate_case(s, sc.material, sc.res, identity_mode, label=f"four_cubes_3d idm{identity_mode} after 30")
