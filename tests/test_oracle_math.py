"""Pins the oracle's small-matrix math (oracle/nclr_oracle.hpp) against numpy/scipy FP64 and
closed forms.  These checks are what make the oracle trustworthy as the parity reference for the
GPU kernels: each would fail on a plausible mistake (a transposed factor, a dropped sign fix, a
wrong sort, lambda on the identity instead of all-ones).
"""
import numpy as np
import pytest
import scipy.linalg

import oracle as O


def rand_mats(d, n, seed, kind="normal"):
    rng = np.random.default_rng(seed)
    if kind == "normal":
        return rng.normal(size=(n, d, d))
    if kind == "near_identity":
        return np.eye(d) + 0.05 * rng.normal(size=(n, d, d))
    if kind == "rank_deficient":
        a = rng.normal(size=(n, d, d))
        u, s, vt = np.linalg.svd(a)
        s[:, -1] = 0.0
        return np.einsum("nij,nj,njk->nik", u, s, vt)
    if kind == "repeated":
        q1 = np.linalg.qr(rng.normal(size=(n, d, d)))[0]
        q2 = np.linalg.qr(rng.normal(size=(n, d, d)))[0]
        s = np.ones((n, d)) * 1.3
        s[:, -1] = 0.4
        return np.einsum("nij,nj,nkj->nik", q1, s, q2)
    if kind == "negative_det":
        a = rng.normal(size=(n, d, d))
        det = np.linalg.det(a)
        a[det > 0, 0, :] *= -1
        return a
    raise ValueError(kind)


KINDS = ["normal", "near_identity", "rank_deficient", "repeated", "negative_det"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("d", [2, 3])
def test_svd_fp64_against_numpy(d, kind):
    A = rand_mats(d, 400, seed=d * 10 + KINDS.index(kind), kind=kind)
    U, S, V = O.svd(A, fp64=True)
    rec = np.einsum("nij,nj,nkj->nik", U, S, V)
    assert np.abs(rec - A).max() < 1e-12 * max(1.0, np.abs(A).max())
    eye = np.eye(d)
    assert np.abs(np.einsum("nji,njk->nik", U, U) - eye).max() < 1e-12
    assert np.abs(np.einsum("nji,njk->nik", V, V) - eye).max() < 1e-12
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.abs(np.abs(S) - ref).max() < 1e-12 * max(1.0, ref.max())
    mags = np.abs(S)
    assert (mags[:, :-1] >= mags[:, 1:] - 1e-14).all(), "singular values must be sorted by magnitude"
    if d == 3:
        # nclr_svd's column-2 flips (src/nclr_math.h:63-71): U, V in SO(3), sign of det on sigma_min
        assert np.abs(np.linalg.det(U) - 1).max() < 1e-12
        assert np.abs(np.linalg.det(V) - 1).max() < 1e-12
        assert (S[:, :2] >= 0).all()
        det = np.linalg.det(A)
        nz = np.abs(det) > 1e-10
        assert (np.sign(S[nz, 2]) == np.sign(det[nz])).all()
    else:
        # 2D: the out-of-bounds col(2) flip is treated as a no-op, singular values stay >= 0
        assert (S >= 0).all()
        assert np.abs(np.abs(np.linalg.det(U)) - 1).max() < 1e-12


@pytest.mark.parametrize("d", [2, 3])
def test_svd_fp32_accuracy(d):
    A = rand_mats(d, 2000, seed=7).astype(np.float32)
    U, S, V = O.svd(A)
    rec = np.einsum("nij,nj,nkj->nik", U.astype(np.float64), S.astype(np.float64), V.astype(np.float64))
    err = np.abs(rec - A).max(axis=(1, 2)) / np.abs(A).max(axis=(1, 2))
    assert err.max() < 2e-6
    ref = np.linalg.svd(A.astype(np.float64), compute_uv=False)
    assert (np.abs(np.abs(S) - ref).max(axis=1) / ref[:, 0]).max() < 2e-6


def test_svd_special_matrices():
    Z = np.zeros((1, 3, 3))
    U, S, V = O.svd(Z, fp64=True)
    assert (S == 0).all() and np.allclose(U, np.eye(3)) and np.allclose(V, np.eye(3))
    D = np.diag([1.0, 1.0, 0.0])[None]  # the 3D reference-mode F_0 (Q1)
    U, S, V = O.svd(D, fp64=True)
    assert np.allclose(S, [1, 1, 0])
    assert np.allclose(U @ np.diag(S[0]) @ V[0].T, D[0])
    R = scipy.linalg.expm(np.array([[0, -0.3, 0.2], [0.3, 0, -0.1], [-0.2, 0.1, 0]]))[None]
    U, S, V = O.svd(R, fp64=True)
    assert np.allclose(S, 1.0)
    N = np.array([[[np.nan, 0, 0], [0, 1, 0], [0, 0, 1]]])
    U, S, V = O.svd(N, fp64=True)
    assert np.isnan(S).all()


@pytest.mark.parametrize("d", [2, 3])
def test_polar_against_scipy(d):
    A = rand_mats(d, 300, seed=11, kind="near_identity")
    A = A[np.linalg.det(A) > 0]
    R = O.polar(A, fp64=True)
    for a, r in zip(A, R):
        u, p = scipy.linalg.polar(a)
        assert np.abs(r - u).max() < 1e-10
    assert np.abs(np.linalg.det(R) - 1).max() < 1e-12


def test_polar_2d_closed_form_and_3d_sign_convention():
    # 2D closed form (src/nclr_math.h:80-89) is the rotation maximising tr(R^T A)
    A = rand_mats(2, 300, seed=3)
    R = O.polar(A, fp64=True)
    assert np.abs(np.einsum("nji,njk->nik", R, R) - np.eye(2)).max() < 1e-12
    assert np.abs(np.linalg.det(R) - 1).max() < 1e-12
    S = np.einsum("nji,njk->nik", R, A)
    assert np.abs(S - S.transpose(0, 2, 1)).max() < 1e-12
    assert (np.trace(S, axis1=1, axis2=2) >= -1e-12).all()
    # 3D with det < 0: R = U V^T is still a rotation and R^T A is symmetric with the negative
    # eigenvalue on the smallest singular direction
    A = rand_mats(3, 300, seed=4, kind="negative_det")
    R = O.polar(A, fp64=True)
    assert np.abs(np.linalg.det(R) - 1).max() < 1e-12
    S = np.einsum("nji,njk->nik", R, A)
    assert np.abs(S - S.transpose(0, 2, 1)).max() < 1e-10
    ev = np.linalg.eigvalsh(S)
    sv = np.linalg.svd(A, compute_uv=False)
    assert np.abs(np.sort(ev, 1) - np.sort(np.stack([sv[:, 0], sv[:, 1], -sv[:, 2]], 1), 1)).max() < 1e-10


def _stress_expected(F, C, Jp, m, V0, mu, lam, dt, inv_dx, R):
    J = np.linalg.det(F)
    PF = 2 * mu * (F - R) @ F.T + lam * (J - 1) * J * np.ones_like(F)  # lambda term on every entry (Q2)
    return -(dt * V0) * 4 * inv_dx ** 2 * PF + m * C


# the hardening constants are `real` (float) in the reference (src/nclr.h:67-69), also in FP64 runs
@pytest.mark.parametrize("mat,e", [(O.JELLY, float(np.float32(0.3))), (O.LIQUID, 1.0)])
def test_stress_closed_form_scaled_identity(mat, e):
    E, nu, res, dt = 1e4, 0.2, 64, 1e-4
    mu0 = E / (2 * (1 + nu))
    lam0 = E * nu / ((1 + nu) * (1 - 2 * nu))
    for s in (0.9, 1.0, 1.07):
        F = np.eye(2) * s
        C = np.array([[0.3, -1.2], [0.7, 0.1]])
        got = O.stress(F[None], C[None], [1.0], [2.0], [0.5], mat, res=res, dt=dt, E=E, nu=nu, fp64=True)[0]
        exp = _stress_expected(F, C, 1.0, 2.0, 0.5, mu0 * e, lam0 * e, dt, res, np.eye(2))
        assert np.abs(got - exp).max() <= 1e-9 * max(1.0, np.abs(exp).max())
        if s != 1.0:
            # all-ones lambda term: the off-diagonal carries lambda (J-1) J too
            assert abs(got[0, 1] - 2.0 * C[0, 1]) > 1.0


def test_stress_closed_form_3d_and_rotation():
    E, nu, res, dt = 5e3, 0.3, 32, 1e-4
    mu0 = E / (2 * (1 + nu))
    lam0 = E * nu / ((1 + nu) * (1 - 2 * nu))
    rng = np.random.default_rng(5)
    for _ in range(20):
        Q = scipy.linalg.expm(np.cross(np.eye(3), rng.normal(size=3) * 0.5))
        S = np.diag(rng.uniform(0.8, 1.2, 3))
        F = Q @ S
        C = rng.normal(size=(3, 3))
        got = O.stress(F[None], C[None], [1.0], [1.5], [0.8], O.JELLY, res=res, dt=dt, E=E, nu=nu, fp64=True)[0]
        e = float(np.float32(0.3))
        exp = _stress_expected(F, C, 1.0, 1.5, 0.8, mu0 * e, lam0 * e, dt, res, Q)
        assert np.abs(got - exp).max() <= 1e-8 * max(1.0, np.abs(exp).max())


def test_snow_hardening_exponent():
    E, nu, res, dt = 1e4, 0.2, 64, 1e-4
    mu0 = E / (2 * (1 + nu))
    lam0 = E * nu / ((1 + nu) * (1 - 2 * nu))
    F = np.array([[1.02, 0.01], [-0.03, 0.97]])
    R = scipy.linalg.polar(F)[0]
    for Jp in (1.0, 0.9, 1.05):
        e = np.exp(10 * (1 - Jp))
        got = O.stress(F[None], np.zeros((1, 2, 2)), [Jp], [1.0], [1.0], O.SNOW, res=res, dt=dt, E=E, nu=nu,
                       fp64=True)[0]
        exp = _stress_expected(F, np.zeros((2, 2)), Jp, 1.0, 1.0, mu0 * e, lam0 * e, dt, res, R)
        assert np.abs(got - exp).max() <= 1e-9 * np.abs(exp).max()


def test_stress_zero_for_reference_identity_3d():
    """Q1: the 3D particle ctor's F = diag(1,1,0) is singular; the first-step stress is zero."""
    F = np.diag([1.0, 1.0, 0.0])[None]
    for mat in (O.SNOW, O.JELLY, O.LIQUID):
        got = O.stress(F, np.zeros((1, 3, 3)), [1.0], [1.0], [1.0], mat, res=128, fp64=True)
        assert np.abs(got).max() < 1e-9


def test_linspaced_matches_lattice_definition():
    for n, lo, hi in [(50, 0.4, 0.6), (64, 0.1, 0.3), (2, 0.25, 0.75), (7, -0.5, 0.25), (5, 0.9, 0.1)]:
        v = O.linspaced(n, lo, hi, fp64=True)
        assert v[0] == pytest.approx(lo, abs=1e-15) and v[-1] == hi
        assert np.allclose(np.diff(v), (hi - lo) / (n - 1), atol=1e-14)
    assert O.linspaced(1, 0.2, 0.7, fp64=True)[0] == 0.7
    v32 = O.linspaced(50, np.float32(0.4), np.float32(0.6))
    assert v32.dtype == np.float32 and v32[-1] == np.float32(0.6)
    assert np.abs(v32 - np.linspace(0.4, 0.6, 50)).max() < 1e-7


def test_polar_2d_of_a_scaled_reflection_is_not_finite():
    """The reference's 2D polar closed form (src/nclr_math.h:76-98) divides by
    sqrt((F00 + F11)^2 + (F10 - F01)^2), which is zero for a scaled reflection [[c, s], [s, -c]]:
    R is NaN.  2D snow reaches such an F when both singular values of a reflected F^ (nclr_svd's
    2D singular values are non-negative) are clamped to the same bound, e.g. 0.975 U V^T."""
    c, s = np.float32(0.6), np.float32(0.8)
    F = (np.float32(0.975) * np.array([[c, s], [s, -c]], np.float32)).astype(np.float32)
    assert F[0, 0] + F[1, 1] == 0 and F[1, 0] - F[0, 1] == 0
    R = O.polar(F[None])
    assert not np.isfinite(R).any()
