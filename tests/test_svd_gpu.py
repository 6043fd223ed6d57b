"""The device SVD of the step (nclr_svd, src/nclr_math.h:50-74) through mpm_svd.

exact mode: bitwise the oracle's FP32 restatement of nclr_svd (Eigen's two-sided Jacobi order).
fast mode (3D default): pinned by what the decomposition must satisfy -- A = U diag(s) V^T, U and
V proper rotations, |s| descending with the sign of det A on the last value, the singular values
of numpy's FP64 SVD -- and by the convention-independent products the step uses (the polar factor
U V^T and the snow clamp U clamp(s) V^T) against numpy's FP64 SVD.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

from nuclearmpm_amd._lib import check, lib, ptr

pytestmark = pytest.mark.gpu


def device_svd(A, exact):
    A = np.ascontiguousarray(A, np.float32)
    n, d, _ = A.shape
    U, V = np.empty_like(A), np.empty_like(A)
    S = np.empty((n, d), np.float32)
    check(lib().mpm_svd(d, n, ptr(A), ptr(U), ptr(S), ptr(V), int(exact), 0))
    return U, S, V


def matrices(d, n=4000, seed=0):
    rng = np.random.default_rng(seed)
    out = [rng.normal(size=(n, d, d)),
           np.eye(d)[None] + rng.normal(0, 0.01, (n, d, d)),
           np.eye(d)[None] + rng.normal(0, 0.3, (n, d, d))]
    if d == 3:
        r2 = np.diag([1.0, 1.0, 0.0])[None] + rng.normal(0, 0.05, (n, 3, 3))
        r2[:, 2, :] = 0.0  # the reference default F = diag(1,1,0) and its descendants
        out.append(r2)
        q1 = np.eye(3)[None] + rng.normal(0, 0.005, (n, 3, 3))
        q1[:, 2, :] *= 1e-3  # snow F-hat under diag(1,1,0): a tiny third row
        out.append(q1)
        neg = np.diag([1.0, 1.0, -1.0])[None] + rng.normal(0, 0.1, (n, 3, 3))
        out.append(neg)
    return np.concatenate(out).astype(np.float32)


@pytest.mark.parametrize("d", [2, 3])
def test_exact_mode_is_bitwise_the_oracle(d):
    A = matrices(d, 2000, seed=1)
    U, S, V = device_svd(A, exact=True)
    Uo, So, Vo = O.svd(A)
    assert np.array_equal(S, So) and np.array_equal(U, Uo) and np.array_equal(V, Vo)


def numpy_svd_so3(A):
    A = A.astype(np.float64)
    U, s, Vt = np.linalg.svd(A)
    fix = np.linalg.det(U) < 0
    U[fix, :, 2] *= -1
    s[fix, 2] *= -1
    fix = np.linalg.det(Vt) < 0
    Vt[fix, 2, :] *= -1
    s[fix, 2] *= -1
    return U, s, np.swapaxes(Vt, 1, 2)


def test_fast_mode_3d_properties_and_products():
    A = matrices(3, 4000, seed=2)
    U, S, V = device_svd(A, exact=False)
    A64, U64, S64, V64 = (x.astype(np.float64) for x in (A, U, S, V))
    scale = np.abs(A64).max(axis=(1, 2))
    rec = np.einsum("nij,nj,nkj->nik", U64, S64, V64)
    assert (np.abs(rec - A64).max(axis=(1, 2)) <= 4e-6 * scale).all()
    eye = np.eye(3)[None]
    assert np.abs(np.einsum("nji,njk->nik", U64, U64) - eye).max() < 2e-6
    assert np.abs(np.einsum("nji,njk->nik", V64, V64) - eye).max() < 2e-6
    assert (np.linalg.det(U64) > 0).all() and (np.linalg.det(V64) > 0).all()
    assert (S64[:, 0] >= S64[:, 1] - 1e-6 * scale).all() and (S64[:, 1] >= np.abs(S64[:, 2]) - 1e-6 * scale).all()
    detA = np.linalg.det(A64)
    clear = np.abs(detA) > 1e-4 * scale ** 3
    assert (np.sign(S64[clear, 2]) == np.sign(detA[clear])).all()
    Un, sn, Vn = numpy_svd_so3(A64)
    assert (np.abs(S64 - sn).max(axis=1) <= 2e-6 * scale).all()
    # the step's products: polar factor and snow clamp, away from the ill-conditioned set
    # sigma_2 + sigma_3 ~ 0 (where U V^T is not continuous)
    ok = (sn[:, 1] + sn[:, 2]) > 1e-2 * scale
    R = np.einsum("nij,nkj->nik", U64, V64)
    Rn = np.einsum("nij,nkj->nik", Un, Vn)
    assert np.abs(R - Rn)[ok].max() < 2e-5
    cl = lambda s: np.clip(s, np.float32(1 - 2.5e-2), np.float32(1 + 4.5e-3))
    Fc = np.einsum("nij,nj,nkj->nik", U64, cl(S64), V64)
    Fn = np.einsum("nij,nj,nkj->nik", Un, cl(sn), Vn)
    assert np.abs(Fc - Fn)[ok].max() < 2e-5


def test_fast_mode_3d_degenerate_inputs_stay_finite_and_orthonormal():
    A = np.zeros((6, 3, 3), np.float32)
    A[1] = np.diag([1.0, 0.0, 0.0])
    A[2] = np.diag([2.0, 2.0, 2.0])
    A[3] = np.diag([1.0, 1.0, 0.0])
    A[4] = -np.eye(3)
    A[5] = np.outer([1.0, 2.0, 3.0], [0.5, -1.0, 2.0])
    U, S, V = device_svd(A, exact=False)
    assert np.isfinite(U).all() and np.isfinite(S).all() and np.isfinite(V).all()
    for i in range(len(A)):
        np.testing.assert_allclose(U[i].T @ U[i], np.eye(3), atol=2e-6)
        np.testing.assert_allclose(V[i].T @ V[i], np.eye(3), atol=2e-6)
        np.testing.assert_allclose((U[i] * S[i]) @ V[i].T, A[i], atol=4e-6 * max(1.0, np.abs(A[i]).max()))
    np.testing.assert_allclose(S[4], [1, 1, -1], atol=1e-6)
    np.testing.assert_allclose(S[3], [1, 1, 0], atol=1e-6)
