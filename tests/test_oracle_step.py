"""Pins the oracle's step (oracle/nclr_oracle.hpp, restating src/nclr.h:104-310) with
conservation laws, closed forms and exact-reproduction properties of the quadratic B-spline
APIC transfer.  All checks run the FP64 oracle unless stated otherwise.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle as O


def random_state(n, d, seed, lo=0.3, hi=0.7, strain=0.05, identity_mode=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, size=(n, d))
    F = np.eye(d) + strain * rng.normal(size=(n, d, d))
    return O.State(x=x, v=rng.normal(size=(n, d)), F=F, C=rng.normal(size=(n, d, d)) * 3,
                   Jp=rng.uniform(0.95, 1.05, n), mass=rng.uniform(0.5, 2.0, n), volume=rng.uniform(0.5, 2.0, n))


def node_positions(d, res):
    side = np.arange(res + 1) / res
    return np.stack(np.meshgrid(*([side] * d), indexing="ij"), -1).reshape(-1, d)


@pytest.mark.parametrize("mat", [O.SNOW, O.JELLY, O.LIQUID])
@pytest.mark.parametrize("d", [2, 3])
def test_p2g_conserves_mass_momentum_and_angular_momentum(d, mat):
    res = 32 if d == 3 else 64
    s = random_state(200, d, seed=d + 3 * mat)
    mom, m = O.p2g(s, mat, res=res, identity_mode=1, fp64=True)
    assert abs(m.sum() - s.mass.sum()) < 1e-12 * s.mass.sum()
    # the affine term sums to zero since sum_i w_i (x_i - x_p) = 0
    p = (s.mass[:, None] * s.v).sum(0)
    assert np.abs(mom.sum(0) - p).max() < 1e-10 * np.abs(p).max()
    xn = node_positions(d, res)
    # first moment: sum_i m_i x_i = sum_p m_p x_p
    assert np.abs((m[:, None] * xn).sum(0) - (s.mass[:, None] * s.x).sum(0)).max() < 1e-11
    # angular momentum: sum_i x_i x mom_i = sum_p [x_p x m_p v_p + (dx^2/4) eps : A_p] since
    # sum_i w_i (x_i - x_p)(x_i - x_p)^T = (dx^2/4) I for the quadratic B-spline; this pins the
    # orientation of A.dpos (a transposed A flips the sign of the last term)
    A = O.stress(s.F, s.C, s.Jp, s.mass, s.volume, mat, res=res, fp64=True)
    q = 0.25 / res ** 2
    if d == 2:
        Lg = (xn[:, 0] * mom[:, 1] - xn[:, 1] * mom[:, 0]).sum()
        Lp = (s.x[:, 0] * s.mass * s.v[:, 1] - s.x[:, 1] * s.mass * s.v[:, 0]).sum() + q * (A[:, 1, 0] - A[:, 0, 1]).sum()
        assert abs(Lg - Lp) < 1e-10 * max(1.0, abs(Lp))
    else:
        Lg = np.cross(xn, mom).sum(0)
        skew = np.stack([A[:, 2, 1] - A[:, 1, 2], A[:, 0, 2] - A[:, 2, 0], A[:, 1, 0] - A[:, 0, 1]], 1)
        Lp = np.cross(s.x, s.mass[:, None] * s.v).sum(0) + q * skew.sum(0)
        assert np.abs(Lg - Lp).max() < 1e-10 * max(1.0, np.abs(Lp).max())


@pytest.mark.parametrize("d", [2, 3])
def test_p2g_single_particle_weights(d):
    """Mass of one particle spreads with the quadratic B-spline: 3^d nodes, second moment dx^2/4."""
    res = 16
    rng = np.random.default_rng(1)
    for _ in range(10):
        x = rng.uniform(0.2, 0.8, size=(1, d))
        s = O.default_state(x.astype(np.float32), identity_mode=1).astype(np.float64)
        s.x = x
        mom, m = O.p2g(s, O.JELLY, res=res, identity_mode=1, fp64=True)
        assert (m > 0).sum() <= 3 ** d
        dxn = node_positions(d, res) - x
        assert abs(m.sum() - 1) < 1e-14
        second = (m[:, None, None] * dxn[:, :, None] * dxn[:, None, :]).sum(0)
        assert np.abs(second - np.eye(d) * 0.25 / res ** 2).max() < 1e-14
        base = np.trunc(x[0] * res - 0.5).astype(int)  # Q4: truncation, here equal to floor
        idx = np.nonzero(m)[0]
        coords = np.stack(np.unravel_index(idx, (res + 1,) * d), 1)
        assert (coords >= base).all() and (coords <= base + 2).all()


def _grid(d, res):
    nodes = (res + 1) ** d
    return np.zeros((nodes, d)), np.zeros(nodes)


def _flat(n, res):
    return int(np.ravel_multi_index(tuple(n), (res + 1,) * len(n)))


@pytest.mark.parametrize("d", [2, 3])
def test_grid_op_closed_form(d):
    res, dt, g = 64, 1e-4, -100.0
    allowed = np.float32(np.float64(np.float32(1.0 / res)) * 0.9 / np.float64(np.float32(dt)))
    mom, m = _grid(d, res)
    mid = [res // 2] * d
    i_mid = _flat(mid, res)
    mom[i_mid] = np.arange(1, d + 1) * 4.0
    m[i_mid] = 2.0
    fast = list(mid)
    fast[0] += 1
    i_fast = _flat(fast, res)
    mom[i_fast] = 1e9
    m[i_fast] = 1.0
    # sticky band (kBoundary = 3): node index < 3 moving outward, or >= res-2 moving outward
    cases = []
    for axis in range(d):
        for n_a, vel, zeroed in [(2, -1.0, True), (2, 1.0, False), (3, -1.0, False), (res - 2, 1.0, True),
                                 (res - 3, 1.0, False), (res - 2, -1.0, False)]:
            n = [res // 2 + 5] * d
            n[axis] = n_a
            n[(axis + 1) % d] += 3 + len(cases)
            i = _flat(n, res)
            mom[i] = 0.0
            mom[i, axis] = vel
            if axis == 1:
                mom[i, 1] = vel + 1.0 if vel > 0 else vel  # keep the sign after gravity
            m[i] = 1.0
            cases.append((i, axis, zeroed))
    # a massless node with momentum is not normalised; the box acts on the raw value
    raw = [1] * d
    raw[0] = 1
    i_raw = _flat([1] + [res // 2] * (d - 1), res)
    mom[i_raw, 0] = -3.0
    v, mm = O.grid_op(mom, m, d, res, dt=dt, gravity=g, fp64=True)
    exp = np.arange(1, d + 1) * 2.0
    exp[1] += dt * g
    assert np.allclose(v[i_mid], exp, rtol=1e-7) and mm[i_mid] == 2.0
    assert (v[i_fast] == np.float64(allowed)).all(), "velocity clamp at 0.9 dx/dt"
    for i, axis, zeroed in cases:
        if zeroed:
            assert (v[i] == 0).all() and mm[i] == 0, (i, axis)
        else:
            assert mm[i] == 1.0 and v[i, axis] != 0, (i, axis)
    assert (v[i_raw] == 0).all() and mm[i_raw] == 0
    # gravity alone pushes a resting bottom node outward -> zeroed (axis 1 only, Q7)
    mom2, m2 = _grid(d, res)
    n = [res // 2] * d
    n[1] = 1
    j = _flat(n, res)
    m2[j] = 1.0
    v2, mm2 = O.grid_op(mom2, m2, d, res, fp64=True)
    assert mm2[j] == 0 and (v2[j] == 0).all()
    n[1] = 10
    j = _flat(n, res)
    m2[:] = 0
    m2[j] = 1.0
    v2, mm2 = O.grid_op(mom2, m2, d, res, fp64=True)
    assert v2[j, 1] == pytest.approx(dt * g) and (np.delete(v2[j], 1) == 0).all()


@pytest.mark.parametrize("d", [2, 3])
def test_g2p_affine_reproduction(d):
    """A linear grid field v(x) = A x + b is reproduced exactly: v_p = A x_p + b and C_p = A."""
    res, dt = 32, 1e-4
    rng = np.random.default_rng(d)
    A = rng.normal(size=(d, d))
    b = rng.normal(size=d)
    xn = node_positions(d, res)
    gv = xn @ A.T + b
    s = random_state(100, d, seed=9)
    out = O.g2p(s, gv, O.JELLY, res=res, dt=dt, identity_mode=1, fp64=True)
    vp = s.x @ A.T + b
    assert np.abs(out.v - vp).max() < 1e-12
    assert np.abs(out.C - A).max() < 1e-10
    assert np.abs(out.x - (s.x + dt * vp)).max() < 1e-15
    Fexp = np.einsum("ij,njk->nik", np.eye(d) + dt * A, s.F)
    assert np.abs(out.F - Fexp).max() < 1e-12
    # uniform field: v_p = u, C = 0
    u = rng.normal(size=d)
    out = O.g2p(s, np.broadcast_to(u, xn.shape).copy(), O.JELLY, res=res, identity_mode=1, fp64=True)
    assert np.abs(out.v - u).max() < 1e-13 and np.abs(out.C).max() < 1e-10


def _clamped_F(Fh, d):
    U, S, Vt = np.linalg.svd(Fh)
    if d == 3:
        # signed SVD: U, V in SO(3), the sign of det on the smallest singular value
        dU, dV = np.linalg.det(U), np.linalg.det(Vt)
        S = S.copy()
        for k in range(len(Fh)):
            if dU[k] < 0:
                U[k, :, 2] *= -1
                S[k, 2] *= -1
            if dV[k] < 0:
                Vt[k, 2, :] *= -1
                S[k, 2] *= -1
    Sc = np.clip(S, float(np.float32(0.975)), float(np.float32(1.0045)))
    return np.einsum("nij,nj,njk->nik", U, Sc, Vt), Sc


@pytest.mark.parametrize("identity_mode", [0, 1])
@pytest.mark.parametrize("d", [2, 3])
def test_g2p_snow_plasticity_against_numpy_svd(d, identity_mode):
    if d == 2 and identity_mode == 0:
        pytest.skip("2D has one identity")
    res, dt = 32, 1e-3
    rng = np.random.default_rng(4)
    A = rng.normal(size=(d, d)) * 20.0  # large enough to push sigma past the clamps
    xn = node_positions(d, res)
    s = random_state(100, d, seed=5, strain=0.01)
    out = O.g2p(s, xn @ A.T, O.SNOW, res=res, dt=dt, identity_mode=identity_mode, fp64=True)
    I = np.eye(d)
    if identity_mode == 0 and d == 3:
        I[2, 2] = 0.0  # Q1: diag<3>(1) = diag(1,1,0)
    Fh = np.einsum("ij,njk->nik", I + dt * A, s.F)
    Fexp, Sc = _clamped_F(Fh, d)
    assert np.abs(out.F - Fexp).max() < 1e-10
    # the clamp bounds are real(0.6), real(20.0): float constants (src/nclr.h:248)
    Jexp = np.clip(s.Jp * np.linalg.det(Fh) / Sc.prod(1), float(np.float32(0.6)), float(np.float32(20.0)))
    assert np.abs(out.Jp - Jexp).max() < 1e-10
    assert (out.Jp >= 0.6).all() and (out.Jp <= 20).all()


@pytest.mark.parametrize("identity_mode", [0, 1])
@pytest.mark.parametrize("d", [2, 3])
def test_g2p_liquid_reset(d, identity_mode):
    res, dt = 32, 1e-3
    rng = np.random.default_rng(6)
    A = rng.normal(size=(d, d)) * 5.0
    xn = node_positions(d, res)
    s = random_state(50, d, seed=8)
    out = O.g2p(s, xn @ A.T, O.LIQUID, res=res, dt=dt, identity_mode=identity_mode, fp64=True)
    I = np.eye(d)
    if identity_mode == 0 and d == 3:
        I[2, 2] = 0.0
    Fh = np.einsum("ij,njk->nik", I + dt * A, s.F)
    det = np.linalg.det(Fh)
    J = det if d == 3 else np.abs(det)  # 2D singular values stay unsigned
    exp = np.broadcast_to(I, Fh.shape).copy()
    exp[:, 0, 0] = J
    assert np.abs(out.F - exp).max() < 1e-10
    assert (out.Jp == s.Jp).all(), "Jp only changes for snow (Q10)"


def test_reference_identity_3d_snow_first_step():
    """Q1 in 3D: F_0 = diag(1,1,0) makes sigma_3 = 0 -> clamped to 0.975, Jp -> 0.6 after one step."""
    lin = O.linspaced(6, 0.4, 0.6)
    x = np.stack(np.meshgrid(lin, lin, lin, indexing="ij"), -1).reshape(-1, 3)
    s = O.default_state(x, identity_mode=0)
    out = O.run(s, O.SNOW, res=32, steps=1)
    assert np.allclose(out.Jp, 0.6)
    assert np.allclose(out.F[:, 2, 2], np.float32(0.975), atol=1e-5)
    s1 = O.default_state(x, identity_mode=1)
    out1 = O.run(s1, O.SNOW, res=32, steps=1, identity_mode=1)
    assert np.allclose(out1.Jp, 1.0, atol=1e-5)


def test_out_of_range_semantics():
    """std::vector::at throws on a flattened overflow (axis 0 past the grid); the faster axes alias."""
    s = O.default_state(np.array([[0.5, 0.5], [0.99, 0.5]], np.float32))
    with pytest.raises(O.OracleOutOfRange):
        O.run(s, O.SNOW, res=16)
    s = O.default_state(np.array([[0.5, 0.5], [0.5, 0.99]], np.float32))
    O.run(s, O.SNOW, res=16)  # aliases into the next row, no exception (as in the reference)
    s = O.default_state(np.array([[np.nan, 0.5]], np.float32))
    with pytest.raises(O.OracleOutOfRange):
        O.run(s, O.JELLY, res=16)


def test_fp32_tracks_fp64_over_a_short_run():
    lin = O.linspaced(30, 0.4, 0.6)
    x = np.stack(np.meshgrid(lin, lin, indexing="ij"), -1).reshape(-1, 2)
    s = O.default_state(x)
    a = O.run(s, O.SNOW, res=64, steps=20)
    b = O.run(s, O.SNOW, res=64, steps=20, fp64=True)
    assert np.abs(a.x - b.x).max() < 1e-6
    assert np.abs(a.F - b.F).max() < 1e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_parallel_step_matches_serial(dim):
    """The OpenMP step (the multi-core CPU baseline and the FP64 finite-difference rollouts) is
    bitwise the serial step: each thread adds the contributions that land in its band of grid rows
    in particle order.  FP32 fast build and exact FP64 build, every material, with thread counts
    that split the grid unevenly."""
    side = 20 if dim == 2 else 8
    g = np.linspace(0.3, 0.45, side, dtype=np.float32)
    a = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), -1).reshape(-1, dim)
    x = np.concatenate([a, a + np.float32(0.2)]).astype(np.float32)
    n = len(x)
    v = np.zeros_like(x)
    v[: n // 2, 0] = 2.0
    v[n // 2:, 0] = -2.0
    F = np.broadcast_to(np.eye(dim, dtype=np.float32), (n, dim, dim)).copy()
    st = O.State(x=x, v=v, F=F, C=np.zeros_like(F), Jp=np.ones(n, np.float32), mass=np.ones(n, np.float32),
                 volume=np.ones(n, np.float32))
    for mat in (O.SNOW, O.JELLY, O.LIQUID):
        for fp64 in (False, True):
            ser = O.run_threads(st, mat, res=32, steps=3, identity_mode=1, threads=1, fp64=fp64)
            for threads in (3, 7):
                par = O.run_threads(st, mat, res=32, steps=3, identity_mode=1, threads=threads, fp64=fp64)
                for k in ("x", "v", "F", "C", "Jp"):
                    np.testing.assert_array_equal(getattr(par, k), getattr(ser, k), err_msg=f"{mat} {k} {fp64}")
    # the exact FP64 build's threaded step is the exact serial step
    ref = O.run(st, O.JELLY, res=32, steps=3, identity_mode=1, fp64=True)
    par = O.run_threads(st, O.JELLY, res=32, steps=3, identity_mode=1, threads=5, fp64=True)
    np.testing.assert_array_equal(par.x, ref.x)


def test_parallel_step_raises_out_of_range():
    x = np.array([[0.5, 0.5], [1.2, 0.5], [0.3, 0.3]], np.float32)
    st = O.default_state(x)
    with pytest.raises(O.OracleOutOfRange):
        O.run_threads(st, O.SNOW, res=16, steps=1, threads=4, fp64=True)
