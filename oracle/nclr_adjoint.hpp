// Test infrastructure only: reverse-mode adjoint of one step of the CPU oracle (nclr_oracle.hpp).
//
// The reference has no derivatives; this is the FP64 CPU adjoint the GPU adjoint kernels are
// checked against.  It differentiates the restated step (src/nclr.h:104-310) for every material,
// holding mass and volume fixed (shown for jelly, F' = (I' + dt C') F, src/nclr.h:230-234):
//
//   P2G   mom_N = sum_p w (m v + A dpos),  mass_N = sum_p w m,
//         A = -dt V0 4 inv_dx^2 [2 mu (F - R) F^T + lambda (J-1) J 1] + m C,  R = polar(F)
//   grid  v_N = clamp(mom_N / mass_N + dt g e_1), zeroed in the sticky band
//   G2P   v' = sum w v_N,  C' = 4 inv_dx sum w v_N (o - fx)^T,  x' = x + dt v',  F' = (I' + dt C') F
//
// with the weights w (quadratic B-spline of fx = x inv_dx - base) differentiated as well.  The
// clamp and the sticky band contribute zero derivative where active; the polar factor's
// derivative uses R^T dR = Y skew with S Y + Y S = skew(R^T dF) - ..., i.e. for a cotangent
// Rbar: Fbar += 2 R Y, S Y + Y S = skew(R^T Rbar), S = R^T F.
//
// Snow (src/nclr.h:235-251): F^ = (I' + dt C') F = U S V^T (nclr_svd, signed S), F' = U f(S) V^T with
// f = clamp(., 1 - 2.5e-2, 1 + 4.5e-3), Jp' = clamp(Jp det F^ / det F', 0.6, 20), and the P2G
// hardening mu = mu0 e, lambda = lambda0 e, e = exp(10 (1 - Jp)).  The backward pass of
// U f(S) V^T: with P = U^T Fbar' V, the cotangent of F^ is U Q V^T, Q_ii = f'(s_i) P_ii and, for
// i != j, Q_ij = a_ij (P_ij + P_ji) / 2 + b_ij (P_ij - P_ji) / 2 with the divided differences
// a_ij = (f_i - f_j) / (s_i - s_j) (f'(s_i) when s_i = s_j) and b_ij = (f_i + f_j) / (s_i + s_j).
// Liquid (src/nclr.h:252-258): F' = diag<dim>(1) with F'_00 = J = prod S (= |det F^| in 2D, where
// nclr_svd's singular values are non-negative, det F^ in 3D), so Fbar^ = Fbar'_00 dJ/dF^.
//
// Pinned by central finite differences of the FP64 oracle step (tests/test_oracle_adjoint.py).
#pragma once

#include "nclr_oracle.hpp"

namespace oracle {

template <typename T, int D>
struct StepGrads {
    std::vector<T> x, v, F, C, Jp;  // particle-major: x[n][D], F[n][D][D], Jp[n]
    T E = T(0);
    T E_abs = T(0);  // sum over particles of |their contribution to E|: the scale of E's rounding
    explicit StepGrads(size_t n = 0)
        : x(n * D, T(0)), v(n * D, T(0)), F(n * D * D, T(0)), C(n * D * D, T(0)), Jp(n, T(0)) {}
};

// Fbar += the cotangent of A through A -> U f(S) V^T (the snow clamp of the singular values).
template <typename T, int D>
inline void clamp_svd_backward(const T U[D][D], const T sig[D], const T V[D][D], T lo, T hi, const T Gbar[D][D],
                               T Fbar[D][D]) {
    T P[D][D], Q[D][D], f[D], fp[D];
    for (int i = 0; i < D; ++i) {
        f[i] = std::clamp(sig[i], lo, hi);
        fp[i] = (sig[i] > lo && sig[i] < hi) ? T(1) : T(0);
    }
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            T s = T(0);
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) s += U[a][i] * Gbar[a][b] * V[b][j];
            P[i][j] = s;
        }
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            if (i == j) {
                Q[i][i] = fp[i] * P[i][i];
                continue;
            }
            const T ds = sig[i] - sig[j];
            const T a = ds != T(0) ? (f[i] - f[j]) / ds : fp[i];
            const T b = (f[i] + f[j]) / (sig[i] + sig[j]);
            Q[i][j] = a * T(0.5) * (P[i][j] + P[j][i]) + b * T(0.5) * (P[i][j] - P[j][i]);
        }
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) {
            T s = T(0);
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j) s += U[a][i] * Q[i][j] * V[b][j];
            Fbar[a][b] += s;
        }
}

// Quadratic B-spline weight derivatives d w_k / d fx.
template <typename T>
inline void bspline_dweights(T fx, T dw[3]) {
    dw[0] = -(T(1.5) - fx);
    dw[1] = -T(2) * (fx - T(1));
    dw[2] = fx - T(0.5);
}

template <typename T>
inline void cofactor(const T a[2][2], T c[2][2]) {
    c[0][0] = a[1][1];
    c[0][1] = -a[1][0];
    c[1][0] = -a[0][1];
    c[1][1] = a[0][0];
}

template <typename T>
inline void cofactor(const T a[3][3], T c[3][3]) {
    c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    c[0][1] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    c[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    c[1][0] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    c[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    c[1][2] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    c[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    c[2][1] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
}

// Fbar contribution of a cotangent Rbar on the rotation factor R of F (polar decomposition).
template <typename T, int D>
inline void polar_backward(const T F[D][D], const T R[D][D], const T Rbar[D][D], T Fbar[D][D]) {
    T S[D][D], M[D][D];
    // S = R^T F, M = R^T Rbar
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            T s = T(0), m = T(0);
            for (int k = 0; k < D; ++k) {
                s += R[k][i] * F[k][j];
                m += R[k][i] * Rbar[k][j];
            }
            S[i][j] = s;
            M[i][j] = m;
        }
    T Y[D][D] = {};
    if constexpr (D == 2) {
        const T m = T(0.5) * (M[1][0] - M[0][1]);
        const T y = m / (S[0][0] + S[1][1]);
        Y[1][0] = y;
        Y[0][1] = -y;
    } else {
        // skew(M) <-> vector m; S Y + Y S = skew(M)  <=>  (tr(S) I - S) y = m
        const T mv[3] = {T(0.5) * (M[2][1] - M[1][2]), T(0.5) * (M[0][2] - M[2][0]), T(0.5) * (M[1][0] - M[0][1])};
        const T tr = S[0][0] + S[1][1] + S[2][2];
        T Q[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Q[i][j] = (i == j ? tr : T(0)) - T(0.5) * (S[i][j] + S[j][i]);
        T Qc[3][3];
        cofactor<T>(Q, Qc);
        const T detQ = det<T>(Q);
        T y[3];
        for (int i = 0; i < 3; ++i) y[i] = (Qc[0][i] * mv[0] + Qc[1][i] * mv[1] + Qc[2][i] * mv[2]) / detQ;
        Y[2][1] = y[0];
        Y[1][2] = -y[0];
        Y[0][2] = y[1];
        Y[2][0] = -y[1];
        Y[1][0] = y[2];
        Y[0][1] = -y[2];
    }
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            T s = T(0);
            for (int k = 0; k < D; ++k) s += R[i][k] * Y[k][j];
            Fbar[i][j] += T(2) * s;
        }
}

// One-step vector-Jacobian product: given the cotangents `out` of the state after the step,
// returns the cotangents `in` of the state before it and of E.
template <typename T, int D>
void adjoint_step(const std::vector<Particle<T, D>>& ps, int material, int res, T dt, T E, T nu, T gravity, int idm,
                  const StepGrads<T, D>& out, StepGrads<T, D>& in) {
    const size_t n = ps.size();
    in = StepGrads<T, D>(n);
    Simulation<T, D> sim(ps, material, res, dt, E, nu, gravity, idm);
    sim.p2g();
    const std::vector<Cell<T, D>> pre = sim.grid();  // momentum, mass
    sim.grid_op();
    const std::vector<Cell<T, D>> post = sim.grid();  // velocity, mass after the sticky box
    const T inv_dx = sim.inv_dx(), dx = sim.dx();
    const T four_inv_dx = T(4) * inv_dx;
    std::vector<Cell<T, D>> vbar(pre.size(), Cell<T, D>{});
    for (auto& c : vbar) {
        for (int a = 0; a < D; ++a) c.v[a] = T(0);
        c.m = T(0);
    }
    T Iprime[D][D];
    diag_ref<T, D>(Iprime, T(1), idm);

    // --- G2P backward ------------------------------------------------------------------------
    for (size_t p = 0; p < n; ++p) {
        const auto& P = ps[p];
        int base[D];
        T fx[D], w[D][3], dw[D][3];
        sim.base_and_weights(P.x, base, fx, w);
        for (int a = 0; a < D; ++a) bspline_dweights<T>(fx[a], dw[a]);
        // forward C' (needed for F' = (I' + dt C') F)
        T Cn[D][D] = {};
        sim.each_offset([&](const int off[D]) {
            const auto& g = post[size_t(sim.node_index(base, off))];
            T weight = w[0][off[0]];
            for (int a = 1; a < D; ++a) weight *= w[a][off[a]];
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) Cn[a][b] += four_inv_dx * weight * g.v[a] * (T(off[b]) - fx[b]);
        });
        // F^ = (I' + dt C') F and the cotangent of F^ from those of F' and Jp'
        T Mh[D][D], Fh[D][D], Fhbar[D][D];
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) Mh[a][b] = Iprime[a][b] + dt * Cn[a][b];
        matmul<T, D>(Mh, P.F, Fh);
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) Fhbar[a][b] = material == kJelly ? out.F[(p * D + a) * D + b] : T(0);
        in.Jp[p] += material == kSnow ? T(0) : out.Jp[p];  // jelly and liquid keep Jp
        if (material == kLiquid) {
            T U[D][D], sig[D], V[D][D], cof[D][D];
            nclr_svd<T, D>(Fh, U, sig, V);
            T J = T(1);
            for (int d = 0; d < D; ++d) J *= sig[d];
            cofactor<T>(Fh, cof);
            const T sgn = J * det<T>(Fh) < T(0) ? T(-1) : T(1);
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) Fhbar[a][b] = out.F[(p * D) * D] * sgn * cof[a][b];
        } else if (material == kSnow) {
            T U[D][D], sig[D], V[D][D];
            nclr_svd<T, D>(Fh, U, sig, V);
            const T lo = T(float(1.0 - 2.5e-2));
            const T hi = T(float(1.0 + 4.5e-3));
            T US[D][D], Fn[D][D];
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) US[a][b] = U[a][b] * std::clamp(sig[b], lo, hi);
            matmul_bt<T, D>(US, V, Fn);
            const T old_J = det<T>(Fh), new_J = det<T>(Fn);
            const T r = P.Jp * old_J / new_J;
            const T rbar = (r > T(0.6f) && r < T(20.0f)) ? out.Jp[p] : T(0);
            in.Jp[p] += rbar * old_J / new_J;
            T cof_h[D][D], cof_n[D][D], Gn[D][D];
            cofactor<T>(Fh, cof_h);
            cofactor<T>(Fn, cof_n);
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) {
                    Fhbar[a][b] += rbar * P.Jp / new_J * cof_h[a][b];
                    Gn[a][b] = out.F[(p * D + a) * D + b] - rbar * P.Jp * old_J / (new_J * new_J) * cof_n[a][b];
                }
            clamp_svd_backward<T, D>(U, sig, V, lo, hi, Gn, Fhbar);
        }
        T vbar_t[D], Cbar_t[D][D];
        for (int a = 0; a < D; ++a) {
            vbar_t[a] = out.v[p * D + a] + dt * out.x[p * D + a];
            in.x[p * D + a] += out.x[p * D + a];
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                T s = T(0), fb = T(0);
                for (int k = 0; k < D; ++k) {
                    s += Fhbar[a][k] * P.F[b][k];               // (Fbar^ F^T)_ab
                    fb += Mh[k][a] * Fhbar[k][b];               // (M^T Fbar^)_ab
                }
                Cbar_t[a][b] = out.C[(p * D + a) * D + b] + dt * s;
                in.F[(p * D + a) * D + b] += fb;
            }
        T fxbar[D] = {};
        sim.each_offset([&](const int off[D]) {
            auto& gb = vbar[size_t(sim.node_index(base, off))];
            const auto& g = post[size_t(sim.node_index(base, off))];
            T weight = w[0][off[0]];
            for (int a = 1; a < D; ++a) weight *= w[a][off[a]];
            T d[D];
            for (int a = 0; a < D; ++a) d[a] = T(off[a]) - fx[a];
            T wbar = T(0);
            for (int a = 0; a < D; ++a) {
                T cd = T(0);
                for (int b = 0; b < D; ++b) cd += Cbar_t[a][b] * d[b];
                gb.v[a] += weight * vbar_t[a] + four_inv_dx * weight * cd;
                wbar += g.v[a] * vbar_t[a] + four_inv_dx * g.v[a] * cd;
            }
            for (int b = 0; b < D; ++b) {
                T ctv = T(0);
                for (int a = 0; a < D; ++a) ctv += Cbar_t[a][b] * g.v[a];
                fxbar[b] -= four_inv_dx * weight * ctv;
            }
            for (int a = 0; a < D; ++a) {
                T dwa = dw[a][off[a]];
                for (int b = 0; b < D; ++b)
                    if (b != a) dwa *= w[b][off[b]];
                fxbar[a] += wbar * dwa;
            }
        });
        for (int a = 0; a < D; ++a) in.x[p * D + a] += inv_dx * fxbar[a];
    }

    // --- grid backward -------------------------------------------------------------------------
    std::vector<Cell<T, D>> gbar(pre.size(), Cell<T, D>{});
    const T allowed = sim.allowed_velocity();
    for (size_t i = 0; i < pre.size(); ++i) {
        for (int a = 0; a < D; ++a) gbar[i].v[a] = T(0);
        gbar[i].m = T(0);
        const bool sticky_zero = post[i].m == T(0) && pre[i].m > T(0);
        const bool zero_velocity = [&] {
            for (int a = 0; a < D; ++a)
                if (post[i].v[a] != T(0)) return false;
            return true;
        }();
        if (sticky_zero || (zero_velocity && post[i].m == T(0))) continue;  // zeroed by the sticky box
        if (pre[i].m > T(0)) {
            const T m = pre[i].m;
            T mbar = T(0);
            for (int a = 0; a < D; ++a) {
                T vt = pre[i].v[a] / m;
                if (a == 1) vt += dt * gravity;
                const bool clamped = vt < -allowed || vt > allowed;
                const T vtb = clamped ? T(0) : vbar[i].v[a];
                gbar[i].v[a] = vtb / m;
                mbar -= vtb * pre[i].v[a] / (m * m);
            }
            gbar[i].m = mbar;
        } else {
            for (int a = 0; a < D; ++a) gbar[i].v[a] = vbar[i].v[a];
        }
    }

    // --- P2G backward ------------------------------------------------------------------------
    T mu, lambda;
    for (size_t p = 0; p < n; ++p) {
        const auto& P = ps[p];
        sim.hardening(P, &mu, &lambda);
        int base[D];
        T fx[D], w[D][3], dw[D][3];
        sim.base_and_weights(P.x, base, fx, w);
        for (int a = 0; a < D; ++a) bspline_dweights<T>(fx[a], dw[a]);
        T A[D][D];
        sim.first_piola_kirchoff_stress(P, A);
        T Abar[D][D] = {};
        T fxbar[D] = {};
        sim.each_offset([&](const int off[D]) {
            const auto& gb = gbar[size_t(sim.node_index(base, off))];
            T weight = w[0][off[0]];
            for (int a = 1; a < D; ++a) weight *= w[a][off[a]];
            T dpos[D];
            for (int a = 0; a < D; ++a) dpos[a] = (T(off[a]) - fx[a]) * dx;
            T wbar = gb.m * P.mass;
            for (int r = 0; r < D; ++r) {
                T ad = T(0);
                for (int c = 0; c < D; ++c) ad += A[r][c] * dpos[c];
                wbar += gb.v[r] * (P.mass * P.v[r] + ad);
                in.v[p * D + r] += weight * P.mass * gb.v[r];
                for (int c = 0; c < D; ++c) Abar[r][c] += weight * gb.v[r] * dpos[c];
            }
            for (int c = 0; c < D; ++c) {
                T atm = T(0);
                for (int r = 0; r < D; ++r) atm += A[r][c] * gb.v[r];
                fxbar[c] -= dx * weight * atm;
            }
            for (int a = 0; a < D; ++a) {
                T dwa = dw[a][off[a]];
                for (int b = 0; b < D; ++b)
                    if (b != a) dwa *= w[b][off[b]];
                fxbar[a] += wbar * dwa;
            }
        });
        for (int a = 0; a < D; ++a) in.x[p * D + a] += inv_dx * fxbar[a];
        // A = -(dt V0) Dinv PF + m C
        const T Dinv = T(4) * inv_dx * inv_dx;
        T G[D][D];
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                in.C[(p * D + a) * D + b] += P.mass * Abar[a][b];
                G[a][b] = -(dt * P.volume) * Dinv * Abar[a][b];
            }
        // PF = 2 mu (F - R) F^T + lambda (J - 1) J * ones
        T R[D][D];
        nclr_polar<T, D>(P.F, R);
        const T J = det<T>(P.F);
        T Fbar[D][D] = {}, Rbar[D][D] = {};
        T sG = T(0), mubar = T(0);
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                sG += G[a][b];
                T gf = T(0), gtfr = T(0), fmrf = T(0);
                for (int k = 0; k < D; ++k) {
                    gf += G[a][k] * P.F[k][b];                      // (G F)_ab
                    gtfr += G[k][a] * (P.F[k][b] - R[k][b]);        // (G^T (F - R))_ab
                    fmrf += (P.F[a][k] - R[a][k]) * P.F[b][k];      // ((F - R) F^T)_ab
                }
                Fbar[a][b] += T(2) * mu * (gf + gtfr);
                Rbar[a][b] = -T(2) * mu * gf;
                mubar += G[a][b] * T(2) * fmrf;
            }
        polar_backward<T, D>(P.F, R, Rbar, Fbar);
        T cof[D][D];
        cofactor<T>(P.F, cof);
        const T Jbar = sG * lambda * (T(2) * J - T(1));
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) in.F[(p * D + a) * D + b] += Fbar[a][b] + Jbar * cof[a][b];
        const T lambar = sG * (J - T(1)) * J;
        const T e = material == kSnow ? T(std::exp(10.0 * (1.0 - double(P.Jp))))
                                      : (material == kJelly ? T(0.3f) : T(1.0f));
        // snow: mu = mu0 e, lambda = lambda0 e with de/dJp = -10 e
        if (material == kSnow) in.Jp[p] += (mubar * sim.mu_0 + lambar * sim.lambda_0) * (T(-10) * e);
        const T ep = e * (mubar / (T(2) * (T(1) + nu)) + lambar * nu / ((T(1) + nu) * (T(1) - T(2) * nu)));
        in.E += ep;
        in.E_abs += ep < T(0) ? -ep : ep;
    }
}

}  // namespace oracle
