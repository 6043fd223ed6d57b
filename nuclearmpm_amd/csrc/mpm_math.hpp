// Device small-matrix math for the MLS-MPM kernels (gfx950, FP32 in registers).
//
// Every routine here is fully unrolled over compile-time indices so that the
// matrices stay in VGPRs (no scratch).  The translation unit is compiled with
// -ffp-contract=off and IEEE-rounded division / square root, so each
// expression rounds exactly as the reference's float code does on a default
// x86-64 build (src/nclr.h, src/nclr_math.h: -Ofast without -mfma emits no
// fused multiply-adds); operation order follows the reference's expressions
// left to right.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nclr_dev {

#define NCLR_DI __device__ __forceinline__

constexpr float kFltMin = 1.17549435e-38f;   // std::numeric_limits<float>::min()
constexpr float kFltEps = 1.1920928955e-7f;  // std::numeric_limits<float>::epsilon()
constexpr int kMaxJacobiSweeps = 32;

template <int D>
struct M {
    float a[D][D];
};

NCLR_DI float det(const M<2>& m) { return m.a[0][0] * m.a[1][1] - m.a[0][1] * m.a[1][0]; }

// Cofactor expansion along the first row.
NCLR_DI float det(const M<3>& m) {
    const auto& a = m.a;
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

template <int D>
NCLR_DI M<D> matmul(const M<D>& x, const M<D>& y) {
    M<D> o;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float s = x.a[i][0] * y.a[0][j];
#pragma unroll
            for (int k = 1; k < D; ++k) s = s + x.a[i][k] * y.a[k][j];
            o.a[i][j] = s;
        }
    return o;
}

// x * y^T
template <int D>
NCLR_DI M<D> matmul_bt(const M<D>& x, const M<D>& y) {
    M<D> o;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float s = x.a[i][0] * y.a[j][0];
#pragma unroll
            for (int k = 1; k < D; ++k) s = s + x.a[i][k] * y.a[j][k];
            o.a[i][j] = s;
        }
    return o;
}

// ---------------------------------------------------------------------------
// Two-sided cyclic Jacobi SVD (the algorithm nclr_svd gets from Eigen's
// JacobiSVD, src/nclr_math.h:57): per pair (p,q) a symmetrising rotation and
// a symmetric 2x2 Jacobi rotation; U accumulates the transposed left
// rotations, V the right ones; the diagonal is made non-negative, unscaled
// and selection-sorted descending.
// ---------------------------------------------------------------------------
struct Rot {
    float c, s;
};

// Correctly rounded sqrt(x) for x >= 1 (or +inf): the hardware square root (within 1 ulp) moved to
// a neighbour when the residual x - s' s at that neighbour s' says so.  This is the rounding step
// hipcc emits for an IEEE sqrtf; its small-argument scaling and its special-value fixup do
// nothing for these arguments and are left out.
NCLR_DI float sqrt_ge1(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __int_as_float(__float_as_int(s) - 1);
    const float up = __int_as_float(__float_as_int(s) + 1);
    const float r = fmaf(-dn, s, x) <= 0.0f ? dn : s;
    return fmaf(-up, s, x) > 0.0f ? up : r;
}

NCLR_DI Rot make_jacobi(float x, float y, float z) {
    const float deno = 2.0f * fabsf(y);
    if (deno < kFltMin) return {1.0f, 0.0f};
    const float tau = (x - z) / deno;
    const float w = sqrt_ge1(tau * tau + 1.0f);
    const float t = 1.0f / (tau > 0.0f ? tau + w : tau - w);
    const float sign_t = t > 0.0f ? 1.0f : -1.0f;
    const float n = 1.0f / sqrt_ge1(t * t + 1.0f);
    // y / |y| is exactly +-1 for the finite, nonzero y that reaches here
    const float s = -sign_t * copysignf(1.0f, y) * fabsf(t) * n;
    return {n, s};
}

template <int D, int P, int Q>
NCLR_DI void jacobi_pair(M<D>& W, M<D>& U, M<D>& V, float& max_diag, bool& finished) {
    const float threshold = fmaxf(kFltMin, (2.0f * kFltEps) * max_diag);
    if (!(fabsf(W.a[P][Q]) > threshold || fabsf(W.a[Q][P]) > threshold)) return;
    finished = false;
    const float m00 = W.a[P][P], m01 = W.a[P][Q], m10 = W.a[Q][P], m11 = W.a[Q][Q];
    Rot rot1;
    const float t = m00 + m11;
    const float d = m10 - m01;
    if (fabsf(d) < kFltMin) {
        rot1 = {1.0f, 0.0f};
    } else {
        const float u = t / d;
        const float tmp = sqrt_ge1(1.0f + u * u);
        rot1 = {u / tmp, 1.0f / tmp};
    }
    const float a00 = rot1.c * m00 + rot1.s * m10;
    const float a01 = rot1.c * m01 + rot1.s * m11;
    const float a11 = -rot1.s * m01 + rot1.c * m11;
    const Rot jr = make_jacobi(a00, a01, a11);
    const float rc = jr.c, rs = -jr.s;
    const Rot jl = {rot1.c * rc - rot1.s * rs, rot1.c * rs + rot1.s * rc};
    // rows P,Q of W <- jl * rows
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float x = W.a[P][k], y = W.a[Q][k];
        W.a[P][k] = jl.c * x + jl.s * y;
        W.a[Q][k] = -jl.s * x + jl.c * y;
    }
    // cols P,Q of U <- cols * jl^T
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float x = U.a[k][P], y = U.a[k][Q];
        U.a[k][P] = jl.c * x + jl.s * y;
        U.a[k][Q] = -jl.s * x + jl.c * y;
    }
    // cols P,Q of W and V <- cols * jr
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float x = W.a[k][P], y = W.a[k][Q];
        W.a[k][P] = jr.c * x - jr.s * y;
        W.a[k][Q] = jr.s * x + jr.c * y;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float x = V.a[k][P], y = V.a[k][Q];
        V.a[k][P] = jr.c * x - jr.s * y;
        V.a[k][Q] = jr.s * x + jr.c * y;
    }
    max_diag = fmaxf(max_diag, fmaxf(fabsf(W.a[P][P]), fabsf(W.a[Q][Q])));
}

template <int D, int I, int J>
NCLR_DI void swap_cols(float (&sig)[D], M<D>& U, M<D>& V) {
    const float t = sig[I];
    sig[I] = sig[J];
    sig[J] = t;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float u = U.a[k][I];
        U.a[k][I] = U.a[k][J];
        U.a[k][J] = u;
        const float v = V.a[k][I];
        V.a[k][I] = V.a[k][J];
        V.a[k][J] = v;
    }
}

// nclr_svd (src/nclr_math.h:50-74).  In 3D the column-2 flips make U, V proper
// rotations with the sign of det(A) on the smallest singular value.  In 2D the
// reference's col(2) access is out of bounds; it is treated as a no-op.
template <int D>
NCLR_DI void nclr_svd(const M<D>& A, M<D>& U, float (&sig)[D], M<D>& V) {
    float scale = 0.0f;
    bool finite = true;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            finite = finite && isfinite(A.a[i][j]);
            scale = fmaxf(scale, fabsf(A.a[i][j]));
        }
    if (!finite) {
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int i = 0; i < D; ++i) {
            sig[i] = nan;
#pragma unroll
            for (int j = 0; j < D; ++j) U.a[i][j] = V.a[i][j] = nan;
        }
        return;
    }
    if (scale == 0.0f) scale = 1.0f;
    M<D> W;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            W.a[i][j] = A.a[i][j] / scale;
            U.a[i][j] = V.a[i][j] = (i == j) ? 1.0f : 0.0f;
        }
    float max_diag = 0.0f;
#pragma unroll
    for (int i = 0; i < D; ++i) max_diag = fmaxf(max_diag, fabsf(W.a[i][i]));
    bool finished = false;
    for (int sweep = 0; !finished && sweep < kMaxJacobiSweeps; ++sweep) {
        finished = true;
        jacobi_pair<D, 1, 0>(W, U, V, max_diag, finished);
        if constexpr (D == 3) {
            jacobi_pair<D, 2, 0>(W, U, V, max_diag, finished);
            jacobi_pair<D, 2, 1>(W, U, V, max_diag, finished);
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const float a = W.a[i][i];
        sig[i] = fabsf(a);
        if (a < 0.0f) {
#pragma unroll
            for (int k = 0; k < D; ++k) U.a[k][i] = -U.a[k][i];
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) sig[i] = sig[i] * scale;
    // selection sort, descending, first maximum wins; stops at an all-zero tail
    if constexpr (D == 2) {
        if (sig[0] != 0.0f || sig[1] != 0.0f) {
            if (sig[1] > sig[0]) swap_cols<2, 0, 1>(sig, U, V);
        }
    } else {
        int pos = 0;
        if (sig[1] > sig[pos]) pos = 1;
        if (sig[2] > (pos == 0 ? sig[0] : sig[1])) pos = 2;
        const float mx = pos == 0 ? sig[0] : (pos == 1 ? sig[1] : sig[2]);
        if (mx != 0.0f) {
            if (pos == 1) swap_cols<3, 0, 1>(sig, U, V);
            else if (pos == 2) swap_cols<3, 0, 2>(sig, U, V);
            const bool second = sig[2] > sig[1];
            const float mx2 = second ? sig[2] : sig[1];
            if (mx2 != 0.0f && second) swap_cols<3, 1, 2>(sig, U, V);
        }
        if (det(U) < 0.0f) {
#pragma unroll
            for (int k = 0; k < 3; ++k) U.a[k][2] = -U.a[k][2];
            sig[2] = -sig[2];
        }
        if (det(V) < 0.0f) {
#pragma unroll
            for (int k = 0; k < 3; ++k) V.a[k][2] = -V.a[k][2];
            sig[2] = -sig[2];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Fast 3x3 SVD with nclr_svd's convention (U, V in SO(3), singular values in descending order of
// magnitude, the sign of det A on the last one), for the default (non-bitwise) mode.
//
// Cyclic Jacobi eigen-decomposition of the symmetric B = A^T A (Golub & Van Loan, Matrix
// Computations, 4th ed., Algorithm 8.5.2 with the 2x2 symmetric Schur step of Algorithm 8.5.1),
// a fixed four sweeps (quadratic convergence reaches FP32 rounding in three; the fixed count keeps
// every lane of a wave on the same path), then U from A V by Gram-Schmidt with u3 = u1 x u2, and
// sigma_i = u_i^T A v_i.  Singular vectors of small singular values come from the eigenvectors of
// B (well separated from the large ones) rather than from A v / sigma, so rank-deficient inputs
// (the reference's default F = diag(1, 1, 0) in 3D) keep orthonormal factors.
// Rotation, reciprocal and square-root steps use the hardware approximations (1 ulp): the
// factors are orthonormal to FP32 rounding by construction (c = 1/sqrt(1 + t^2), s = t c).
// ---------------------------------------------------------------------------------------------
NCLR_DI float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
NCLR_DI float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
NCLR_DI float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int P, int Q>
NCLR_DI void sym_jacobi(float (&B)[3][3], M<3>& V) {
    constexpr int R = 3 - P - Q;
    const float bpq = B[P][Q];
    const float bpp = B[P][P], bqq = B[Q][Q];
    // tan of the angle that zeroes B[P][Q]; |tau| -> inf gives t -> 0
    const float tau = (bqq - bpp) * fast_rcp(2.0f * bpq);
    float t = copysignf(1.0f, tau) * fast_rcp(fabsf(tau) + fast_sqrt(fmaf(tau, tau, 1.0f)));
    t = (bpq != 0.0f && t == t) ? t : 0.0f;
    const float c = fast_rsq(fmaf(t, t, 1.0f));
    const float s = t * c;
    B[P][P] = fmaf(-t, bpq, bpp);
    B[Q][Q] = fmaf(t, bpq, bqq);
    B[P][Q] = B[Q][P] = 0.0f;
    const float brp = B[R][P], brq = B[R][Q];
    B[R][P] = B[P][R] = fmaf(c, brp, -s * brq);
    B[R][Q] = B[Q][R] = fmaf(s, brp, c * brq);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float vp = V.a[k][P], vq = V.a[k][Q];
        V.a[k][P] = fmaf(c, vp, -s * vq);
        V.a[k][Q] = fmaf(s, vp, c * vq);
    }
}

// Orders eigenpairs I < J by decreasing eigenvalue; a column swap is paired with a sign change so
// that det V stays +1.
template <int I, int J>
NCLR_DI void sort_pair(float (&l)[3], M<3>& V) {
    const bool sw = l[I] < l[J];
    const float li = l[I], lj = l[J];
    l[I] = sw ? lj : li;
    l[J] = sw ? li : lj;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float vi = V.a[k][I], vj = V.a[k][J];
        V.a[k][I] = sw ? vj : vi;
        V.a[k][J] = sw ? -vi : vj;
    }
}

NCLR_DI void fast_svd3(const M<3>& A, M<3>& U, float (&sig)[3], M<3>& V) {
    float B[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            float s = A.a[0][i] * A.a[0][j];
            s = fmaf(A.a[1][i], A.a[1][j], s);
            s = fmaf(A.a[2][i], A.a[2][j], s);
            B[i][j] = s;
            B[j][i] = s;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) V.a[i][j] = i == j ? 1.0f : 0.0f;
#pragma unroll
    for (int sweep = 0; sweep < 4; ++sweep) {
        sym_jacobi<0, 1>(B, V);
        sym_jacobi<0, 2>(B, V);
        sym_jacobi<1, 2>(B, V);
    }
    float l[3] = {B[0][0], B[1][1], B[2][2]};
    sort_pair<0, 1>(l, V);
    sort_pair<0, 2>(l, V);
    sort_pair<1, 2>(l, V);
    // A v_i
    float av[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float s = A.a[r][0] * V.a[0][i];
            s = fmaf(A.a[r][1], V.a[1][i], s);
            s = fmaf(A.a[r][2], V.a[2][i], s);
            av[i][r] = s;
        }
    float u1[3], u2[3], u3[3];
    {
        const float n2 = fmaf(av[0][0], av[0][0], fmaf(av[0][1], av[0][1], av[0][2] * av[0][2]));
        const bool ok = n2 > 0.0f;
        const float inv = ok ? fast_rsq(n2) : 0.0f;
        u1[0] = ok ? av[0][0] * inv : 1.0f;
        u1[1] = ok ? av[0][1] * inv : 0.0f;
        u1[2] = ok ? av[0][2] * inv : 0.0f;
    }
    {
        const float d = fmaf(u1[0], av[1][0], fmaf(u1[1], av[1][1], u1[2] * av[1][2]));
        float w[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) w[r] = fmaf(-d, u1[r], av[1][r]);
        const float n2 = fmaf(w[0], w[0], fmaf(w[1], w[1], w[2] * w[2]));
        if (n2 > 0.0f) {
            const float inv = fast_rsq(n2);
#pragma unroll
            for (int r = 0; r < 3; ++r) u2[r] = w[r] * inv;
        } else {  // rank <= 1: any unit vector orthogonal to u1
            const bool ex = fabsf(u1[0]) < 0.9f;
            const float e[3] = {ex ? 1.0f : 0.0f, ex ? 0.0f : 1.0f, 0.0f};
            const float de = fmaf(u1[0], e[0], u1[1] * e[1]);
#pragma unroll
            for (int r = 0; r < 3; ++r) w[r] = fmaf(-de, u1[r], e[r]);
            const float inv = fast_rsq(fmaf(w[0], w[0], fmaf(w[1], w[1], w[2] * w[2])));
#pragma unroll
            for (int r = 0; r < 3; ++r) u2[r] = w[r] * inv;
        }
    }
    u3[0] = fmaf(u1[1], u2[2], -u1[2] * u2[1]);
    u3[1] = fmaf(u1[2], u2[0], -u1[0] * u2[2]);
    u3[2] = fmaf(u1[0], u2[1], -u1[1] * u2[0]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        U.a[r][0] = u1[r];
        U.a[r][1] = u2[r];
        U.a[r][2] = u3[r];
    }
    sig[0] = fmaf(u1[0], av[0][0], fmaf(u1[1], av[0][1], u1[2] * av[0][2]));
    sig[1] = fmaf(u2[0], av[1][0], fmaf(u2[1], av[1][1], u2[2] * av[1][2]));
    sig[2] = fmaf(u3[0], av[2][0], fmaf(u3[1], av[2][1], u3[2] * av[2][2]));
}

// ---------------------------------------------------------------------------------------------
// Accurate mode: the 3x3 polar factor and SVD in double precision, more accurate than the FP32
// two-sided Jacobi (whose R carries the FP32 rounding of its rotations, amplified by the stress's
// F - R at small strain).  FP64 FMA, add and multiply issue at the FP32 rate on gfx950; the
// reciprocal, square root and reciprocal square root estimates take one Newton step each.
// ---------------------------------------------------------------------------------------------
NCLR_DI double rcp_f64(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    const double r1 = fma(r, e, r);
    return fma(r1, fma(-x, r1, 1.0), r1);
}
NCLR_DI double rsq_f64(double x) {  // x > 0
    const double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x * y;
    const double y1 = fma(y, fma(-h, y, 0.5), y);  // y (1.5 - x y^2 / 2)
    return fma(y1, fma(-0.5 * x * y1, y1, 0.5), y1);
}

NCLR_DI void cofactors(const double (&X)[3][3], double (&c)[3][3]) {
    c[0][0] = fma(X[1][1], X[2][2], -X[1][2] * X[2][1]);
    c[0][1] = fma(X[1][2], X[2][0], -X[1][0] * X[2][2]);
    c[0][2] = fma(X[1][0], X[2][1], -X[1][1] * X[2][0]);
    c[1][0] = fma(X[0][2], X[2][1], -X[0][1] * X[2][2]);
    c[1][1] = fma(X[0][0], X[2][2], -X[0][2] * X[2][0]);
    c[1][2] = fma(X[0][1], X[2][0], -X[0][0] * X[2][1]);
    c[2][0] = fma(X[0][1], X[1][2], -X[0][2] * X[1][1]);
    c[2][1] = fma(X[0][2], X[1][0], -X[0][0] * X[1][2]);
    c[2][2] = fma(X[0][0], X[1][1], -X[0][1] * X[1][0]);
}

// F - R and R, R the polar rotation of F (the reference's nclr_polar,
// src/nclr_math.h:76-98, for det F > 0), by Newton's iteration X <- (X + X^-T) / 2 =
// (X + cof(X) / det X) / 2 from X = F in double, for F with ||F||_F^3 <= 40 det F (condition number
// below ~8; false otherwise, and the caller takes the Jacobi SVD).  The iteration converges
// quadratically to R; once a step changes X by at most 1e-6 one more step leaves an error below
// 1e-12, and F - R is rounded once to float, so it keeps its relative precision at small strain.
NCLR_DI bool polar_newton(const M<3>& F, M<3>& fmr, M<3>& rot) {
    double X[3][3], c[3][3];
    double nf2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            X[i][j] = double(F.a[i][j]);
            nf2 = fma(X[i][j], X[i][j], nf2);
        }
    cofactors(X, c);
    double det = fma(X[0][0], c[0][0], fma(X[0][1], c[0][1], X[0][2] * c[0][2]));
    if (!(det > 0.0) || !(nf2 * nf2 * nf2 <= 1600.0 * det * det)) return false;
    bool conv = false;
    for (int it = 0; it < 10; ++it) {
        const double half_inv = 0.5 * rcp_f64(det);
        double mx = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double xn = fma(c[i][j], half_inv, 0.5 * X[i][j]);
                mx = fmax(mx, fabs(xn - X[i][j]));
                X[i][j] = xn;
            }
        cofactors(X, c);
        det = fma(X[0][0], c[0][0], fma(X[0][1], c[0][1], X[0][2] * c[0][2]));
        if (mx <= 1e-6) {
            conv = true;
            break;
        }
    }
    if (!conv) return false;
    const double half_inv = 0.5 * rcp_f64(det);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double r = fma(c[i][j], half_inv, 0.5 * X[i][j]);
            fmr.a[i][j] = float(double(F.a[i][j]) - r);
            rot.a[i][j] = float(r);
        }
    return true;
}
template <bool ROTATION = false>
NCLR_DI bool f_minus_polar_newton(const M<3>& F, M<3>& out) {
    M<3> unused;
    return ROTATION ? polar_newton(F, unused, out) : polar_newton(F, out, unused);
}

// One rotation of the cyclic Jacobi eigen-decomposition of the symmetric B (Golub & Van Loan,
// Matrix Computations, Algorithm 8.5.2 with the symmetric Schur step 8.5.1) in double,
// accumulated into the columns P, Q of W.
template <int P, int Q>
NCLR_DI void sym_jacobi_f64(double (&B)[3][3], double (&W)[3][3]) {
    constexpr int R = 3 - P - Q;
    const double bpq = B[P][Q];
    if (bpq == 0.0) return;
    const double bpp = B[P][P], bqq = B[Q][Q];
    const double tau = (bqq - bpp) * rcp_f64(2.0 * bpq);
    const double root = fma(tau, tau, 1.0);
    // |tau| above 1e150: t = 1 / (2 tau) to double precision, and tau^2 would overflow
    const double t = fabs(tau) > 1e150 ? 0.5 * rcp_f64(tau) : copysign(1.0, tau) * rcp_f64(fabs(tau) + root * rsq_f64(root));
    const double c = rsq_f64(fma(t, t, 1.0));
    const double s = t * c;
    B[P][P] = fma(-t, bpq, bpp);
    B[Q][Q] = fma(t, bpq, bqq);
    B[P][Q] = B[Q][P] = 0.0;
    const double brp = B[R][P], brq = B[R][Q];
    B[R][P] = B[P][R] = fma(c, brp, -s * brq);
    B[R][Q] = B[Q][R] = fma(s, brp, c * brq);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double wp = W[k][P], wq = W[k][Q];
        W[k][P] = fma(c, wp, -s * wq);
        W[k][Q] = fma(s, wp, c * wq);
    }
}

// Orders eigenpairs I < J by decreasing eigenvalue; a swap negates one column (det W stays +1).
template <int I, int J>
NCLR_DI void sort_pair_f64(double (&l)[3], double (&W)[3][3]) {
    const bool sw = l[I] < l[J];
    const double li = l[I], lj = l[J];
    l[I] = sw ? lj : li;
    l[J] = sw ? li : lj;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double wi = W[k][I], wj = W[k][J];
        W[k][I] = sw ? wj : wi;
        W[k][J] = sw ? -wi : wj;
    }
}

// 3x3 SVD with nclr_svd's convention (U, V proper rotations, singular values in descending order
// of magnitude, the sign of det A on the last one), in double: Jacobi eigen-decomposition of
// A^T A until the off-diagonal part is below 1e-15 of the diagonal (at most 8 sweeps), then
// u_1 = A v_1 / |A v_1|, u_2 from A v_2 by Gram-Schmidt, u_3 = u_1 x u_2 and sigma_i = u_i^T A v_i,
// rounded to float.  The singular vectors of a small singular value come from the eigenvectors of
// A^T A and the cross product, so rank-deficient inputs (the reference's F = diag(1, 1, 0) in 3D)
// keep orthonormal factors.  Non-finite input gives NaN factors, like nclr_svd.
NCLR_DI void accurate_svd3(const M<3>& A, M<3>& U, float (&sig)[3], M<3>& V) {
    double a[3][3], B[3][3], W[3][3];
    bool finite = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            a[i][j] = double(A.a[i][j]);
            finite = finite && isfinite(A.a[i][j]);
            W[i][j] = i == j ? 1.0 : 0.0;
        }
    if (!finite) {
        const float nan = __builtin_nanf("");
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            sig[i] = nan;
#pragma unroll
            for (int j = 0; j < 3; ++j) U.a[i][j] = V.a[i][j] = nan;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const double v = fma(a[0][i], a[0][j], fma(a[1][i], a[1][j], a[2][i] * a[2][j]));
            B[i][j] = v;
            B[j][i] = v;
        }
    for (int sweep = 0; sweep < 8; ++sweep) {
        const double off = fma(B[0][1], B[0][1], fma(B[0][2], B[0][2], B[1][2] * B[1][2]));
        const double diag = fma(B[0][0], B[0][0], fma(B[1][1], B[1][1], B[2][2] * B[2][2]));
        if (off <= 1e-30 * diag) break;
        sym_jacobi_f64<0, 1>(B, W);
        sym_jacobi_f64<0, 2>(B, W);
        sym_jacobi_f64<1, 2>(B, W);
    }
    double l[3] = {B[0][0], B[1][1], B[2][2]};
    sort_pair_f64<0, 1>(l, W);
    sort_pair_f64<0, 2>(l, W);
    sort_pair_f64<1, 2>(l, W);
    double av[3][3];  // av[i] = A v_i
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int r = 0; r < 3; ++r) av[i][r] = fma(a[r][0], W[0][i], fma(a[r][1], W[1][i], a[r][2] * W[2][i]));
    double u1[3], u2[3], u3[3];
    {
        const double n2 = fma(av[0][0], av[0][0], fma(av[0][1], av[0][1], av[0][2] * av[0][2]));
        const bool ok = n2 > 0.0;
        const double inv = ok ? rsq_f64(n2) : 0.0;
        u1[0] = ok ? av[0][0] * inv : 1.0;
        u1[1] = ok ? av[0][1] * inv : 0.0;
        u1[2] = ok ? av[0][2] * inv : 0.0;
    }
    {
        const double d = fma(u1[0], av[1][0], fma(u1[1], av[1][1], u1[2] * av[1][2]));
        double w[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) w[r] = fma(-d, u1[r], av[1][r]);
        double n2 = fma(w[0], w[0], fma(w[1], w[1], w[2] * w[2]));
        if (!(n2 > 1e-60 * fma(av[1][0], av[1][0], fma(av[1][1], av[1][1], fma(av[1][2], av[1][2], 1e-300))))) {
            // rank <= 1: any unit vector orthogonal to u1
            const bool ex = fabs(u1[0]) < 0.9;
            const double e[3] = {ex ? 1.0 : 0.0, ex ? 0.0 : 1.0, 0.0};
            const double de = fma(u1[0], e[0], u1[1] * e[1]);
#pragma unroll
            for (int r = 0; r < 3; ++r) w[r] = fma(-de, u1[r], e[r]);
            n2 = fma(w[0], w[0], fma(w[1], w[1], w[2] * w[2]));
        }
        const double inv = rsq_f64(n2);
#pragma unroll
        for (int r = 0; r < 3; ++r) u2[r] = w[r] * inv;
    }
    u3[0] = fma(u1[1], u2[2], -u1[2] * u2[1]);
    u3[1] = fma(u1[2], u2[0], -u1[0] * u2[2]);
    u3[2] = fma(u1[0], u2[1], -u1[1] * u2[0]);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        U.a[r][0] = float(u1[r]);
        U.a[r][1] = float(u2[r]);
        U.a[r][2] = float(u3[r]);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) V.a[r][cc] = float(W[r][cc]);
    }
    sig[0] = float(fma(u1[0], av[0][0], fma(u1[1], av[0][1], u1[2] * av[0][2])));
    sig[1] = float(fma(u2[0], av[1][0], fma(u2[1], av[1][1], u2[2] * av[1][2])));
    sig[2] = float(fma(u3[0], av[2][0], fma(u3[1], av[2][1], u3[2] * av[2][2])));
}

// SVD modes of the step: 0 nclr_svd's operation order (bitwise the oracle's FP32 arithmetic),
// 1 the fast FP32 path, 2 the accurate FP64 path (2x2 always takes nclr_svd, one Jacobi pair).
constexpr int kSvdExact = 0;
constexpr int kSvdFast = 1;
constexpr int kSvdAccurate = 2;

template <int MODE, int D>
NCLR_DI void step_svd(const M<D>& A, M<D>& U, float (&sig)[D], M<D>& V) {
    if constexpr (D == 3 && MODE == kSvdFast) fast_svd3(A, U, sig, V);
    else if constexpr (D == 3 && MODE == kSvdAccurate) accurate_svd3(A, U, sig, V);
    else nclr_svd<D>(A, U, sig, V);
}

// Rotation factor of nclr_polar (src/nclr_math.h:76-98).
NCLR_DI M<2> polar_rotation(const M<2>& A) {
    const float x = A.a[0][0] + A.a[1][1];
    const float y = A.a[1][0] - A.a[0][1];
    const float scale = 1.0f / sqrtf(x * x + y * y);
    const float c = x * scale, s = y * scale;
    M<2> R;
    R.a[0][0] = c;
    R.a[0][1] = -s;
    R.a[1][0] = s;
    R.a[1][1] = c;
    return R;
}

template <int MODE = kSvdExact>
NCLR_DI M<3> polar_rotation(const M<3>& A) {
    M<3> U, V;
    float sig[3];
    step_svd<MODE, 3>(A, U, sig, V);
    return matmul_bt<3>(U, V);
}

template <int MODE>
NCLR_DI M<2> polar_rotation_mode(const M<2>& A) {
    return polar_rotation(A);
}
template <int MODE>
NCLR_DI M<3> polar_rotation_mode(const M<3>& A) {
    return polar_rotation<MODE>(A);
}

// F - R with R = nclr_polar's rotation factor.  The accurate mode takes Newton's iteration where F
// is well-conditioned with det F > 0, and the two-sided Jacobi SVD elsewhere (in the accurate
// mode the Jacobi result's F - R is that of nclr_svd, so the fallback keeps the reference's
// rotation convention for det F <= 0).
template <int MODE, int D>
NCLR_DI M<D> f_minus_rotation(const M<D>& F) {
    M<D> out;
    if constexpr (D == 3 && MODE == kSvdAccurate) {
        if (f_minus_polar_newton(F, out)) return out;
    }
    const M<D> R = polar_rotation_mode<MODE == kSvdAccurate ? kSvdExact : MODE>(F);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) out.a[a][b] = F.a[a][b] - R.a[a][b];
    return out;
}

// R with nclr_polar's convention: in the accurate mode (3D) Newton's iteration where it applies
// and the two-sided Jacobi elsewhere, as f_minus_rotation; otherwise the mode's SVD.
template <int MODE, int D>
NCLR_DI M<D> rotation_mode(const M<D>& F) {
    if constexpr (D == 3 && MODE == kSvdAccurate) {
        M<D> R;
        if (f_minus_polar_newton<true>(F, R)) return R;
    }
    return polar_rotation_mode<MODE == kSvdAccurate ? kSvdExact : MODE>(F);
}

// Quadratic B-spline weights (src/nclr.h:124-127).
NCLR_DI void bspline_weights(float fx, float (&w)[3]) {
    const float a = 1.5f - fx;
    const float b = fx - 1.0f;
    const float c = fx - 0.5f;
    w[0] = 0.5f * (a * a);
    w[1] = 0.75f - b * b;
    w[2] = 0.5f * (c * c);
}

}  // namespace nclr_dev
