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

NCLR_DI Rot make_jacobi(float x, float y, float z) {
    const float deno = 2.0f * fabsf(y);
    if (deno < kFltMin) return {1.0f, 0.0f};
    const float tau = (x - z) / deno;
    const float w = sqrtf(tau * tau + 1.0f);
    const float t = tau > 0.0f ? 1.0f / (tau + w) : 1.0f / (tau - w);
    const float sign_t = t > 0.0f ? 1.0f : -1.0f;
    const float n = 1.0f / sqrtf(t * t + 1.0f);
    const float s = -sign_t * (y / fabsf(y)) * fabsf(t) * n;
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
        const float tmp = sqrtf(1.0f + u * u);
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

NCLR_DI M<3> polar_rotation(const M<3>& A) {
    M<3> U, V;
    float sig[3];
    nclr_svd<3>(A, U, sig, V);
    return matmul_bt<3>(U, V);
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
