// Per-particle pieces of the step shared by the tiled kernels (mpm_step.hip) and the compact
// step (mpm_compact.hip): P2G's quantised deposit into LDS accumulators and G2P of one particle.
#pragma once

#include <hip/hip_runtime.h>

#include "mpm_device.hpp"

namespace nclr_dev {

// One particle's 3^D contributions weight * (m v + A dpos) (momentum channels 0..D-1) and
// weight * m (channel 3), quantised into the LDS accumulators of its region (offset rb).  The
// products A[c][b] dpos_b of the outer axes are formed once per outer offset and those of the last
// axis once per particle, so every value keeps the operation order of
// weight * (m v_c + ((A_c0 dpos_0 + A_c1 dpos_1) + A_c2 dpos_2)) with weight = (w_0 w_1) w_2.
// CHECKED: contributions at or beyond the fixed-point range flag their node channel instead.
// The accumulators of node (o0, o1[, o2]) of the stencil and channel c are acc[c * rs.chan + rb +
// o0 * rs.s0 + o1 * rs.s1 (+ o2)] (2D: rs.s1 = 1).
struct RegionStrides {
    int s0, s1, chan;
};
template <int D>
NCLR_DI constexpr RegionStrides tile_region() {
    using TL = Tile<D>;
    return D == 3 ? RegionStrides{TL::R * TL::R, TL::R, TL::REGION} : RegionStrides{TL::R, 1, TL::REGION};
}

template <int D, bool CHECKED>
NCLR_DI void p2g_deposit(unsigned long long* __restrict__ acc, uint32_t* __restrict__ nonfinite, int rb,
                         const float (&fx)[D], const M<D>& A, const float (&mv)[D], float m, float dx,
                         const float (&scale)[4], const RegionStrides rs) {
    // The channel scales are powers of two, and scaling by 2^e commutes with every rounding: the
    // per-particle factors are scaled once, so each node's weight * (m v_c + A_c dpos) * 2^e comes
    // out bit for bit without a multiply per node and channel.
    M<D> As;
    float mvs[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        mvs[c] = mv[c] * scale[c];
#pragma unroll
        for (int b = 0; b < D; ++b) As.a[c][b] = A.a[c][b] * scale[c];
    }
    const float ms = m * scale[3];
    float w[D][3], dp[D][3], last[D][3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        bspline_weights(fx[a], w[a]);
#pragma unroll
        for (int o = 0; o < 3; ++o) dp[a][o] = (float(o) - fx[a]) * dx;
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int o = 0; o < 3; ++o) last[c][o] = As.a[c][D - 1] * dp[D - 1][o];
    auto node = [&](int ro, float weight, const float (&outer)[D], int ol) {
        float val[4];
#pragma unroll
        for (int c = 0; c < D; ++c) val[c] = weight * (mvs[c] + (outer[c] + last[c][ol]));
        val[3] = weight * ms;
        uint32_t bad = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (D == 2 && c == 2) continue;
            const float y = val[c];
            if (!CHECKED || fabsf(y) < kFixedLimit) atomicAdd(&acc[c * rs.chan + rb + ro], fixed_point(y));
            else bad |= 1u << c;
        }
        if (CHECKED && bad) atomicOr(&nonfinite[rb + ro], bad);
    };
    if constexpr (D == 3) {
#pragma unroll
        for (int o0 = 0; o0 < 3; ++o0)
#pragma unroll
            for (int o1 = 0; o1 < 3; ++o1) {
                float outer[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) outer[c] = As.a[c][0] * dp[0][o0] + As.a[c][1] * dp[1][o1];
                const float w01 = w[0][o0] * w[1][o1];
#pragma unroll
                for (int o2 = 0; o2 < 3; ++o2) node(o0 * rs.s0 + o1 * rs.s1 + o2, w01 * w[2][o2], outer, o2);
            }
    } else {
#pragma unroll
        for (int o0 = 0; o0 < 3; ++o0) {
            float outer[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) outer[c] = As.a[c][0] * dp[0][o0];
#pragma unroll
            for (int o1 = 0; o1 < 3; ++o1) node(o0 * rs.s0 + o1 * rs.s1, w[0][o0] * w[1][o1], outer, o1);
        }
    }
}
// The accurate mode's deposit (3D): each contribution weight * (m v_c + A_c dpos) and weight * m is
// formed in double from the particle's factors and quantised by one fused multiply-add with
// 1.5 * 2^52 (fixed_point's rounding, without rounding the product to float first): three
// instructions per contribution instead of six, and no float rounding of the contribution.
template <bool CHECKED>
NCLR_DI void p2g_deposit_fused(unsigned long long* __restrict__ acc, uint32_t* __restrict__ nonfinite, int rb,
                               const float (&fx)[3], const M<3>& A, const float (&mv)[3], float m, float dx,
                               const float (&scale)[4], const RegionStrides rs) {
    constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
    double As[3][3], mvs[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        mvs[c] = double(mv[c]) * double(scale[c]);
#pragma unroll
        for (int b = 0; b < 3; ++b) As[c][b] = double(A.a[c][b]) * double(scale[c]);
    }
    const double ms = double(m) * double(scale[3]);
    double w[3][3], dp[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double f = double(fx[a]);
        const double p = 1.5 - f, q = f - 1.0, r = f - 0.5;
        w[a][0] = 0.5 * p * p;
        w[a][1] = fma(-q, q, 0.75);
        w[a][2] = 0.5 * r * r;
#pragma unroll
        for (int o = 0; o < 3; ++o) dp[a][o] = (double(o) - f) * double(dx);
    }
    double last[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int o = 0; o < 3; ++o) last[c][o] = As[c][2] * dp[2][o];
    auto quant = [&](double weight, double x, int slot) {
        const double t = fma(weight, x, kMagic);
        if (CHECKED && !(fabs(t - kMagic) < double(kFixedLimit))) return false;
        const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(t));
        const uint32_t lo = uint32_t(b), hi = uint32_t(b >> 32) - 0x43380000u;
        atomicAdd(&acc[slot], (static_cast<unsigned long long>(hi) << 32) | lo);
        return true;
    };
#pragma unroll
    for (int o0 = 0; o0 < 3; ++o0)
#pragma unroll
        for (int o1 = 0; o1 < 3; ++o1) {
            double outer[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) outer[c] = fma(As[c][1], dp[1][o1], fma(As[c][0], dp[0][o0], mvs[c]));
            const double w01 = w[0][o0] * w[1][o1];
#pragma unroll
            for (int o2 = 0; o2 < 3; ++o2) {
                const int node = rb + o0 * rs.s0 + o1 * rs.s1 + o2;
                const double weight = w01 * w[2][o2];
                uint32_t bad = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (!quant(weight, outer[c] + last[c][o2], c * rs.chan + node)) bad |= 1u << c;
                if (!quant(weight, ms, 3 * rs.chan + node)) bad |= 8u;
                if (CHECKED && bad) atomicOr(&nonfinite[node], bad);
            }
        }
}

// G2P of one particle (src/nclr.h:167-261) from its stencil (fx, w) and a gather functor node(off)
// returning the updated grid's float4 {v, m} at stencil node base + off: v and C from the 3^D
// nodes, x advected, F updated with I' (diag<dim>(1), Q1), then jelly / snow (SVD clamp, Jp) /
// liquid (J in double).  The exact mode follows the reference's operation order.
template <int D, int MODE, typename Gather>
NCLR_DI void g2p_particle(int material, const SceneParams& sp, const float (&fx)[D], const float (&w)[D][3],
                          const Gather& node, float (&x)[D], float (&v)[D], M<D>& C, M<D>& F, float& Jp) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
        v[a] = 0.0f;
#pragma unroll
        for (int b = 0; b < D; ++b) C.a[a][b] = 0.0f;
    }
    const float four_inv_dx = 4.0f * sp.inv_dx;
    constexpr int K = D == 2 ? 9 : 27;
    if constexpr (D == 3 && MODE == kSvdAccurate) {
        // v = sum w g and C = 4 inv_dx sum w g (off - fx)^T = 4 inv_dx (T - v fx^T) with
        // T = sum w g off^T, off in {0, 1, 2}^3: fused multiply-adds, and the (off - fx) products
        // taken once per particle instead of per node
        float T[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) T[a][b] = 0.0f;
#pragma unroll
        for (int o0 = 0; o0 < 3; ++o0)
#pragma unroll
            for (int o1 = 0; o1 < 3; ++o1) {
                const float w01 = w[0][o0] * w[1][o1];
#pragma unroll
                for (int o2 = 0; o2 < 3; ++o2) {
                    const int off[3] = {o0, o1, o2};
                    const float4 g = node(off);
                    const float weight = w01 * w[2][o2];
                    const float wg[3] = {weight * g.x, weight * g.y, weight * g.z};
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        v[a] += wg[a];
#pragma unroll
                        for (int b = 0; b < 3; ++b)
                            if (off[b] != 0) T[a][b] = fmaf(float(off[b]), wg[a], T[a][b]);
                    }
                }
            }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) C.a[a][b] = four_inv_dx * fmaf(-v[a], fx[b], T[a][b]);
    } else {
#pragma unroll
        for (int o = 0; o < K; ++o) {
            int off[D];
            offsets<D>(o, off);
            float dpos[D];
#pragma unroll
            for (int a = 0; a < D; ++a) dpos[a] = float(off[a]) - fx[a];
            const float4 g = node(off);
            const float gvel[3] = {g.x, g.y, g.z};
            float weight = w[0][off[0]];
#pragma unroll
            for (int a = 1; a < D; ++a) weight = weight * w[a][off[a]];
#pragma unroll
            for (int a = 0; a < D; ++a) v[a] = v[a] + weight * gvel[a];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float wv = four_inv_dx * (weight * gvel[a]);
#pragma unroll
                for (int b = 0; b < D; ++b) C.a[a][b] = C.a[a][b] + wv * dpos[b];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = x[a] + sp.dt * v[a];
    M<D> Mx;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            const float id1 = (a == b && (a < 2 || sp.identity_mode == 1)) ? 1.0f : 0.0f;
            Mx.a[a][b] = id1 + sp.dt * C.a[a][b];
        }
    M<D> Fh = matmul<D>(Mx, F);
    if (material == MPM_JELLY) {
        F = Fh;
    } else {
        M<D> U, V;
        float sig[D];
        step_svd<MODE, D>(Fh, U, sig, V);
        if (material == MPM_SNOW) {
            const float lo = float(1.0 - 2.5e-2), hi = float(1.0 + 4.5e-3);
#pragma unroll
            for (int d = 0; d < D; ++d) sig[d] = (sig[d] < lo) ? lo : ((hi < sig[d]) ? hi : sig[d]);
            const float old_J = det(Fh);
            M<D> US;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) US.a[a][b] = U.a[a][b] * sig[b];
            const M<D> Fn = matmul_bt<D>(US, V);
            Jp = Jp * old_J / det(Fn);
            Jp = (Jp < 0.6f) ? 0.6f : ((20.0f < Jp) ? 20.0f : Jp);
            F = Fn;
        } else {
            double J = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) J *= double(sig[d]);
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b)
                    F.a[a][b] = (a == b && (a < 2 || sp.identity_mode == 1)) ? 1.0f : 0.0f;
            F.a[0][0] = float(J);
        }
    }
}

// The grid update of one node (src/nclr.h:263-310): g = {momentum, mass} summed over the
// contributions; node coordinates n.  Returns {velocity, mass}, zero where the boundary
// condition applies.
template <int D>
NCLR_DI float4 grid_node(float4 g, const int (&n)[D], const SceneParams& sp) {
    float vel[D];
    vel[0] = g.x;
    vel[1] = g.y;
    if constexpr (D == 3) vel[2] = g.z;
    float m = g.w;
    if (m > 0.0f) {
#pragma unroll
        for (int a = 0; a < D; ++a) vel[a] = vel[a] / m;
        vel[1] = vel[1] + sp.dt * sp.gravity;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const float lo = -sp.allowed, hi = sp.allowed;
            vel[a] = (vel[a] < lo) ? lo : ((hi < vel[a]) ? hi : vel[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        if ((n[a] < 3 && vel[a] < 0.0f) || (n[a] >= (sp.res + 1) - 3 && vel[a] > 0.0f)) {
#pragma unroll
            for (int b = 0; b < D; ++b) vel[b] = 0.0f;
            m = 0.0f;
        }
    }
    return make_float4(vel[0], vel[1], D == 3 ? vel[2] : 0.0f, m);
}

}  // namespace nclr_dev
