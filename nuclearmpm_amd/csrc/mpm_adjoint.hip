// Reverse-mode (adjoint) kernels of one MLS-MPM step (jelly, snow, liquid) for gfx950.
//
// Given the cotangent of the state after a step (x', v', F', C', Jp') these kernels produce the
// cotangent of the state before it (x, v, F, C, Jp) and of Young's modulus, per scene.  They run after
// the forward sort, P2G and grid kernels have been recomputed for the state before the step, and
// reuse that step's tiles, work items and P2G slots:
//
//   g2p_adjoint   one workgroup per P2G work item: per particle, the cotangents of the gathered
//                 velocity and affine matrix are scattered back to the 3^D nodes (exact fixed-point
//                 sums in LDS, as in P2G, stored per work item); the partial x and F cotangents
//                 that G2P itself contributes go to grad_in.  Snow and liquid first recompute the
//                 forward's F^ = (I' + dt C') F bit for bit and pass the cotangents of F' and Jp'
//                 back through the material update: snow's F' = U clamp(S) V^T (the backward pass
//                 of a function of the singular values) and Jp' = clamp(Jp det F^ / det F'),
//                 liquid's F'_00 = prod S
//   grid_adjoint  one workgroup per node tile: sums both slot sets (or, when the forward grid
//                 kernel kept the pre-update sums, reads those and sums the adjoint slots), recomputes grid_op
//                 (src/nclr.h:263-310) and maps the velocity cotangent to the cotangent of the
//                 pre-update {momentum, mass} (clamped components and sticky nodes pass nothing)
//   p2g_adjoint   per particle: gathers the grid cotangent through the P2G stencil and the fused
//                 affine matrix (src/nclr.h:104-165, 313-337), including the backward pass of the
//                 polar decomposition F = R S: for a cotangent Rbar of R, Fbar = 2 R Y with Y skew
//                 and S Y + Y S = skew(R^T Rbar); snow's hardening e = exp(10 (1 - Jp)) adds the
//                 Jp cotangent of mu = mu0 e and lambda = lambda0 e; completes grad_in and the
//                 per-particle dL/dE
//   e_sum         per scene, a fixed-order double sum of the dL/dE contributions
//
// The same restatement on the CPU in FP64 (the oracle's adjoint) is checked against central
// finite differences of the step.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mpm_device.hpp"

namespace nclr_dev {

// d w_k / d fx of the quadratic B-spline weights (src/nclr.h:124-127)
NCLR_DI void bspline_dweights(float fx, float (&dw)[3]) {
    dw[0] = fx - 1.5f;
    dw[1] = -2.0f * (fx - 1.0f);
    dw[2] = fx - 0.5f;
}

// d det(A) / dA
NCLR_DI M<2> cofactor(const M<2>& a) {
    M<2> c;
    c.a[0][0] = a.a[1][1];
    c.a[0][1] = -a.a[1][0];
    c.a[1][0] = -a.a[0][1];
    c.a[1][1] = a.a[0][0];
    return c;
}

NCLR_DI M<3> cofactor(const M<3>& a) {
    M<3> c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            c.a[i][j] = a.a[i1][j1] * a.a[i2][j2] - a.a[i1][j2] * a.a[i2][j1];
        }
    return c;
}

// Fbar += d<Rbar, R(F)>/dF for the rotation factor R of the polar decomposition F = R S.
template <int D>
NCLR_DI void polar_backward(const M<D>& F, const M<D>& R, const M<D>& Rbar, M<D>& Fbar) {
    M<D> S, Mm;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float s = 0.0f, m = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                s += R.a[k][i] * F.a[k][j];
                m += R.a[k][i] * Rbar.a[k][j];
            }
            S.a[i][j] = s;
            Mm.a[i][j] = m;
        }
    M<D> Y;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Y.a[i][j] = 0.0f;
    if constexpr (D == 2) {
        const float y = 0.5f * (Mm.a[1][0] - Mm.a[0][1]) / (S.a[0][0] + S.a[1][1]);
        Y.a[1][0] = y;
        Y.a[0][1] = -y;
    } else {
        // skew(M) as a vector m: (tr(S) I - sym(S)) y = m
        const float mv[3] = {0.5f * (Mm.a[2][1] - Mm.a[1][2]), 0.5f * (Mm.a[0][2] - Mm.a[2][0]),
                             0.5f * (Mm.a[1][0] - Mm.a[0][1])};
        const float tr = S.a[0][0] + S.a[1][1] + S.a[2][2];
        M<3> Q;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Q.a[i][j] = (i == j ? tr : 0.0f) - 0.5f * (S.a[i][j] + S.a[j][i]);
        const M<3> Qc = cofactor(Q);
        const float inv_det = 1.0f / det(Q);
        float y[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) y[i] = (Qc.a[0][i] * mv[0] + Qc.a[1][i] * mv[1] + Qc.a[2][i] * mv[2]) * inv_det;
        Y.a[2][1] = y[0];
        Y.a[1][2] = -y[0];
        Y.a[0][2] = y[1];
        Y.a[2][0] = -y[1];
        Y.a[1][0] = y[2];
        Y.a[0][1] = -y[2];
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) s += R.a[i][k] * Y.a[k][j];
            Fbar.a[i][j] += 2.0f * s;
        }
}

// Fbar += the cotangent of A through A -> U f(S) V^T, f = the snow clamp of the signed singular
// values (src/nclr.h:241-245).  With P = U^T Gbar V: Q_ii = f'(s_i) P_ii and, for i != j,
// Q_ij = a_ij (P_ij + P_ji) / 2 + b_ij (P_ij - P_ji) / 2 with the divided differences
// a_ij = (f_i - f_j) / (s_i - s_j) (1 when both are inside the clamp, 0 when both are clamped to the
// same bound) and b_ij = (f_i + f_j) / (s_i + s_j); then Fbar += U Q V^T.
template <int D>
NCLR_DI void clamp_svd_backward(const M<D>& U, const float (&sig)[D], const M<D>& V, float lo, float hi,
                                const M<D>& Gbar, M<D>& Fbar) {
    float f[D];
    int side[D];  // -1: at or below lo, 1: at or above hi, 0: inside
#pragma unroll
    for (int i = 0; i < D; ++i) {
        side[i] = sig[i] <= lo ? -1 : (hi <= sig[i] ? 1 : 0);
        f[i] = side[i] < 0 ? lo : (side[i] > 0 ? hi : sig[i]);
    }
    M<D> GV, P;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float t = 0.0f;
#pragma unroll
            for (int b = 0; b < D; ++b) t += Gbar.a[a][b] * V.a[b][j];
            GV.a[a][j] = t;
        }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float t = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) t += U.a[a][i] * GV.a[a][j];
            P.a[i][j] = t;
        }
    M<D> Q;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (i == j) {
                Q.a[i][i] = side[i] == 0 ? P.a[i][i] : 0.0f;
                continue;
            }
            const float a = (side[i] == 0 && side[j] == 0) ? 1.0f
                            : (side[i] == side[j] ? 0.0f : (f[i] - f[j]) / (sig[i] - sig[j]));
            const float b = (f[i] + f[j]) / (sig[i] + sig[j]);
            Q.a[i][j] = 0.5f * (a * (P.a[i][j] + P.a[j][i]) + b * (P.a[i][j] - P.a[j][i]));
        }
    M<D> UQ;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            float t = 0.0f;
#pragma unroll
            for (int i = 0; i < D; ++i) t += U.a[a][i] * Q.a[i][j];
            UQ.a[a][j] = t;
        }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            float t = 0.0f;
#pragma unroll
            for (int j = 0; j < D; ++j) t += UQ.a[a][j] * V.a[b][j];
            Fbar.a[a][b] += t;
        }
}

// Cotangents of F^ = (I' + dt C') F and of the old Jp from those of the material update's outputs
// (src/nclr.h:230-258): jelly F' = F^; snow F' = U clamp(S) V^T and Jp' = clamp(Jp det F^ / det F',
// 0.6, 20); liquid F' = diag<dim>(1) with F'_00 = J = prod S (|det F^| in 2D, det F^ in 3D).  The SVD
// is the forward's (same method and operation order).
template <int D, bool EXACT>
NCLR_DI void material_backward(int material, const M<D>& Fh, float Jp, const M<D>& Fbar_new, float Jbar_new,
                               M<D>& Fhbar, float& Jbar) {
    if (material == MPM_JELLY) {
        Fhbar = Fbar_new;
        Jbar = Jbar_new;
        return;
    }
    M<D> U, V;
    float sig[D];
    step_svd<EXACT ? kSvdExact : kSvdFast, D>(Fh, U, sig, V);
    if (material == MPM_SNOW) {
        const float lo = float(1.0 - 2.5e-2), hi = float(1.0 + 4.5e-3);
        M<D> US;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) US.a[a][b] = U.a[a][b] * ((sig[b] < lo) ? lo : ((hi < sig[b]) ? hi : sig[b]));
        const M<D> Fn = matmul_bt<D>(US, V);
        const float old_J = det(Fh), new_J = det(Fn);
        const float r = Jp * old_J / new_J;
        const float rbar = (r > 0.6f && r < 20.0f) ? Jbar_new : 0.0f;
        Jbar = rbar * old_J / new_J;
        const M<D> cof_h = cofactor(Fh), cof_n = cofactor(Fn);
        const float c1 = rbar * Jp / new_J, c2 = rbar * Jp * old_J / (new_J * new_J);
        M<D> Gn;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                Fhbar.a[a][b] = c1 * cof_h.a[a][b];
                Gn.a[a][b] = Fbar_new.a[a][b] - c2 * cof_n.a[a][b];
            }
        clamp_svd_backward<D>(U, sig, V, lo, hi, Gn, Fhbar);
    } else {
        double J = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) J *= double(sig[d]);
        const float sgn = float(J) * det(Fh) < 0.0f ? -1.0f : 1.0f;
        const M<D> cof = cofactor(Fh);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) Fhbar.a[a][b] = Fbar_new.a[0][0] * sgn * cof.a[a][b];
        Jbar = Jbar_new;
    }
}

NCLR_DI float ldg(const float* __restrict__ g, int64_t gstride, int c, int64_t j) { return g[int64_t(c) * gstride + j]; }

// ---------------------------------------------------------------------------------------------
// G2P adjoint (src/nclr.h:167-258).  GENERAL: any material (the scene's, at run time); otherwise
// every scene is jelly.  A work item of TPB * PPT particles, PPT per thread.  At least two waves per
// SIMD (at most 256 VGPRs): measured 1.45x faster than one wave per SIMD; four (128 VGPRs) with one
// 3D particle per thread.
// ---------------------------------------------------------------------------------------------
template <int D, int PPT, int TPB, bool GENERAL, bool EXACT>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(D == 3 && PPT == 1 ? 4 : 2))) void k_g2p_adjoint(const float* __restrict__ st, const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ item_offset, StepParams p,
                                                     const float4* __restrict__ gv, const float* __restrict__ gout,
                                                     float* __restrict__ gin, int64_t gstride,
                                                     float4* __restrict__ slots_b) {
    using TL = Tile<D>;
    using L = Layout<D>;
    constexpr int REG = TL::REGION;
    constexpr int K = D == 2 ? 9 : 27;
    __shared__ unsigned long long acc[D * REG];
    __shared__ uint32_t nonfinite[REG];
    constexpr int NW = TPB / 64;
    __shared__ float s_bound[NW][D];
    const uint32_t item = blockIdx.x;
    const WorkItem wi = work_item(p, item);  // loaded with the count below, used only if valid
    if (item >= item_offset[p.ntiles]) return;
    for (int r = threadIdx.x; r < D * REG; r += blockDim.x) acc[r] = 0ull;
    for (int r = threadIdx.x; r < REG; r += blockDim.x) nonfinite[r] = 0u;
    __syncthreads();
    const uint32_t tile = wi.tile;
    const SceneParams sp = p.scenes[p.nscenes == 1 ? 0 : scene_of_tile(p, tile)];
    int torg[D];
    tile_coords<D>(tile - sp.tile_base, sp.nt, torg);
#pragma unroll
    for (int a = 0; a < D; ++a) torg[a] *= TL::S;
    const uint32_t j0 = wi.j0, j1 = wi.j1;
    const float four_inv_dx = 4.0f * sp.inv_dx;

    // pass 1: cotangents of the gathered velocity (vb = v' + dt x') and affine matrix
    // (Cb = C' + dt F' F^T) and a bound of the scattered sums per channel
    bool valid[PPT];
    int bases[PPT][D];
    float fxs[PPT][D], vb[PPT][D];
    M<D> Cb[PPT];
    float bound[D];
#pragma unroll
    for (int c = 0; c < D; ++c) bound[c] = 0.0f;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const uint32_t j = j0 + uint32_t(k) * TPB + threadIdx.x;
        valid[k] = false;
        if (j >= j1) continue;
        const int64_t i = perm[j];
        float x[D];
        M<D> F;
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = ldf(st, p.stride, L::X + a, i);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) F.a[a][b] = ldf(st, p.stride, L::F + a * D + b, i);
        int base[D];
        float fx[D], w[D][3];
        if (!stencil<D>(x, sp, base, fx, w)) continue;  // keyed into this tile, so never taken
        valid[k] = true;
        float wsum = 1.0f;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            bases[k][a] = base[a];
            fxs[k][a] = fx[a];
            vb[k][a] = ldg(gout, gstride, L::V + a, j) + sp.dt * ldg(gout, gstride, L::X + a, j);
            wsum *= fabsf(w[a][0]) + fabsf(w[a][1]) + fabsf(w[a][2]);
        }
        // Fb: cotangent of F^ = (I' + dt C') F
        M<D> Fb;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) Fb.a[a][b] = ldg(gout, gstride, L::F + a * D + b, j);
        if constexpr (GENERAL) {
            // the forward's C' (k_g2p_sorted's operation order, so F^ is the forward's bit for bit)
            M<D> Cf;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) Cf.a[a][b] = 0.0f;
            const StencilRows<D> node(gv, base, sp.ntn, sp.node_tile_base);
#pragma unroll
            for (int o = 0; o < K; ++o) {
                int off[D];
                float dpos[D];
                offsets<D>(o, off);
#pragma unroll
                for (int a = 0; a < D; ++a) dpos[a] = float(off[a]) - fx[a];
                const float4 g = node(off);
                const float gvel[3] = {g.x, g.y, g.z};
                float weight = w[0][off[0]];
#pragma unroll
                for (int a = 1; a < D; ++a) weight = weight * w[a][off[a]];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const float wv = four_inv_dx * (weight * gvel[a]);
#pragma unroll
                    for (int b = 0; b < D; ++b) Cf.a[a][b] = Cf.a[a][b] + wv * dpos[b];
                }
            }
            M<D> Mx;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    const float id1 = (a == b && (a < 2 || sp.identity_mode == 1)) ? 1.0f : 0.0f;
                    Mx.a[a][b] = id1 + sp.dt * Cf.a[a][b];
                }
            const M<D> Fh = matmul<D>(Mx, F);
            const float Jp = ldf(st, p.stride, L::JP, i);
            const M<D> Fbn = Fb;
            float Jb;
            material_backward<D, EXACT>(sp.material, Fh, Jp, Fbn, ldg(gout, gstride, L::JP, j), Fb, Jb);
            // F cotangent (I' + dt C')^T Fbar^ and the Jp cotangent of the G2P update
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    float t = 0.0f;
#pragma unroll
                    for (int q = 0; q < D; ++q) t += Mx.a[q][a] * Fb.a[q][b];
                    gin[int64_t(L::F + a * D + b) * gstride + i] = t;
                }
            gin[int64_t(L::JP) * gstride + i] = Jb;
        }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                float s = 0.0f;
#pragma unroll
                for (int q = 0; q < D; ++q) s += Fb.a[a][q] * F.a[b][q];
                Cb[k].a[a][b] = ldg(gout, gstride, L::C + a * D + b, j) + sp.dt * s;
            }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            float s = 0.0f;
#pragma unroll
            for (int b = 0; b < D; ++b) s += fabsf(Cb[k].a[a][b]);
            // |off - fx| <= 1.5 on every axis
            const float bc = wsum * (fabsf(vb[k][a]) + 1.5f * four_inv_dx * s);
            if (bc < 3.0e38f) bound[a] = fmaxf(bound[a], bc);
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
        float b = bound[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) b = fmaxf(b, __shfl_xor(b, off));
        bound[c] = b;
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < D; ++c) s_bound[threadIdx.x >> 6][c] = bound[c];
    __syncthreads();
    float scale[D];
    double inv_scale[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        float b = s_bound[0][c];
#pragma unroll
        for (int w = 1; w < NW; ++w) b = fmaxf(b, s_bound[w][c]);
        fixed_point_scale(b, scale[c], inv_scale[c]);
    }

    // pass 2: gather the updated grid (for the forward C and the weight derivatives), scatter the
    // quantised grid cotangent, write the partial x and F cotangents
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!valid[k]) continue;
        const uint32_t j = j0 + uint32_t(k) * TPB + threadIdx.x;
        float w[D][3], dw[D][3];
        int rb = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            bspline_weights(fxs[k][a], w[a]);
            bspline_dweights(fxs[k][a], dw[a]);
            rb = rb * TL::R + (bases[k][a] - torg[a]);
        }
        M<D> Cn;
        float fxbar[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            fxbar[a] = 0.0f;
#pragma unroll
            for (int b = 0; b < D; ++b) Cn.a[a][b] = 0.0f;
        }
        const StencilRows<D> node(gv, bases[k], sp.ntn, sp.node_tile_base);
        // 3D: unrolled by rows of three nodes (fully unrolled, the early node loads spill)
#pragma unroll(D == 3 ? 3 : K)
        for (int o = 0; o < K; ++o) {
            // products and sums contract to fused multiply-adds here (the kernel is bound by its
            // VALU issue; these derivatives have no bitwise counterpart to match)
#pragma clang fp contract(fast)
            int off[D];
            offsets<D>(o, off);
            float d[D];
            int ro = 0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                d[a] = float(off[a]) - fxs[k][a];
                ro = ro * TL::R + off[a];
            }
            const float4 g = node(off);
            const float gvel[3] = {g.x, g.y, g.z};
            float weight = pick3(w[0], off[0]);
#pragma unroll
            for (int a = 1; a < D; ++a) weight *= pick3(w[a], off[a]);
            float wbar = 0.0f;
            const int r = rb + ro;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                float cd = 0.0f;
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    cd += Cb[k].a[a][b] * d[b];
                    Cn.a[a][b] += four_inv_dx * weight * gvel[a] * d[b];
                }
                const float val = weight * (vb[k][a] + four_inv_dx * cd);
                wbar += gvel[a] * (vb[k][a] + four_inv_dx * cd);
                const float y = val * scale[a];
                if (fabsf(y) < kFixedLimit) atomicAdd(&acc[a * REG + r], fixed_point(y));
                else atomicOr(&nonfinite[r], 1u << a);
            }
#pragma unroll
            for (int b = 0; b < D; ++b) {
                float ctv = 0.0f;
#pragma unroll
                for (int a = 0; a < D; ++a) ctv += Cb[k].a[a][b] * gvel[a];
                fxbar[b] -= four_inv_dx * weight * ctv;
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                float dwa = pick3(dw[a], off[a]);
#pragma unroll
                for (int b = 0; b < D; ++b)
                    if (b != a) dwa *= pick3(w[b], off[b]);
                fxbar[a] += wbar * dwa;
            }
        }
        const int64_t i = perm[j];
#pragma unroll
        for (int a = 0; a < D; ++a)
            gin[int64_t(L::X + a) * gstride + i] = ldg(gout, gstride, L::X + a, j) + sp.inv_dx * fxbar[a];
        if constexpr (!GENERAL) {
            // jelly: F' = (I' + dt C) F  =>  F cotangent (I' + dt C)^T F'; Jp is carried unchanged
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    float s = 0.0f;
#pragma unroll
                    for (int q = 0; q < D; ++q) {
                        const float id1 = (q == a && (a < 2 || sp.identity_mode == 1)) ? 1.0f : 0.0f;
                        s += (id1 + sp.dt * Cn.a[q][a]) * ldg(gout, gstride, L::F + q * D + b, j);
                    }
                    gin[int64_t(L::F + a * D + b) * gstride + i] = s;
                }
            gin[int64_t(L::JP) * gstride + i] = ldg(gout, gstride, L::JP, j);
        }
    }
    __syncthreads();
    float4* dst = slots_b + int64_t(item) * REG;
    const float nan = __builtin_nanf("");
    for (int r = threadIdx.x; r < REG; r += blockDim.x) {
        float out[3] = {0.0f, 0.0f, 0.0f};
        const uint32_t bad = nonfinite[r];
#pragma unroll
        for (int c = 0; c < D; ++c)
            out[c] = (bad >> c) & 1u ? nan : fixed_point_value(acc[c * REG + r], inv_scale[c]);
        dst[r] = make_float4(out[0], out[1], out[2], 0.0f);
    }
}

// ---------------------------------------------------------------------------------------------
// grid adjoint (src/nclr.h:263-310)
// ---------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_grid_adjoint(const float4* __restrict__ slots,
                                                      const float4* __restrict__ slots_b,
                                                      const float4* __restrict__ gpre,
                                                      const uint32_t* __restrict__ item_offset, StepParams p,
                                                      float4* __restrict__ gbar) {
    using TL = Tile<D>;
    constexpr int NC = 1 << D;
    __shared__ int s_first[NC];
    __shared__ int s_count[NC];
    const uint32_t nt_idx = blockIdx.x;
    const SceneParams sp = p.scenes[p.nscenes == 1 ? 0 : scene_of_node_tile(p, nt_idx)];
    int nto[D];
    tile_coords<D>(nt_idx - sp.node_tile_base, sp.ntn, nto);
    bool occupied = false;
    if (threadIdx.x < NC) {
        int t = 0;
        bool valid = true;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int ta = nto[a] - ((threadIdx.x >> a) & 1);
            valid = valid && ta >= 0 && ta < sp.nt;
            t = t * sp.nt + ta;
        }
        const uint32_t gt = sp.tile_base + uint32_t(t);
        const uint32_t first = valid ? item_offset[gt] : 0u;
        const uint32_t count = valid ? item_offset[gt + 1] - first : 0u;
        occupied = count > 0;
        s_first[threadIdx.x] = occupied ? int(first) : -1;
        s_count[threadIdx.x] = int(count);
    }
    if (!__syncthreads_or(occupied)) return;  // no particle's stencil reaches this node tile
    int first[NC], count[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        first[c] = s_first[c];
        count[c] = s_count[c];
    }
    for (int l = threadIdx.x; l < TL::NODES; l += blockDim.x) {
        int n[D];
        {
            int rem = l;
#pragma unroll
            for (int a = D - 1; a >= 0; --a) {
                n[a] = nto[a] * TL::S + (rem & (TL::S - 1));
                rem >>= TL::LOG;
            }
        }
        float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // 2D: the forward grid kernel's four partial sums (item it of a candidate in partial it % 4)
        float4 g1 = g, g2 = g, g3 = g;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (first[c] < 0) continue;
            bool inside = true;
            int r = 0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const int ra = n[a] - (nto[a] - ((c >> a) & 1)) * TL::S;
                inside = inside && ra < TL::R;
                r = r * TL::R + ra;
            }
            if (!inside) continue;
            // items in order (the forward grid kernel's order)
            // rounds of loads in flight before their (ordered) sums; an item past the count adds
            // +0.0, which leaves a sum that started at +0.0 bitwise unchanged
            const int64_t q = int64_t(first[c]) * TL::REGION + r;
            if (gpre) {  // the forward sums are kept: only the adjoint slots, eight per round
                const float4* qb = slots_b + q;
                for (int it = 0; it < count[c]; it += 8, qb += 8 * TL::REGION) {
                    float4 bv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        bv[u] = it + u < count[c] ? qb[u * TL::REGION] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        b.x += bv[u].x;
                        b.y += bv[u].y;
                        b.z += bv[u].z;
                        b.w += bv[u].w;
                    }
                }
                continue;
            }
            // without the tape: both slot sets
            auto add = [&](float4& acc, const float4 sa, const float4 sb) {
                acc.x += sa.x;
                acc.y += sa.y;
                acc.z += sa.z;
                acc.w += sa.w;
                b.x += sb.x;
                b.y += sb.y;
                b.z += sb.z;
            };
            if constexpr (D == 2) {  // four items' loads in flight, into the four partial sums
                int it = 0;
                for (; it + 4 <= count[c]; it += 4) {
                    const int64_t s0 = q + int64_t(it) * TL::REGION;
                    const float4 a0 = slots[s0], a1 = slots[s0 + TL::REGION], a2 = slots[s0 + 2 * TL::REGION],
                                 a3 = slots[s0 + 3 * TL::REGION];
                    const float4 b0 = slots_b[s0], b1 = slots_b[s0 + TL::REGION], b2 = slots_b[s0 + 2 * TL::REGION],
                                 b3 = slots_b[s0 + 3 * TL::REGION];
                    add(g, a0, b0);
                    add(g1, a1, b1);
                    add(g2, a2, b2);
                    add(g3, a3, b3);
                }
                const int64_t s0 = q + int64_t(it) * TL::REGION;
                if (it < count[c]) add(g, slots[s0], slots_b[s0]);
                if (it + 1 < count[c]) add(g1, slots[s0 + TL::REGION], slots_b[s0 + TL::REGION]);
                if (it + 2 < count[c]) add(g2, slots[s0 + 2 * TL::REGION], slots_b[s0 + 2 * TL::REGION]);
            } else {  // items in order, two items' four loads in flight
                int it = 0;
                for (; it + 2 <= count[c]; it += 2) {
                    const int64_t s0 = q + int64_t(it) * TL::REGION, s1 = s0 + TL::REGION;
                    const float4 a0 = slots[s0], b0 = slots_b[s0], a1 = slots[s1], b1 = slots_b[s1];
                    add(g, a0, b0);
                    add(g, a1, b1);
                }
                if (it < count[c]) {
                    const int64_t s0 = q + int64_t(it) * TL::REGION;
                    add(g, slots[s0], slots_b[s0]);
                }
            }
        }
        if constexpr (D == 2) {
            if (!gpre) {  // the partial sums added in order, as the forward grid kernel adds them
                const float4 parts[3] = {g1, g2, g3};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    g.x += parts[k].x;
                    g.y += parts[k].y;
                    g.z += parts[k].z;
                    g.w += parts[k].w;
                }
            }
        }
        if (gpre) g = gpre[int64_t(nt_idx) * TL::NODES + l];  // the same sums, in the same order
        const float mom[3] = {g.x, g.y, g.z};
        const float vbar[3] = {b.x, b.y, b.z};
        const float m = g.w;
        // forward grid_op, remembering which components were clamped and whether the sticky
        // boundary zeroed the node
        float vel[D];
        bool clamped[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            vel[a] = mom[a];
            clamped[a] = false;
        }
        if (m > 0.0f) {
#pragma unroll
            for (int a = 0; a < D; ++a) vel[a] = vel[a] / m;
            vel[1] = vel[1] + sp.dt * sp.gravity;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                clamped[a] = vel[a] < -sp.allowed || sp.allowed < vel[a];
                vel[a] = (vel[a] < -sp.allowed) ? -sp.allowed : ((sp.allowed < vel[a]) ? sp.allowed : vel[a]);
            }
        }
        bool sticky = false;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            if ((n[a] < 3 && vel[a] < 0.0f) || (n[a] >= (sp.res + 1) - 3 && vel[a] > 0.0f)) {
#pragma unroll
                for (int q = 0; q < D; ++q) vel[q] = 0.0f;
                sticky = true;
            }
        }
        float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (m > 0.0f && !sticky) {
            const float inv_m2 = 1.0f / (m * m);
            float mb = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const float vtb = clamped[a] ? 0.0f : vbar[a];
                out[a] = vtb / m;
                mb -= vtb * mom[a] * inv_m2;
            }
            out[3] = mb;
        }
        gbar[int64_t(nt_idx) * TL::NODES + l] = make_float4(out[0], out[1], out[2], out[3]);
    }
}

// ---------------------------------------------------------------------------------------------
// P2G adjoint (src/nclr.h:104-165, 313-372) per sorted position j.  Two waves per SIMD, as the
// G2P adjoint (1.38x faster than one).
// ---------------------------------------------------------------------------------------------
template <int D, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_p2g_adjoint(const float* __restrict__ st, const uint32_t* __restrict__ pid,
                                                     const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ keys, StepParams p,
                                                     const float4* __restrict__ gbar, const float* __restrict__ gout,
                                                     float* __restrict__ gin, int64_t gstride,
                                                     float* __restrict__ epart) {
    using L = Layout<D>;
    constexpr int K = D == 2 ? 9 : 27;
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= p.n) return;
    const int64_t i = perm[j];
    const int64_t s = p.stride;
    const SceneParams sp = p.scenes[p.nscenes == 1 ? 0 : scene_of_sorted(p, keys, j, pid[i])];
    float x[D], v[D];
    M<D> F, C;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        x[a] = ldf(st, s, L::X + a, i);
        v[a] = ldf(st, s, L::V + a, i);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            F.a[a][b] = ldf(st, s, L::F + a * D + b, i);
            C.a[a][b] = ldf(st, s, L::C + a * D + b, i);
        }
    const float Jp = ldf(st, s, L::JP, i);
    const float mass = ldf(st, s, L::MASS, i);
    const float vol = ldf(st, s, L::VOL, i);
    int base[D];
    float fx[D], w[D][3];
    if (!stencil<D>(x, sp, base, fx, w)) {
        // outside the grid the forward step carries the particle through unchanged
        for (int c = 0; c < L::NG; ++c) gin[int64_t(c) * gstride + i] = ldg(gout, gstride, c, j);
        epart[j] = 0.0f;
        return;
    }
    float dw[D][3];
#pragma unroll
    for (int a = 0; a < D; ++a) bspline_dweights(fx[a], dw[a]);
    // the forward's affine matrix from one polar factor: in the accurate mode its F - R is kept,
    // and the stress backward below takes R = F - (F - R) (within one float rounding of the
    // iteration's R) instead of a second Newton iteration; in the other modes R itself is kept and
    // F - R formed as the forward forms it
    M<D> A, FmR, R;
    if constexpr (D == 3 && MODE == kSvdAccurate) {
        FmR = f_minus_rotation<MODE, D>(F);
    } else {
        R = rotation_mode<MODE, D>(F);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) FmR.a[a][b] = F.a[a][b] - R.a[a][b];
    }
    A = fused_affine_from<D>(sp.material, F, FmR, C, Jp, mass, vol, sp);
    M<D> Abar;
    float fxbar[D], vin[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        fxbar[a] = 0.0f;
        vin[a] = 0.0f;
#pragma unroll
        for (int b = 0; b < D; ++b) Abar.a[a][b] = 0.0f;
    }
    const StencilRows<D> node(gbar, base, sp.ntn, sp.node_tile_base);
    // 3D: unrolled by rows of three nodes (fully unrolled, the early node loads spill)
#pragma unroll(D == 3 ? 3 : K)
    for (int o = 0; o < K; ++o) {
        // fused multiply-adds here, as in the G2P adjoint's stencil loop
#pragma clang fp contract(fast)
        int off[D];
        offsets<D>(o, off);
        float dpos[D];
#pragma unroll
        for (int a = 0; a < D; ++a) dpos[a] = (float(off[a]) - fx[a]) * sp.dx;
        const float4 g = node(off);
        const float gm[3] = {g.x, g.y, g.z};
        float weight = pick3(w[0], off[0]);
#pragma unroll
        for (int a = 1; a < D; ++a) weight *= pick3(w[a], off[a]);
        float wbar = g.w * mass;
#pragma unroll
        for (int r = 0; r < D; ++r) {
            float ad = 0.0f;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                ad += A.a[r][c] * dpos[c];
                Abar.a[r][c] += weight * gm[r] * dpos[c];
            }
            wbar += gm[r] * (mass * v[r] + ad);
            vin[r] += weight * mass * gm[r];
        }
#pragma unroll
        for (int c = 0; c < D; ++c) {
            float atm = 0.0f;
#pragma unroll
            for (int r = 0; r < D; ++r) atm += A.a[r][c] * gm[r];
            fxbar[c] -= sp.dx * weight * atm;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            float dwa = pick3(dw[a], off[a]);
#pragma unroll
            for (int b = 0; b < D; ++b)
                if (b != a) dwa *= pick3(w[b], off[b]);
            fxbar[a] += wbar * dwa;
        }
    }
    // A = -(dt V0) Dinv PF + m C  ->  G = dL/dPF
    const float Dinv = 4.0f * sp.inv_dx * sp.inv_dx;
    const float e = hardening(sp.material, Jp);
    const float mu = sp.mu0 * e, lambda = sp.lambda0 * e;
    M<D> G;
    float sG = 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            G.a[a][b] = -(sp.dt * vol) * Dinv * Abar.a[a][b];
            sG += G.a[a][b];
        }
    // PF = 2 mu (F - R) F^T + lambda (J - 1) J on every entry
    if constexpr (D == 3 && MODE == kSvdAccurate) {
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) R.a[a][b] = F.a[a][b] - FmR.a[a][b];
    }
    const float J = det(F);
    M<D> Fbar, Rbar;
    float mubar = 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            float gf = 0.0f, gtfr = 0.0f, fmrf = 0.0f;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                gf += G.a[a][q] * F.a[q][b];
                gtfr += G.a[q][a] * (F.a[q][b] - R.a[q][b]);
                fmrf += (F.a[a][q] - R.a[a][q]) * F.a[b][q];
            }
            Fbar.a[a][b] = 2.0f * mu * (gf + gtfr);
            Rbar.a[a][b] = -2.0f * mu * gf;
            mubar += G.a[a][b] * 2.0f * fmrf;
        }
    polar_backward<D>(F, R, Rbar, Fbar);
    const M<D> cof = cofactor(F);
    const float Jbar = sG * lambda * (2.0f * J - 1.0f);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        gin[int64_t(L::X + a) * gstride + i] += sp.inv_dx * fxbar[a];
        gin[int64_t(L::V + a) * gstride + i] = vin[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            gin[int64_t(L::F + a * D + b) * gstride + i] += Fbar.a[a][b] + Jbar * cof.a[a][b];
            gin[int64_t(L::C + a * D + b) * gstride + i] = mass * Abar.a[a][b];
        }
    const float lambar = sG * (J - 1.0f) * J;
    // snow: mu = mu0 e, lambda = lambda0 e with e = exp(10 (1 - Jp)), de/dJp = -10 e
    if (sp.material == MPM_SNOW) gin[int64_t(L::JP) * gstride + i] += (mubar * sp.mu0 + lambar * sp.lambda0) * (-10.0f * e);
    epart[j] = e * (mubar * sp.dmu_dE + lambar * sp.dlambda_dE);
}

// grad_E[s] += the sum of epart over scene s's sorted particles, in a fixed order, in three
// kernels: the scene's sorted range [lo, hi) (scenes occupy consecutive keys; particles off their
// grid sort last), partial sums of consecutive chunks of kEChunk positions counted from lo (each a
// fixed tree), then the scene's chunks in order.  The partition depends only on the scene's own
// range, so a scene's dL/dE is the same alone and inside a batch.
constexpr int kEChunk = 2048;

// Boundary t of the range table (scene t / 2's first sorted position for even t, one past its last
// for odd t): the first sorted position with keys[.] >= the boundary's key.  One wave per boundary
// searches 64-ary: the answer lies in [lo, hi] (keys[lo - 1] < key, and hi = n or keys[hi] >= key);
// each round the lanes probe 64 evenly spaced positions and the first probe at or above the key
// narrows the interval to one spacing, so 2^21 keys take four rounds of one load per lane.
__global__ __launch_bounds__(64) void k_scene_ranges(const uint32_t* __restrict__ keys, StepParams p,
                                                     int64_t* __restrict__ range) {
    const int t = int(blockIdx.x);
    if (t >= 2 * p.nscenes) return;
    const int sc = t >> 1;
    const uint32_t key = (t & 1) == 0 ? p.scenes[sc].tile_base
                                      : (sc + 1 < p.nscenes ? p.scenes[sc + 1].tile_base : p.ntiles);
    const int lane = int(threadIdx.x);
    int64_t lo = 0, hi = p.n;
    while (hi - lo > 64) {
        const int64_t step = (hi - lo + 63) / 64;
        // probe q; lane 63's is at or past hi - 1, and probes at or past hi load nothing
        const int64_t q = lo + int64_t(lane + 1) * step - 1;
        const bool at_or_above = q >= hi || keys[q] >= key;
        const uint64_t ball = __ballot(at_or_above);
        if (ball == 0ull) {  // every probe below the key, lane 63's at hi - 1: the answer is hi
            lo = lo + 64 * step;
        } else {
            const int f = __ffsll((unsigned long long)ball) - 1;
            const int64_t qf = lo + int64_t(f + 1) * step - 1;
            const int64_t new_lo = f == 0 ? lo : lo + int64_t(f) * step;  // one past probe f - 1
            hi = qf < hi ? qf : hi;
            lo = new_lo;
        }
    }
    // at most 64 candidates [lo, hi): the first at or above the key, else hi
    const int64_t q = lo + lane;
    const uint64_t ball = __ballot(q < hi && keys[q] >= key);
    if (lane == 0) range[t] = ball == 0ull ? hi : lo + (__ffsll((unsigned long long)ball) - 1);
}

// Sum of part[0..256) by a fixed tree; the result is in part[0] for thread 0.
NCLR_DI void tree_sum_256(double* part) {
    for (int off = 128; off > 0; off >>= 1) {
        __syncthreads();
        if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
    }
}

// Scene sc's sorted range: from the range table, or with range == nullptr (a single scene)
// [0, meta[0]), the particles on the grid, which the sort places first.
NCLR_DI void scene_range(const int64_t* __restrict__ range, const uint32_t* __restrict__ meta, int sc,
                         int64_t& lo, int64_t& hi) {
    if (range) {
        lo = range[2 * sc];
        hi = range[2 * sc + 1];
    } else {
        lo = 0;
        hi = meta[0];
    }
}

// Grid (chunks, scenes): chunk blockIdx.x of scene blockIdx.y -> partial[scene * chunks + chunk].
__global__ __launch_bounds__(256) void k_scene_e_partial(const float* __restrict__ epart,
                                                         const int64_t* __restrict__ range,
                                                         const uint32_t* __restrict__ meta, int chunks,
                                                         double* __restrict__ partial) {
    __shared__ double part[256];
    const int sc = int(blockIdx.y);
    int64_t s_lo, s_hi;
    scene_range(range, meta, sc, s_lo, s_hi);
    const int64_t lo = s_lo + int64_t(blockIdx.x) * kEChunk;
    const int64_t hi = min(s_hi, lo + int64_t(kEChunk));
    // the thread's elements j, j + 256, ... summed in that order, eight loads in flight
    double acc = 0.0;
    int64_t j = lo + threadIdx.x;
    for (; j + 7 * 256 < hi; j += 8 * 256) {
        float e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) e[u] = epart[j + u * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += double(e[u]);
    }
    for (; j < hi; j += 256) acc += double(epart[j]);
    part[threadIdx.x] = acc;
    tree_sum_256(part);
    if (threadIdx.x == 0) partial[int64_t(sc) * chunks + blockIdx.x] = part[0];
}

// One workgroup per scene: its used chunks, ceil((hi - lo) / kEChunk), in order.
__global__ __launch_bounds__(256) void k_scene_e_final(const double* __restrict__ partial,
                                                       const int64_t* __restrict__ range,
                                                       const uint32_t* __restrict__ meta, int chunks,
                                                       double* __restrict__ grad_E) {
    __shared__ double part[256];
    const int sc = int(blockIdx.x);
    int64_t s_lo, s_hi;
    scene_range(range, meta, sc, s_lo, s_hi);
    const int64_t len = s_hi - s_lo;
    const int used = int(min(int64_t(chunks), (len + kEChunk - 1) / kEChunk));
    double acc = 0.0;
    for (int c = int(threadIdx.x); c < used; c += 256) acc += partial[int64_t(sc) * chunks + c];
    part[threadIdx.x] = acc;
    tree_sum_256(part);
    if (threadIdx.x == 0) grad_E[sc] += part[0];
}

// dst[c][to(i)] = src[c][from(i)]: with to_particle_order, slot i goes to particle pid[i];
// otherwise slot i takes particle pid[i].
__global__ __launch_bounds__(256) void k_permute_channels(const uint32_t* __restrict__ pid, const float* __restrict__ src,
                                                          int64_t src_stride, float* __restrict__ dst,
                                                          int64_t dst_stride, int channels, int64_t n,
                                                          int to_particle_order) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = pid[i];
    const int64_t si = to_particle_order ? i : u;
    const int64_t di = to_particle_order ? u : i;
    for (int c = 0; c < channels; ++c) dst[int64_t(c) * dst_stride + di] = src[int64_t(c) * src_stride + si];
}

}  // namespace nclr_dev

namespace nclr_host {

using namespace nclr_dev;

static inline unsigned blocks_of(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

template <int D, int PPT, int TPB, bool GENERAL, bool EXACT>
static hipError_t g2p_adjoint_one(const LaunchArgs& a, const AdjointArgs& b) {
    // the forward's work items: TPB * PPT particles each
    const int64_t items = (a.p.n + TPB * PPT - 1) / (TPB * PPT) + a.p.ntiles;
    k_g2p_adjoint<D, PPT, TPB, GENERAL, EXACT><<<unsigned(items), TPB, 0, a.stream>>>(
        a.state_in, a.perm, a.item_offset, a.p, a.gv, b.grad_out, b.grad_in, b.gstride, b.slots_b);
    return hipGetLastError();
}

template <int D, int PPT, int TPB>
static hipError_t g2p_adjoint_material(const LaunchArgs& a, const AdjointArgs& b) {
    if (a.material == MPM_JELLY) return g2p_adjoint_one<D, PPT, TPB, false, true>(a, b);
    // 2D SVDs have one method; in 3D the forward's choice
    if (D == 2 || (a.svd_mode != 1)) return g2p_adjoint_one<D, PPT, TPB, true, true>(a, b);
    return g2p_adjoint_one<D, PPT, TPB, true, false>(a, b);
}

// 3D: one particle per thread (a workgroup of chunk threads), at most 128 VGPRs and four waves per
// SIMD, where two particles per thread need 144 and fit three.  2D: likewise when the particles
// fill fewer work items than the chip has CUs (half the latency of a work item), else 256
// threads and chunk / 256 particles each, as P2G.
template <int D>
static hipError_t g2p_adjoint_dispatch(const LaunchArgs& a, const AdjointArgs& b) {
    if constexpr (D == 2) {
        if (a.p.n >= int64_t(kFewItems) * a.p.chunk) {
            switch (a.p.chunk) {
                case 256: return g2p_adjoint_material<D, 1, 256>(a, b);
                case 1024: return g2p_adjoint_material<D, 4, 256>(a, b);
                default: return g2p_adjoint_material<D, 2, 256>(a, b);
            }
        }
    }
    switch (a.p.chunk) {
        case 256: return g2p_adjoint_material<D, 1, 256>(a, b);
        case 1024: return g2p_adjoint_material<D, 1, 1024>(a, b);
        default: return g2p_adjoint_material<D, 1, 512>(a, b);
    }
}

hipError_t launch_g2p_adjoint(const LaunchArgs& a, const AdjointArgs& b) {
    if (a.p.n == 0) return hipSuccess;
    return a.dim == 2 ? g2p_adjoint_dispatch<2>(a, b) : g2p_adjoint_dispatch<3>(a, b);
}

hipError_t launch_grid_adjoint(const LaunchArgs& a, const AdjointArgs& b) {
    const unsigned tiles = a.p.node_tiles;
    if (tiles == 0 || a.p.n == 0) return hipSuccess;
    if (a.dim == 2)
        k_grid_adjoint<2><<<tiles, 256, 0, a.stream>>>(a.slots, b.slots_b, a.gpre, a.item_offset, a.p, b.gbar);
    else
        k_grid_adjoint<3><<<tiles, 256, 0, a.stream>>>(a.slots, b.slots_b, a.gpre, a.item_offset, a.p, b.gbar);
    return hipGetLastError();
}

hipError_t launch_p2g_adjoint(const LaunchArgs& a, const AdjointArgs& b) {
    if (a.p.n > 0) {
        if (a.dim == 2)
            k_p2g_adjoint<2, kSvdExact><<<blocks_of(a.p.n, 256), 256, 0, a.stream>>>(
                a.state_in, a.pid_in, a.perm, a.keys, a.p, b.gbar, b.grad_out, b.grad_in, b.gstride, b.epart);
        else if (a.svd_mode == kSvdAccurate)
            k_p2g_adjoint<3, kSvdAccurate><<<blocks_of(a.p.n, 256), 256, 0, a.stream>>>(
                a.state_in, a.pid_in, a.perm, a.keys, a.p, b.gbar, b.grad_out, b.grad_in, b.gstride, b.epart);
        else if (a.svd_mode != kSvdFast)
            k_p2g_adjoint<3, kSvdExact><<<blocks_of(a.p.n, 256), 256, 0, a.stream>>>(
                a.state_in, a.pid_in, a.perm, a.keys, a.p, b.gbar, b.grad_out, b.grad_in, b.gstride, b.epart);
        else
            k_p2g_adjoint<3, kSvdFast><<<blocks_of(a.p.n, 256), 256, 0, a.stream>>>(
                a.state_in, a.pid_in, a.perm, a.keys, a.p, b.gbar, b.grad_out, b.grad_in, b.gstride, b.epart);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (a.p.n > 0 && b.grad_E) {
        // one scene: its range is [0, meta[0]), no search needed
        const int64_t* range = a.p.nscenes == 1 ? nullptr : b.erange;
        if (range) k_scene_ranges<<<unsigned(2 * a.p.nscenes), 64, 0, a.stream>>>(a.keys, a.p, b.erange);
        k_scene_e_partial<<<dim3(unsigned(b.echunks), unsigned(a.p.nscenes)), 256, 0, a.stream>>>(
            b.epart, range, a.meta, b.echunks, b.epartial);
        k_scene_e_final<<<unsigned(a.p.nscenes), 256, 0, a.stream>>>(b.epartial, range, a.meta, b.echunks, b.grad_E);
    }
    return hipGetLastError();
}

int e_chunks(int64_t n) { return int(std::max<int64_t>(1, (n + kEChunk - 1) / kEChunk)); }

hipError_t launch_permute_channels(const uint32_t* pid, const float* src, int64_t src_stride, float* dst,
                                   int64_t dst_stride, int channels, int64_t n, bool to_particle_order,
                                   hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_permute_channels<<<blocks_of(n, 256), 256, 0, s>>>(pid, src, src_stride, dst, dst_stride, channels, n,
                                                         to_particle_order ? 1 : 0);
    return hipGetLastError();
}

}  // namespace nclr_host
