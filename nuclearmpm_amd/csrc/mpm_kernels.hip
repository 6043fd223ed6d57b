// HIP kernels of the MLS-MPM step for gfx950 (MI355X).
//
// Reference step: MPMSimulation<dim>::advance() = p2g(); grid_op(); g2p();
// (src/nclr.h:80-84).  Data layout on the device:
//   particles  structure-of-arrays, field f of particle i at st[f*stride + i]
//              (fields: x[D] v[D] F[DxD] C[DxD] Jp mass volume, see Layout)
//   grid       blocked dense array of float4 {v0, v1, v2|0, m}: 3D blocks of
//              4x4x4 nodes, 2D blocks of 8x8 nodes, 64 nodes (one wave) per
//              block, blocks in x-major order.  `acc` receives P2G momentum and
//              mass and is all-zero between steps; `gv` holds the updated grid.
//   stamps     one uint32 per node block: the step stamp of the last P2G that
//              touched the block; the grid update and the dense export skip
//              blocks whose stamp is stale, so only the occupied part of the
//              grid is ever read or written.
#include <hip/hip_runtime.h>

#include "mpm_internal.hpp"
#include "mpm_math.hpp"

namespace nclr_dev {

template <int D>
struct Layout {
    static constexpr int X = 0;
    static constexpr int V = D;
    static constexpr int F = 2 * D;
    static constexpr int C = 2 * D + D * D;
    static constexpr int JP = 2 * D + 2 * D * D;
    static constexpr int MASS = JP + 1;
    static constexpr int VOL = JP + 2;
    static constexpr int NF = JP + 3;
};

template <int D>
struct BlockShape;
template <>
struct BlockShape<2> {
    static constexpr int LOG = 3;  // 8 x 8 nodes
};
template <>
struct BlockShape<3> {
    static constexpr int LOG = 2;  // 4 x 4 x 4 nodes
};

template <int D>
NCLR_DI int64_t node_index(const int (&n)[D], int nb) {
    constexpr int LB = BlockShape<D>::LOG;
    constexpr int MASK = (1 << LB) - 1;
    if constexpr (D == 2) {
        const int64_t blk = int64_t(n[0] >> LB) * nb + (n[1] >> LB);
        return blk * 64 + ((n[0] & MASK) << LB) + (n[1] & MASK);
    } else {
        const int64_t blk = (int64_t(n[0] >> LB) * nb + (n[1] >> LB)) * nb + (n[2] >> LB);
        return blk * 64 + ((n[0] & MASK) << (2 * LB)) + ((n[1] & MASK) << LB) + (n[2] & MASK);
    }
}

// Node coordinates of lane `lane` of node block `blk`.
template <int D>
NCLR_DI void node_coords(int64_t blk, int lane, int nb, int (&n)[D]) {
    constexpr int LB = BlockShape<D>::LOG;
    constexpr int MASK = (1 << LB) - 1;
    if constexpr (D == 2) {
        const int b1 = int(blk % nb), b0 = int(blk / nb);
        n[0] = (b0 << LB) + (lane >> LB);
        n[1] = (b1 << LB) + (lane & MASK);
    } else {
        const int b2 = int(blk % nb);
        const int64_t r = blk / nb;
        const int b1 = int(r % nb), b0 = int(r / nb);
        n[0] = (b0 << LB) + (lane >> (2 * LB));
        n[1] = (b1 << LB) + ((lane >> LB) & MASK);
        n[2] = (b2 << LB) + (lane & MASK);
    }
}

// base = trunc(x*inv_dx - 0.5), fx = x*inv_dx - base (src/nclr.h:115,121), with the
// stencil nodes base..base+2 required to lie in [0, res] on every axis.
template <int D>
NCLR_DI bool stencil(const float (&x)[D], const StepParams& p, int (&base)[D], float (&fx)[D],
                     float (&w)[D][3]) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const float s = x[a] * p.inv_dx;
        const float b = s - 0.5f;
        ok = ok && (b > -1.0f) && (b < float(p.res - 1));
        base[a] = ok ? int(b) : 0;
        fx[a] = s - float(base[a]);
        bspline_weights(fx[a], w[a]);
    }
    return ok;
}

template <int D>
NCLR_DI float ld(const float* __restrict__ st, int64_t stride, int f, int64_t i) {
    return st[int64_t(f) * stride + i];
}

template <int MAT>
NCLR_DI float hardening(float Jp) {
    if constexpr (MAT == MPM_SNOW) {
        return float(exp(10.0 * (1.0 - double(Jp))));
    } else if constexpr (MAT == MPM_JELLY) {
        return 0.3f;
    } else {
        return 1.0f;
    }
}

// first_piola_kirchoff_stress (src/nclr.h:313-337): -(dt V0)(4 inv_dx^2 PF) + m C with
// PF = 2 mu (F - R) F^T + lambda (J - 1) J on every entry.
template <int D, int MAT>
NCLR_DI M<D> fused_affine(const M<D>& F, const M<D>& C, float Jp, float mass, float vol, const StepParams& p) {
    const float e = hardening<MAT>(Jp);
    const float mu = p.mu0 * e;
    const float lambda = p.lambda0 * e;
    const float J = det(F);
    const M<D> R = polar_rotation(F);
    const float Dinv = 4.0f * p.inv_dx * p.inv_dx;
    const float two_mu = 2.0f * mu;
    M<D> FmR;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) FmR.a[a][b] = two_mu * (F.a[a][b] - R.a[a][b]);
    const M<D> PF = matmul_bt<D>(FmR, F);
    const float lam_term = lambda * (J - 1.0f) * J;
    const float neg_dtv = -(p.dt * vol);
    M<D> A;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) A.a[a][b] = neg_dtv * (Dinv * (PF.a[a][b] + lam_term)) + mass * C.a[a][b];
    return A;
}

template <int D>
NCLR_DI void load_particle(const float* __restrict__ st, int64_t stride, int64_t i, float (&x)[D], float (&v)[D],
                           M<D>& F, M<D>& C) {
    using L = Layout<D>;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        x[a] = ld<D>(st, stride, L::X + a, i);
        v[a] = ld<D>(st, stride, L::V + a, i);
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            F.a[a][b] = ld<D>(st, stride, L::F + a * D + b, i);
            C.a[a][b] = ld<D>(st, stride, L::C + a * D + b, i);
        }
}

template <int D>
NCLR_DI void mark_blocks(const int (&base)[D], const StepParams& p, uint32_t* __restrict__ stamps) {
    constexpr int LB = BlockShape<D>::LOG;
    int lo[D], hi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        lo[a] = base[a] >> LB;
        hi[a] = (base[a] + 2) >> LB;
    }
    if constexpr (D == 2) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int b0 = (c & 1) ? hi[0] : lo[0], b1 = (c & 2) ? hi[1] : lo[1];
            stamps[int64_t(b0) * p.nb + b1] = p.stamp;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int b0 = (c & 1) ? hi[0] : lo[0], b1 = (c & 2) ? hi[1] : lo[1], b2 = (c & 4) ? hi[2] : lo[2];
            stamps[(int64_t(b0) * p.nb + b1) * p.nb + b2] = p.stamp;
        }
    }
}

// P2G, one thread per particle, 3^D global float atomics per channel
// (src/nclr.h:104-165 with compute_fused_momentum 160-165).
template <int D, int MAT>
__global__ __launch_bounds__(256) void k_p2g_atomic(const float* __restrict__ st, StepParams p,
                                                    float4* __restrict__ acc, uint32_t* __restrict__ stamps,
                                                    uint32_t* __restrict__ status) {
    using L = Layout<D>;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    float x[D], v[D];
    M<D> F, C;
    load_particle<D>(st, p.stride, i, x, v, F, C);
    int base[D];
    float fx[D], w[D][3];
    if (!stencil<D>(x, p, base, fx, w)) {
        atomicOr(status, kStatusOutOfRange);
        return;
    }
    const float Jp = ld<D>(st, p.stride, L::JP, i);
    const float mass = ld<D>(st, p.stride, L::MASS, i);
    const float vol = ld<D>(st, p.stride, L::VOL, i);
    const M<D> A = fused_affine<D, MAT>(F, C, Jp, mass, vol, p);
    float mv[D];
#pragma unroll
    for (int a = 0; a < D; ++a) mv[a] = v[a] * mass;
    mark_blocks<D>(base, p, stamps);
    constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
    for (int o = 0; o < K; ++o) {
        int off[D];
        if constexpr (D == 2) {
            off[0] = o / 3;
            off[1] = o % 3;
        } else {
            off[0] = o / 9;
            off[1] = (o / 3) % 3;
            off[2] = o % 3;
        }
        float dpos[D];
        int nidx[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            dpos[a] = (float(off[a]) - fx[a]) * p.dx;
            nidx[a] = base[a] + off[a];
        }
        float weight = w[0][off[0]];
#pragma unroll
        for (int a = 1; a < D; ++a) weight = weight * w[a][off[a]];
        float* node = reinterpret_cast<float*>(acc + node_index<D>(nidx, p.nb));
#pragma unroll
        for (int r = 0; r < D; ++r) {
            float ad = A.a[r][0] * dpos[0];
#pragma unroll
            for (int c = 1; c < D; ++c) ad = ad + A.a[r][c] * dpos[c];
            unsafeAtomicAdd(node + r, weight * (mv[r] + ad));
        }
        unsafeAtomicAdd(node + 3, weight * mass);
    }
}

// grid_op (src/nclr.h:263-310) over the node blocks stamped by this step's P2G.  One wave per
// block; reads `acc`, writes the updated node to `gv` and clears `acc` for the next step.
template <int D>
__global__ __launch_bounds__(256) void k_grid_update(float4* __restrict__ acc, float4* __restrict__ gv,
                                                     const uint32_t* __restrict__ stamps, StepParams p,
                                                     int64_t nblocks, uint32_t* __restrict__ active_count) {
    const int64_t blk = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) active_count[(p.stamp + 1) & 1] = 0;
    if (blk >= nblocks) return;
    if (stamps[blk] != p.stamp) return;
    const int64_t node = blk * 64 + lane;
    const float4 g = acc[node];
    acc[node] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int n[D];
    node_coords<D>(blk, lane, p.nb, n);
    float vel[D];
    vel[0] = g.x;
    vel[1] = g.y;
    if constexpr (D == 3) vel[2] = g.z;
    float m = g.w;
    const uint64_t massive = __ballot(m > 0.0f);
    if (lane == 0 && massive) atomicAdd(active_count + (p.stamp & 1), uint32_t(__popcll(massive)));
    if (m > 0.0f) {
#pragma unroll
        for (int a = 0; a < D; ++a) vel[a] = vel[a] / m;
        vel[1] = vel[1] + p.dt * p.gravity;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const float lo = -p.allowed, hi = p.allowed;
            vel[a] = (vel[a] < lo) ? lo : ((hi < vel[a]) ? hi : vel[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        if ((n[a] < 3 && vel[a] < 0.0f) || (n[a] >= (p.res + 1) - 3 && vel[a] > 0.0f)) {
#pragma unroll
            for (int b = 0; b < D; ++b) vel[b] = 0.0f;
            m = 0.0f;
        }
    }
    gv[node] = make_float4(vel[0], vel[1], D == 3 ? vel[2] : 0.0f, m);
}

// G2P (src/nclr.h:167-261): gather v and the APIC C, advect, update F with the material law.
template <int D, int MAT>
__global__ __launch_bounds__(256) void k_g2p(float* __restrict__ st, StepParams p, const float4* __restrict__ gv,
                                             uint32_t* __restrict__ status) {
    using L = Layout<D>;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    float x[D];
    M<D> F;
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = ld<D>(st, p.stride, L::X + a, i);
    int base[D];
    float fx[D], w[D][3];
    if (!stencil<D>(x, p, base, fx, w)) {
        atomicOr(status, kStatusOutOfRange);
        return;
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) F.a[a][b] = ld<D>(st, p.stride, L::F + a * D + b, i);
    float v[D];
    M<D> C;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        v[a] = 0.0f;
#pragma unroll
        for (int b = 0; b < D; ++b) C.a[a][b] = 0.0f;
    }
    const float four_inv_dx = 4.0f * p.inv_dx;
    constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
    for (int o = 0; o < K; ++o) {
        int off[D];
        if constexpr (D == 2) {
            off[0] = o / 3;
            off[1] = o % 3;
        } else {
            off[0] = o / 9;
            off[1] = (o / 3) % 3;
            off[2] = o % 3;
        }
        float dpos[D];
        int nidx[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            dpos[a] = float(off[a]) - fx[a];
            nidx[a] = base[a] + off[a];
        }
        const float4 g = gv[node_index<D>(nidx, p.nb)];
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
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = x[a] + p.dt * v[a];
    M<D> Mx;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            const float id = (a == b && (a < 2 || p.identity_mode == 1)) ? 1.0f : 0.0f;
            Mx.a[a][b] = id + p.dt * C.a[a][b];
        }
    M<D> Fh = matmul<D>(Mx, F);
    if constexpr (MAT == MPM_JELLY) {
        F = Fh;
    } else {
        M<D> U, V;
        float sig[D];
        nclr_svd<D>(Fh, U, sig, V);
        if constexpr (MAT == MPM_SNOW) {
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
            float Jp = ld<D>(st, p.stride, L::JP, i);
            Jp = Jp * old_J / det(Fn);
            Jp = (Jp < 0.6f) ? 0.6f : ((20.0f < Jp) ? 20.0f : Jp);
            st[int64_t(L::JP) * p.stride + i] = Jp;
            F = Fn;
        } else {
            double J = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) J *= double(sig[d]);
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) F.a[a][b] = (a == b && (a < 2 || p.identity_mode == 1)) ? 1.0f : 0.0f;
            F.a[0][0] = float(J);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        st[int64_t(L::X + a) * p.stride + i] = x[a];
        st[int64_t(L::V + a) * p.stride + i] = v[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            st[int64_t(L::C + a * D + b) * p.stride + i] = C.a[a][b];
            st[int64_t(L::F + a * D + b) * p.stride + i] = F.a[a][b];
        }
}

// Dense export of the updated grid in the reference's flat x-major order (src/nclr.h:87):
// nodes of blocks not stamped by the last step are empty cells (zero).
template <int D>
__global__ __launch_bounds__(256) void k_export_grid(const float4* __restrict__ gv,
                                                     const uint32_t* __restrict__ stamps, StepParams p,
                                                     float* __restrict__ vel, float* __restrict__ mass) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t side = p.res + 1;
    int64_t total = side * side;
    if constexpr (D == 3) total *= side;
    if (idx >= total) return;
    int n[D];
    int64_t r = idx;
#pragma unroll
    for (int a = D - 1; a >= 0; --a) {
        n[a] = int(r % side);
        r /= side;
    }
    const int64_t node = node_index<D>(n, p.nb);
    const bool live = p.stamp != 0 && stamps[node >> 6] == p.stamp;
    const float4 g = live ? gv[node] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    vel[idx * D + 0] = g.x;
    vel[idx * D + 1] = g.y;
    if constexpr (D == 3) vel[idx * D + 2] = g.z;
    mass[idx] = g.w;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t value) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void k_fill_f4(float4* __restrict__ p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace nclr_dev

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace nclr_host {

using namespace nclr_dev;

static inline unsigned grid1d(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

template <int D>
static hipError_t p2g_dispatch(const LaunchArgs& a) {
    const unsigned blocks = grid1d(a.p.n, 256);
    if (blocks == 0) return hipSuccess;
    switch (a.material) {
        case MPM_SNOW:
            k_p2g_atomic<D, MPM_SNOW><<<blocks, 256, 0, a.stream>>>(a.state, a.p, a.acc, a.stamps, a.status);
            break;
        case MPM_JELLY:
            k_p2g_atomic<D, MPM_JELLY><<<blocks, 256, 0, a.stream>>>(a.state, a.p, a.acc, a.stamps, a.status);
            break;
        default:
            k_p2g_atomic<D, MPM_LIQUID><<<blocks, 256, 0, a.stream>>>(a.state, a.p, a.acc, a.stamps, a.status);
            break;
    }
    return hipGetLastError();
}

template <int D>
static hipError_t g2p_dispatch(const LaunchArgs& a) {
    const unsigned blocks = grid1d(a.p.n, 256);
    if (blocks == 0) return hipSuccess;
    switch (a.material) {
        case MPM_SNOW:
            k_g2p<D, MPM_SNOW><<<blocks, 256, 0, a.stream>>>(a.state, a.p, a.gv, a.status);
            break;
        case MPM_JELLY:
            k_g2p<D, MPM_JELLY><<<blocks, 256, 0, a.stream>>>(a.state, a.p, a.gv, a.status);
            break;
        default:
            k_g2p<D, MPM_LIQUID><<<blocks, 256, 0, a.stream>>>(a.state, a.p, a.gv, a.status);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_p2g(const LaunchArgs& a) { return a.dim == 2 ? p2g_dispatch<2>(a) : p2g_dispatch<3>(a); }

hipError_t launch_grid(const LaunchArgs& a) {
    const unsigned blocks = grid1d(a.nblocks, 4);
    if (a.dim == 2)
        k_grid_update<2><<<blocks, 256, 0, a.stream>>>(a.acc, a.gv, a.stamps, a.p, a.nblocks, a.active_count);
    else
        k_grid_update<3><<<blocks, 256, 0, a.stream>>>(a.acc, a.gv, a.stamps, a.p, a.nblocks, a.active_count);
    return hipGetLastError();
}

hipError_t launch_g2p(const LaunchArgs& a) { return a.dim == 2 ? g2p_dispatch<2>(a) : g2p_dispatch<3>(a); }

hipError_t launch_export_grid(const LaunchArgs& a, float* vel, float* mass, int64_t total) {
    const unsigned blocks = grid1d(total, 256);
    if (a.dim == 2)
        k_export_grid<2><<<blocks, 256, 0, a.stream>>>(a.gv, a.stamps, a.p, vel, mass);
    else
        k_export_grid<3><<<blocks, 256, 0, a.stream>>>(a.gv, a.stamps, a.p, vel, mass);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<grid1d(n, 256), 256, 0, s>>>(p, n, value);
    return hipGetLastError();
}

hipError_t launch_zero_f4(float4* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_f4<<<grid1d(n, 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

int num_fields(int dim) { return dim == 2 ? Layout<2>::NF : Layout<3>::NF; }

}  // namespace nclr_host
