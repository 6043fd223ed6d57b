// Device helpers shared by the step kernels (mpm_step.hip) and their adjoints (mpm_adjoint.hip):
// the particle field layout, scene lookup, the interpolation stencil, hardening, the fused
// affine matrix of P2G and the tile / node-tile index maps.
#pragma once

#include <hip/hip_runtime.h>

#include "mpm_internal.hpp"
#include "mpm_math.hpp"

namespace nclr_dev {

// Particle fields: x[D] v[D] F[DxD] C[DxD] Jp mass volume, field f of slot i at st[f*stride + i].
// The first NG = 2D + 2D^2 + 1 fields (x, v, F, C, Jp) are the channels of a state image and of a
// particle-state cotangent; mass and volume are constants of the particle.
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
    static constexpr int NG = JP + 1;
};

// Largest scene index s with key(s) <= v (scenes ordered by every base).
template <typename KeyFn>
NCLR_DI int scene_search(int nscenes, int64_t v, KeyFn key) {
    int lo = 0, hi = nscenes;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (key(mid) <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

NCLR_DI int scene_of_particle(const StepParams& p, uint32_t pid) {
    return scene_search(p.nscenes, pid, [&](int s) { return p.scenes[s].particle_base; });
}
// Tiles and node tiles look their scene up in the static maps built at creation (one load
// instead of a search over the scene table).
NCLR_DI int scene_of_tile(const StepParams& p, uint32_t tile) { return int(p.tile_scene[tile]); }
NCLR_DI int scene_of_node_tile(const StepParams& p, uint32_t nt) { return int(p.node_tile_scene[nt]); }
// The particle at sorted position j: the scene of its tile (keys[j], this step's sort), or for a
// particle off its grid (key ntiles) a search by its original index `id`.
NCLR_DI int scene_of_sorted(const StepParams& p, const uint32_t* __restrict__ keys, int64_t j, uint32_t id) {
    const uint32_t key = keys[j];
    return key < p.ntiles ? int(p.tile_scene[key]) : scene_of_particle(p, id);
}

// A stopped scene (src/nclr.h:163-164: std::vector::at throws out of that scene's advance(), which
// leaves its particles as they were before the step) takes no further part in a step.
NCLR_DI bool scene_stopped(const StepParams& p, int s) { return (p.scene_flags[s] | p.scene_frozen[s]) != 0u; }
NCLR_DI void stop_scene(const StepParams& p, int s, uint32_t bits) { atomicOr(&p.scene_flags[s], bits); }

// base = trunc(x*inv_dx - 0.5), fx = x*inv_dx - base (src/nclr.h:115,121), with the stencil
// nodes base..base+2 required to lie in [0, res] on every axis.
template <int D>
NCLR_DI bool stencil(const float (&x)[D], const SceneParams& sp, int (&base)[D], float (&fx)[D], float (&w)[D][3]) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const float s = x[a] * sp.inv_dx;
        const float b = s - 0.5f;
        ok = ok && (b > -1.0f) && (b < float(sp.res - 1));
        base[a] = ok ? int(b) : 0;
        fx[a] = s - float(base[a]);
        bspline_weights(fx[a], w[a]);
    }
    return ok;
}

NCLR_DI float ldf(const float* __restrict__ st, int64_t stride, int f, int64_t i) { return st[int64_t(f) * stride + i]; }

// hardening (src/nclr.h:351-372): snow e = exp(10 (1 - Jp)) evaluated in double
NCLR_DI float hardening(int material, float Jp) {
    if (material == MPM_SNOW) return float(exp(10.0 * (1.0 - double(Jp))));
    if (material == MPM_JELLY) return 0.3f;
    return 1.0f;
}

// first_piola_kirchoff_stress (src/nclr.h:313-337): -(dt V0)(4 inv_dx^2 PF) + m C with
// PF = 2 mu (F - R) F^T + lambda (J - 1) J on every entry.  MODE: the SVD mode of the polar factor
// (kSvdExact: the reference's float arithmetic bit for bit).
template <int D>
NCLR_DI M<D> fused_affine_from(int material, const M<D>& F, const M<D>& diff, const M<D>& C, float Jp, float mass,
                               float vol, const SceneParams& sp) {
    const float e = hardening(material, Jp);
    const float mu = sp.mu0 * e;
    const float lambda = sp.lambda0 * e;
    const float J = det(F);
    const float Dinv = 4.0f * sp.inv_dx * sp.inv_dx;
    const float two_mu = 2.0f * mu;
    M<D> FmR;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) FmR.a[a][b] = two_mu * diff.a[a][b];
    const M<D> PF = matmul_bt<D>(FmR, F);
    const float lam_term = lambda * (J - 1.0f) * J;
    const float neg_dtv = -(sp.dt * vol);
    M<D> A;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) A.a[a][b] = neg_dtv * (Dinv * (PF.a[a][b] + lam_term)) + mass * C.a[a][b];
    return A;
}
template <int D, int MODE>
NCLR_DI M<D> fused_affine(int material, const M<D>& F, const M<D>& C, float Jp, float mass, float vol,
                          const SceneParams& sp) {
    return fused_affine_from<D>(material, F, f_minus_rotation<MODE, D>(F), C, Jp, mass, vol, sp);
}

template <int D>
NCLR_DI void offsets(int o, int (&off)[D]) {
    if constexpr (D == 2) {
        off[0] = o / 3;
        off[1] = o % 3;
    } else {
        off[0] = o / 9;
        off[1] = (o / 3) % 3;
        off[2] = o % 3;
    }
}

// Index of node n of the scene whose node tiles start at node_tile_base in the updated grid.
template <int D>
NCLR_DI int64_t gv_index(const int (&n)[D], int ntn, uint32_t node_tile_base) {
    using TL = Tile<D>;
    int64_t t = 0;
    int local = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        t = t * ntn + (n[a] >> TL::LOG);
        local = local * TL::S + (n[a] & (TL::S - 1));
    }
    return (int64_t(node_tile_base) + t) * TL::NODES + local;
}

// v[i] for i in {0, 1, 2} by selects: no private-array indexing (which would live in scratch)
// when i is not a compile-time constant; folds to v[i] when it is.
template <typename T>
NCLR_DI T pick3(const T (&v)[3], int i) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// The 3^D stencil nodes base + off of one particle in the updated grid, addressed per axis: node
// n's byte offset inside its scene's node tiles is sum_a t_a ntn^(D-1-a) NODES 16 +
// l_a S^(D-1-a) 16 with t_a = n_a >> LOG, l_a = n_a & (S-1) (gv_index's layout), so rows[a][o]
// holds axis a's term for n_a = base[a] + o and a node costs one 32-bit add of D terms.  A scene's
// node tiles span less than 2^32 bytes (res <= 512 in 3D: 2.25e9; res <= 8192 in 2D: 1.08e9).
template <int D>
struct StencilRows {
    const char* scene;
    uint32_t rows[D][3];

    NCLR_DI StencilRows(const float4* __restrict__ g, const int (&base)[D], int ntn, uint32_t node_tile_base) {
        using TL = Tile<D>;
        scene = reinterpret_cast<const char*>(g + int64_t(node_tile_base) * TL::NODES);
        uint32_t tstride = TL::NODES * uint32_t(sizeof(float4)), lstride = uint32_t(sizeof(float4));
#pragma unroll
        for (int a = D - 1; a >= 0; --a) {
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                const uint32_t n = uint32_t(base[a] + o);
                rows[a][o] = (n >> TL::LOG) * tstride + (n & uint32_t(TL::S - 1)) * lstride;
            }
            tstride *= uint32_t(ntn);
            lstride *= uint32_t(TL::S);
        }
    }

    NCLR_DI float4 operator()(const int (&off)[D]) const {
        uint32_t b = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) b += pick3(rows[a], off[a]);
        return *reinterpret_cast<const float4*>(scene + b);
    }
};

template <int D>
NCLR_DI uint32_t tile_key(const int (&base)[D], int nt) {
    uint32_t t = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) t = t * uint32_t(nt) + uint32_t(base[a] >> Tile<D>::LOG);
    return t;
}

template <int D>
NCLR_DI void tile_coords(uint32_t t, int n, int (&c)[D]) {
#pragma unroll
    for (int a = D - 1; a >= 0; --a) {
        c[a] = int(t % uint32_t(n));
        t /= uint32_t(n);
    }
}

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin by workgroup id.
// Launches that read neighbouring data from neighbouring blocks round their grid up to a multiple
// of 8 and map workgroup b to logical block (b % 8) * (grid / 8) + b / 8, so that consecutive
// logical blocks share an XCD and its L2.  Logical blocks at or beyond the real count exit.
constexpr uint32_t kXcds = 8;
NCLR_DI uint32_t xcd_block() {
    const uint32_t per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
inline unsigned xcd_grid(int64_t blocks) { return unsigned((blocks + kXcds - 1) / kXcds * kXcds); }

// Work item -> its tile and sorted range [j0, j1), from the item table written with the item scan
// (one load, uniform over the workgroup).
struct WorkItem {
    uint32_t tile, j0, j1;
};
NCLR_DI WorkItem work_item(const StepParams& p, uint32_t item) {
    const uint4 w = p.items[item];
    return {w.x, w.y, w.z};
}

// Records items [first, first + count) of a tile whose particles are the sorted positions
// [start, start + particles): item k covers [start + k chunk, min(start + (k + 1) chunk, end)).
NCLR_DI void write_items(uint4* __restrict__ items, uint32_t first, uint32_t count, uint32_t tile, uint32_t start,
                         uint32_t particles, uint32_t chunk) {
    const uint32_t end = start + particles;
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t j0 = start + k * chunk;
        items[first + k] = make_uint4(tile, j0, min(j0 + chunk, end), 0u);
    }
}

// Scans over tiles, in rounds of kScanRound tiles by a workgroup of 1024 threads (4 consecutive
// tiles per thread, coalesced).  A tile's particle count c gives ceil(c / chunk) P2G work items.

// Workgroup-wide exclusive scans of two values per thread: ea / eb the thread's prefixes, ta / tb
// the workgroup's totals.
NCLR_DI void block_scan2(uint32_t a, uint32_t b, uint32_t& ea, uint32_t& eb, uint32_t& ta, uint32_t& tb) {
    __shared__ uint32_t wsum[2][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t xa = a, xb = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ya = __shfl_up(xa, off), yb = __shfl_up(xb, off);
        if (lane >= off) {
            xa += ya;
            xb += yb;
        }
    }
    if (lane == 63) {
        wsum[0][wave] = xa;
        wsum[1][wave] = xb;
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t wa = lane < 16 ? wsum[0][lane] : 0u, wb = lane < 16 ? wsum[1][lane] : 0u;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            const uint32_t ya = __shfl_up(wa, off), yb = __shfl_up(wb, off);
            if (lane >= off) {
                wa += ya;
                wb += yb;
            }
        }
        if (lane < 16) {
            wsum[0][lane] = wa;
            wsum[1][lane] = wb;
        }
    }
    __syncthreads();
    ea = (wave ? wsum[0][wave - 1] : 0u) + xa - a;
    eb = (wave ? wsum[1][wave - 1] : 0u) + xb - b;
    ta = wsum[0][15];
    tb = wsum[1][15];
    __syncthreads();  // wsum is reused by the next call
}

// Round [b0, b0 + kScanRound) of the tile scans with carries (cp, ci) in: visit(tile,
// first_particle, count, first_item, items) for each tile below ntiles; tp / ti: the round's totals.
template <typename Count, typename Visit>
NCLR_DI void scan_round(uint32_t b0, uint32_t ntiles, uint32_t chunk, uint32_t cp, uint32_t ci, Count count,
                        Visit visit, uint32_t& tp, uint32_t& ti) {
    uint32_t cnt[4], its[4], sp = 0, si = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t tile = b0 + uint32_t(threadIdx.x) * 4 + k;
        cnt[k] = tile < ntiles ? count(tile) : 0u;
        its[k] = (cnt[k] + chunk - 1) / chunk;
        sp += cnt[k];
        si += its[k];
    }
    uint32_t ep, ei;
    block_scan2(sp, si, ep, ei, tp, ti);
    uint32_t rp = cp + ep, ri = ci + ei;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t tile = b0 + uint32_t(threadIdx.x) * 4 + k;
        if (tile < ntiles) visit(tile, rp, cnt[k], ri, its[k]);
        rp += cnt[k];
        ri += its[k];
    }
}

// Totals of round `round` (the first pass of a scan with one workgroup per round).
template <typename Count>
NCLR_DI void round_totals(uint32_t round, uint32_t ntiles, uint32_t chunk, Count count, uint32_t& tp, uint32_t& ti) {
    uint32_t sp = 0, si = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t tile = round * kScanRound + uint32_t(threadIdx.x) * 4 + k;
        const uint32_t c = tile < ntiles ? count(tile) : 0u;
        sp += c;
        si += (c + chunk - 1) / chunk;
    }
    uint32_t ep, ei;
    block_scan2(sp, si, ep, ei, tp, ti);
}

// Carries into round `round`: the sums of the totals of rounds 0 .. round-1 (rtot[2 r], rtot[2 r + 1]).
NCLR_DI void round_carries(const uint32_t* __restrict__ rtot, uint32_t round, uint32_t& cp, uint32_t& ci) {
    uint32_t a = 0, b = 0;
    for (uint32_t r = threadIdx.x; r < round; r += blockDim.x) {
        a += rtot[2 * r];
        b += rtot[2 * r + 1];
    }
    uint32_t ea, eb;
    block_scan2(a, b, ea, eb, cp, ci);
}

// Exact LDS reductions in fixed point.  A workgroup bounds every particle's |contributions| per
// channel (each single contribution is at most that particle's bound) and scales the channel by
// 2^e with (largest particle bound) * 2^e < 2^50: every quantised contribution is below 2^50 in
// magnitude, so a node's sum over the at most 1024 particles of a work item stays below 2^60 and
// the 64-bit integer sum is exact and independent of the order of the atomics.  The quantum, 2^-50
// of the largest particle bound, is below the FP32 rounding of every node sum above 1e-8 of that
// bound.  A particle whose bound is not finite is deposited with a check per contribution:
// contributions at or beyond kFixedLimit flag their node channel (NaN) instead.
constexpr float kFixedLimit = 1.12589991e15f;  // 2^50

NCLR_DI void fixed_point_scale(float bound_max, float& scale, double& inv_scale) {
    int ex = 0;
    (void)frexpf(bound_max > 0.0f && bound_max < 3.0e38f ? bound_max : 1.0f, &ex);  // bound_max < 2^ex
    const int e = min(127, 50 - ex);                                                  // in [-78, 127]
    scale = ldexpf(1.0f, e);
    inv_scale = ldexp(1.0, -e);
}

// y (|y| < 2^51) rounded to the nearest integer (ties to even), as a two's complement 64-bit
// value: in double, y + 1.5 * 2^52 lies in [2^52, 2^53), where the unit in the last place is 1, so
// its significand holds round(y) + 2^51 and its bit pattern minus that of 1.5 * 2^52 is round(y).
// The constant's low word is zero: only the high word is adjusted.  Three instructions
// (convert, add, subtract); deterministic, which is all the exact sums need.
NCLR_DI unsigned long long fixed_point(float y) {
    const double t = double(y) + 6755399441055744.0;  // 1.5 * 2^52
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(t));
    const uint32_t lo = uint32_t(b), hi = uint32_t(b >> 32) - 0x43380000u;
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

NCLR_DI float fixed_point_value(unsigned long long q, double inv_scale) {
    return float(double((long long)q) * inv_scale);
}

}  // namespace nclr_dev
