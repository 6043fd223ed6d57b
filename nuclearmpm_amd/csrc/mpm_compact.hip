// The compact step: many steps of one small scene in one cooperative launch.
//
// A scene whose particles stay inside a box of a few thousand grid nodes does not need the tile
// sort: every workgroup takes a fixed range of particles (in slot order), deposits them into a
// fixed-point copy of the box's nodes in its own LDS, and stores that copy as float4 {momentum,
// mass} in its slot; after a grid-wide barrier every node sums the workgroups' slots in
// workgroup order and applies grid_op (src/nclr.h:263-310); after a second barrier every
// workgroup runs G2P (src/nclr.h:312-360) on its particles in place.  Two barriers per step, one
// launch for all steps.
//
// The box of step s: the stencil bases of the particles before step s lie within one cell of
// those before step s - 1 (a particle moves less than one cell per step: the grid velocity is
// clamped to 0.9 dx / dt per component, src/nclr.h:285, and G2P's weights are non-negative and
// sum to one), so the box is the previous positions' base range widened by one cell (the exact
// range for the first step).  G2P accumulates the range of the positions it writes, for the step
// after next.  When the box outgrows the LDS, every workgroup stops at the same step (they all read
// the same range) and the host continues with the tiled step; when the whole grid fits the LDS a
// particle outside the box (never expected) retries its step over the whole grid instead.  A
// particle off its grid stops the scene as the tiled step does: the step that finds it leaves
// every particle as it was, the grid keeps the last completed step and the stamp still counts
// every step of the launch.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <climits>

#include "mpm_device.hpp"
#include "mpm_step_common.hpp"

namespace nclr_dev {

// Four fixed-point channel planes per node (momentum x, y, z, mass; z unused in 2D, as in the
// tiles) and a flag word: 36 bytes per node, so a box holds up to 4266 nodes (65^2 in 2D, 16^3
// in 3D).
template <int D>
struct Compact {
    static constexpr int CH = 4;
    static constexpr int BYTES_PER_NODE = CH * 8 + 4;
    static constexpr int CAP = kCompactLdsBytes / BYTES_PER_NODE;
};

// Box of the stencil bases, stored as [min[D], max[D]] (atomicMin / atomicMax).
NCLR_DI void box_reset(int* b, int d) {
    for (int a = 0; a < d; ++a) {
        b[a] = INT_MAX;
        b[d + a] = INT_MIN;
    }
}

template <int D>
NCLR_DI void box_accumulate(int* __restrict__ gbox, int* __restrict__ sbox, const int (&lo)[D], const int (&hi)[D]) {
    // wave reduction, then one LDS atomic per wave, then one global atomic per workgroup
    int l[D], h[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        l[a] = lo[a];
        h[a] = hi[a];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            l[a] = min(l[a], __shfl_xor(l[a], off));
            h[a] = max(h[a], __shfl_xor(h[a], off));
        }
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int a = 0; a < D; ++a) {
            atomicMin(&sbox[a], l[a]);
            atomicMax(&sbox[D + a], h[a]);
        }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < D; ++a) {
            atomicMin(&gbox[a], sbox[a]);
            atomicMax(&gbox[D + a], sbox[D + a]);
        }
}

// Grid-wide barrier; the fences make every workgroup's global writes before it visible to every
// workgroup after it (the XCDs' L2 caches are not coherent with each other).
NCLR_DI void grid_barrier(cooperative_groups::grid_group& grid) {
    __threadfence();
    grid.sync();
    __threadfence();
}

__global__ void k_compact_init(int* __restrict__ box, int* __restrict__ info, int dim) {
    const int t = int(threadIdx.x);
    if (t < kInfoWords) info[t] = 0;
    if (t < dim) {
        box[t] = INT_MAX;
        box[dim + t] = INT_MIN;
    }
}

template <int D, int MAT, int MODE, int PPT>
__global__ __launch_bounds__(256) void k_compact_steps(CompactArgs a) {
    using L = Layout<D>;
    using CP = Compact<D>;
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    extern __shared__ unsigned long long lds[];
    __shared__ float s_bound[4][4];
    __shared__ int s_box[2 * D];
    const SceneParams sp = a.sp;
    const int material = MAT == kMixed ? sp.material : MAT;
    const int nwg = int(gridDim.x), wg = int(blockIdx.x);
    const int64_t s = a.stride;
    const int64_t p0 = int64_t(wg) * 256 * PPT;
    float* st = a.st;
    int64_t whole_nodes = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) whole_nodes *= sp.res + 1;
    const bool whole_fits = whole_nodes <= CP::CAP;

    // the exact base range of the particles before the first step
    {
        if (threadIdx.x == 0) box_reset(s_box, D);
        __syncthreads();
        int lo[D], hi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            lo[d] = INT_MAX;
            hi[d] = INT_MIN;
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int64_t i = p0 + int64_t(k) * 256 + threadIdx.x;
            if (i >= a.n) continue;
            float x[D], fx[D], w[D][3];
            int base[D];
#pragma unroll
            for (int d = 0; d < D; ++d) x[d] = ldf(st, s, L::X + d, i);
            if (!stencil<D>(x, sp, base, fx, w)) continue;  // found again by the first step
#pragma unroll
            for (int d = 0; d < D; ++d) {
                lo[d] = min(lo[d], base[d]);
                hi[d] = max(hi[d], base[d]);
            }
        }
        box_accumulate<D>(a.box, s_box, lo, hi);
    }
    grid_barrier(grid);

    int done = 0, reason = kCompactDone;
    bool whole = false;
    int lo[D], ext[D];
    int glo[D], gext[D], gnodes = 0;  // box of the last completed step's grid
    uint32_t gstamp = 0;               // its stamp (workgroup 0, thread 0)
#pragma unroll
    for (int d = 0; d < D; ++d) {
        glo[d] = 0;
        gext[d] = 1;
    }
    for (int step = 0; step < a.nsteps;) {
        // ---- the box of this step ----------------------------------------------------------
        int nodes = 1;
        if (whole) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                lo[d] = 0;
                ext[d] = sp.res + 1;
                nodes *= ext[d];
            }
        } else {
            const int* b = a.box + (step == 0 ? 0 : ((step - 1) & 1)) * 2 * D;
            const int pad = step == 0 ? 0 : 1;
            bool empty = false;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int bmin = b[d], bmax = b[D + d];
                empty = empty || bmin > bmax;
                lo[d] = max(0, bmin - pad);
                const int hi = min(sp.res, bmax + 2 + pad);
                ext[d] = max(1, hi - lo[d] + 1);
                nodes *= ext[d];
            }
            if (empty) {  // every particle off its grid: found below as out of range
                nodes = 1;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    lo[d] = 0;
                    ext[d] = 1;
                }
            }
        }
        if (nodes > CP::CAP) {  // the same decision in every workgroup
            reason = kCompactOverflow;
            break;
        }
        RegionStrides rs;
        rs.chan = nodes;
        rs.s0 = D == 3 ? ext[1] * ext[2] : ext[1];
        rs.s1 = D == 3 ? ext[2] : 1;
        unsigned long long* acc = lds;
        uint32_t* nonfinite = reinterpret_cast<uint32_t*>(lds + CP::CH * nodes);

        // ---- P2G into this workgroup's box copy ----------------------------------------------
        for (int r = threadIdx.x; r < CP::CH * nodes; r += 256) acc[r] = 0ull;
        for (int r = threadIdx.x; r < nodes; r += 256) nonfinite[r] = 0u;
        bool valid[PPT], finite[PPT];
        int rb[PPT];
        float fxs[PPT][D], mvs[PPT][D], ms[PPT];
        M<D> As[PPT];
        float bound[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool off_grid = false, outside = false;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int64_t i = p0 + int64_t(k) * 256 + threadIdx.x;
            valid[k] = false;
            finite[k] = true;
            rb[k] = 0;
            ms[k] = 0.0f;
            if (i >= a.n) continue;
            float x[D], v[D], fx[D], w[D][3];
            M<D> F, C;
            int base[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                x[d] = ldf(st, s, L::X + d, i);
                v[d] = ldf(st, s, L::V + d, i);
            }
            if (!stencil<D>(x, sp, base, fx, w)) {
                off_grid = true;
                continue;
            }
            bool inside = true;
            int r = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int q = base[d] - lo[d];
                inside = inside && q >= 0 && q + 2 < ext[d];
                r = r * ext[d] + q;
            }
            if (!inside) {
                outside = true;
                continue;
            }
#pragma unroll
            for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    F.a[r2][c] = ldf(st, s, L::F + r2 * D + c, i);
                    C.a[r2][c] = ldf(st, s, L::C + r2 * D + c, i);
                }
            const float Jp = ldf(st, s, L::JP, i);
            const float mass = ldf(st, s, L::MASS, i);
            const float vol = ldf(st, s, L::VOL, i);
            valid[k] = true;
            rb[k] = r;
            As[k] = fused_affine<D, MODE>(material, F, C, Jp, mass, vol, sp);
            ms[k] = mass;
            float wsum = 1.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                fxs[k][d] = fx[d];
                mvs[k][d] = v[d] * mass;
                wsum *= fabsf(w[d][0]) + fabsf(w[d][1]) + fabsf(w[d][2]);
            }
#pragma unroll
            for (int c = 0; c < D; ++c) {
                float sum = 0.0f;
#pragma unroll
                for (int b = 0; b < D; ++b) sum += fabsf(As[k].a[c][b]);
                const float bc = wsum * (fabsf(mvs[k][c]) + 2.5f * sp.dx * sum);
                finite[k] = finite[k] && bc < 3.0e38f;
                if (bc < 3.0e38f) bound[c] = fmaxf(bound[c], bc);
            }
            const float bm = wsum * fabsf(mass);
            finite[k] = finite[k] && bm < 3.0e38f;
            if (bm < 3.0e38f) bound[3] = fmaxf(bound[3], bm);
        }
        if (off_grid) stop_scene(a.step, 0, kStatusOutOfRange);
        if (outside) atomicOr(&a.info[kInfoOutside], 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float b = bound[c];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) b = fmaxf(b, __shfl_xor(b, off));
            bound[c] = b;
        }
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int c = 0; c < 4; ++c) s_bound[threadIdx.x >> 6][c] = bound[c];
        __syncthreads();  // also completes the zeroing
        float scale[4];
        double inv_scale[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            fixed_point_scale(fmaxf(fmaxf(s_bound[0][c], s_bound[1][c]), fmaxf(s_bound[2][c], s_bound[3][c])), scale[c],
                              inv_scale[c]);
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (!valid[k]) continue;
            if (finite[k]) {
                if constexpr (D == 3 && MODE == kSvdAccurate)
                    p2g_deposit_fused<false>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale, rs);
                else
                    p2g_deposit<D, false>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale, rs);
            } else if constexpr (D == 3 && MODE == kSvdAccurate) {
                p2g_deposit_fused<true>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale, rs);
            } else {
                p2g_deposit<D, true>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale, rs);
            }
        }
        __syncthreads();
        {
            float4* dst = a.slots + int64_t(wg) * CP::CAP;
            const float nan = __builtin_nanf("");
            for (int r = threadIdx.x; r < nodes; r += 256) {
                float out[4];
                const uint32_t bad = nonfinite[r];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    out[c] = (D == 2 && c == 2) ? 0.0f
                                                : ((bad >> c) & 1u ? nan : fixed_point_value(acc[c * nodes + r], inv_scale[c]));
                dst[r] = make_float4(out[0], out[1], out[2], out[3]);
            }
        }
        grid_barrier(grid);

        // ---- stop, retry or grid update over the box ------------------------------------------
        if ((a.step.scene_flags[0] | a.step.scene_frozen[0]) != 0u) {
            // a particle off its grid: the scene stops with every particle as before this step;
            // the stamp counts this and the remaining steps, as the tiled step's does
            if (wg == 0 && threadIdx.x == 0) {
                if (a.step.scene_frozen[0] == 0u) a.step.scene_frozen[0] = *a.stamp + 1u;
                *a.stamp = *a.stamp + uint32_t(a.nsteps - step);
            }
            reason = kCompactStopped;
            break;
        }
        if (!whole && a.info[kInfoOutside] != 0) {
            // a particle outside the box (the word stays raised: every workgroup decides alike)
            if (!whole_fits) {
                reason = kCompactOverflow;
                break;
            }
            whole = true;
            continue;
        }
        if (wg == 0 && threadIdx.x == 0) {
            box_reset(a.box + ((step + 1) & 1) * 2 * D, D);
            gstamp = *a.stamp + 1u;
            *a.stamp = gstamp;
        }
        for (int r = wg * 256 + int(threadIdx.x); r < nodes; r += nwg * 256) {
            float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float4* q = a.slots + r;
            int wq = 0;
            for (; wq + 4 <= nwg; wq += 4, q += 4 * CP::CAP) {  // four loads in flight, summed in order
                const float4 s0 = q[0], s1 = q[CP::CAP], s2 = q[2 * CP::CAP], s3 = q[3 * CP::CAP];
                g.x += s0.x;
                g.y += s0.y;
                g.z += s0.z;
                g.w += s0.w;
                g.x += s1.x;
                g.y += s1.y;
                g.z += s1.z;
                g.w += s1.w;
                g.x += s2.x;
                g.y += s2.y;
                g.z += s2.z;
                g.w += s2.w;
                g.x += s3.x;
                g.y += s3.y;
                g.z += s3.z;
                g.w += s3.w;
            }
            for (; wq < nwg; ++wq, q += CP::CAP) {
                const float4 t = q[0];
                g.x += t.x;
                g.y += t.y;
                g.z += t.z;
                g.w += t.w;
            }
            int n[D];
            {
                int rem = r;
#pragma unroll
                for (int d = D - 1; d >= 0; --d) {
                    n[d] = lo[d] + rem % ext[d];
                    rem /= ext[d];
                }
            }
            a.massive[r] = g.w > 0.0f ? uint8_t(1) : uint8_t(0);
            a.grid[r] = grid_node<D>(g, n, sp);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            glo[d] = lo[d];
            gext[d] = ext[d];
        }
        gnodes = nodes;
        grid_barrier(grid);

        // ---- G2P in place, and the base range of the new positions --------------------------
        {
            if (threadIdx.x == 0) box_reset(s_box, D);
            __syncthreads();
            int blo[D], bhi[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                blo[d] = INT_MAX;
                bhi[d] = INT_MIN;
            }
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int64_t i = p0 + int64_t(k) * 256 + threadIdx.x;
                if (i >= a.n || !valid[k]) continue;  // every particle of the scene is valid here
                float x[D], fx[D], w[D][3];
                int base[D];
                M<D> F;
#pragma unroll
                for (int d = 0; d < D; ++d) x[d] = ldf(st, s, L::X + d, i);
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) F.a[r][c] = ldf(st, s, L::F + r * D + c, i);
                float Jp = ldf(st, s, L::JP, i);
                (void)stencil<D>(x, sp, base, fx, w);  // on the grid: P2G found it there
                const float4* g0 = a.grid + rb[k];
                auto node = [&](const int (&off)[D]) {
                    const int o = D == 3 ? off[0] * rs.s0 + off[1] * rs.s1 + off[D - 1] : off[0] * rs.s0 + off[1];
                    return g0[o];
                };
                float v[D];
                M<D> C;
                g2p_particle<D, MODE>(material, sp, fx, w, node, x, v, C, F, Jp);
                st[int64_t(L::JP) * s + i] = Jp;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    st[int64_t(L::X + d) * s + i] = x[d];
                    st[int64_t(L::V + d) * s + i] = v[d];
                }
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) {
                        st[int64_t(L::C + r * D + c) * s + i] = C.a[r][c];
                        st[int64_t(L::F + r * D + c) * s + i] = F.a[r][c];
                    }
                int nb[D];
                float nfx[D], nw[D][3];
                if (stencil<D>(x, sp, nb, nfx, nw)) {
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        blo[d] = min(blo[d], nb[d]);
                        bhi[d] = max(bhi[d], nb[d]);
                    }
                }
            }
            box_accumulate<D>(a.box + ((step + 1) & 1) * 2 * D, s_box, blo, bhi);
        }
        ++step;
        done = step;
    }
    if (wg == 0 && threadIdx.x == 0) {
        a.info[kInfoDone] = done;
        a.info[kInfoReason] = reason;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            a.info[kInfoLo + d] = glo[d];
            a.info[kInfoExt + d] = gext[d];
        }
        a.info[kInfoNodes] = gnodes;
        a.info[kInfoStamp] = int(gstamp);
    }
}

// The compact step's last grid written into the tiled grid of a single-scene engine: one
// workgroup per node tile.  Node tiles that meet the box take its values (zero outside it), the
// stamp of its step and their count of nodes with mass; the others count no node with mass and
// keep an older stamp, so the grid export reads them as empty.  A launch that completed no step
// leaves the tiled grid as it was.
template <int D>
__global__ __launch_bounds__(256) void k_compact_to_tiles(const float4* __restrict__ box_grid,
                                                          const uint8_t* __restrict__ massive,
                                                          const int* __restrict__ info, int ntn,
                                                          uint32_t node_tile_base, uint32_t tiles,
                                                          float4* __restrict__ gv, uint32_t* __restrict__ stamps,
                                                          uint32_t* __restrict__ active) {
    using TL = Tile<D>;
    __shared__ uint32_t s_count;
    const uint32_t t = blockIdx.x;
    if (t >= tiles || info[kInfoDone] == 0) return;
    int to[D], lo[D], ext[D];
    tile_coords<D>(t, ntn, to);
    bool touch = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = info[kInfoLo + d];
        ext[d] = info[kInfoExt + d];
        touch = touch && to[d] * TL::S < lo[d] + ext[d] && lo[d] < (to[d] + 1) * TL::S;
    }
    const uint32_t nt = node_tile_base + t;
    if (!touch) {
        if (threadIdx.x == 0) active[nt] = 0u;
        return;
    }
    if (threadIdx.x == 0) s_count = 0u;
    __syncthreads();
    uint32_t count = 0;
    for (int l = threadIdx.x; l < TL::NODES; l += 256) {
        bool inside = true;
        int r = 0;
        {
            int rem = l;
            int q[D];
#pragma unroll
            for (int d = D - 1; d >= 0; --d) {
                q[d] = to[d] * TL::S + (rem & (TL::S - 1)) - lo[d];
                rem >>= TL::LOG;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                inside = inside && q[d] >= 0 && q[d] < ext[d];
                r = r * ext[d] + q[d];
            }
        }
        const float4 g = inside ? box_grid[r] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        count += (inside && massive[r]) ? 1u : 0u;
        gv[int64_t(nt) * TL::NODES + l] = g;
    }
    if (count) atomicAdd(&s_count, count);
    __syncthreads();
    if (threadIdx.x == 0) {
        active[nt] = s_count;
        stamps[nt] = uint32_t(info[kInfoStamp]);
    }
}

}  // namespace nclr_dev

namespace nclr_host {

using namespace nclr_dev;

int compact_capacity(int dim) { return dim == 2 ? Compact<2>::CAP : Compact<3>::CAP; }

template <int D, int MAT, int MODE, int PPT>
static hipError_t compact_one(const CompactArgs& a, int nwg, hipStream_t s) {
    auto kern = k_compact_steps<D, MAT, MODE, PPT>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       kCompactLdsBytes);
    if (e != hipSuccess) return e;
    k_compact_init<<<1, 64, 0, s>>>(a.box, a.info, D);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    CompactArgs args = a;
    void* params[] = {&args};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(unsigned(nwg)), dim3(256), params,
                                      unsigned(kCompactLdsBytes), s);
}

template <int D, int MAT, int MODE>
static hipError_t compact_ppt(const CompactArgs& a, int ppt, int nwg, hipStream_t s) {
    // 3D: 1 or 2 particles per thread (4 would spill)
    if (ppt == 1) return compact_one<D, MAT, MODE, 1>(a, nwg, s);
    if constexpr (D == 2)
        if (ppt == 4) return compact_one<D, MAT, MODE, 4>(a, nwg, s);
    return compact_one<D, MAT, MODE, 2>(a, nwg, s);
}

template <int D, int MAT>
static hipError_t compact_mode(const CompactArgs& a, int mode, int ppt, int nwg, hipStream_t s) {
    if (D == 3 && mode == kSvdFast) return compact_ppt<D, MAT, kSvdFast>(a, ppt, nwg, s);
    if (D == 3 && mode == kSvdAccurate) return compact_ppt<D, MAT, kSvdAccurate>(a, ppt, nwg, s);
    return compact_ppt<D, MAT, kSvdExact>(a, ppt, nwg, s);
}

template <int D>
static hipError_t compact_dim(const CompactArgs& a, int material, int mode, int ppt, int nwg, hipStream_t s) {
    switch (material) {
        case MPM_SNOW: return compact_mode<D, MPM_SNOW>(a, mode, ppt, nwg, s);
        case MPM_JELLY: return compact_mode<D, MPM_JELLY>(a, mode, ppt, nwg, s);
        default: return compact_mode<D, MPM_LIQUID>(a, mode, ppt, nwg, s);
    }
}

hipError_t launch_compact_steps(int dim, int material, int mode, int ppt, int nwg, const CompactArgs& a, hipStream_t s) {
    const bool ppt_ok = ppt == 1 || ppt == 2 || (dim == 2 && ppt == 4);
    if (!ppt_ok || nwg < 1 || nwg > kCompactMaxWorkgroups || int64_t(nwg) * 256 * ppt < a.n || a.n <= 0)
        return hipErrorInvalidValue;
    return dim == 2 ? compact_dim<2>(a, material, mode, ppt, nwg, s) : compact_dim<3>(a, material, mode, ppt, nwg, s);
}

hipError_t launch_compact_to_tiles(int dim, const CompactArgs& a, float4* gv, uint32_t* stamps, uint32_t* active,
                                   hipStream_t s) {
    const int ntn = a.sp.ntn;
    uint32_t tiles = uint32_t(ntn) * uint32_t(ntn);
    if (dim == 3) tiles *= uint32_t(ntn);
    if (dim == 2)
        k_compact_to_tiles<2><<<tiles, 256, 0, s>>>(a.grid, a.massive, a.info, ntn, a.sp.node_tile_base, tiles, gv,
                                                    stamps, active);
    else
        k_compact_to_tiles<3><<<tiles, 256, 0, s>>>(a.grid, a.massive, a.info, ntn, a.sp.node_tile_base, tiles, gv,
                                                    stamps, active);
    return hipGetLastError();
}

}  // namespace nclr_host
