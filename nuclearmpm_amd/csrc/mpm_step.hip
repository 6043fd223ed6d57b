// HIP kernels of the MLS-MPM step for gfx950 (MI355X).
//
// Reference step: MPMSimulation<dim>::advance() = p2g(); grid_op(); g2p(); (src/nclr.h:80-84).
// One launch sequence advances every scene of a batch (a single simulation is a batch of one).
//
// Device data layout
//   particles   structure-of-arrays, field f of slot i at st[f*stride + i] (Layout<D>), two
//               ping-pong buffers; pid[i] is the original particle index of slot i, scenes own
//               consecutive original indices.
//   tiles       particles are keyed by the tile of their stencil base cell (S^D cells, S = 16 in
//               2D, 8 in 3D), tiles numbered scene after scene, and counting-sorted every step
//               (mpm_sort.hip): perm maps sorted position -> slot, [tile_start, tile_end) per tile.
//   slots       per P2G work item (a chunk of one tile), its region of (S+2)^D nodes as float4
//               {momentum, mass}.
//   gv          the updated grid in node tiles of S^D nodes, float4 {v, m}; stamps[node tile] ==
//               *stamp marks the node tiles the current step touched.
//
// Kernels per step
//   sort   stable counting sort of the slots by tile: after the first step the incremental bucket
//          sort of mpm_bucket.hip (a particle's tile is a neighbour of its previous one); at
//          creation and after the state is replaced, keys (tile key per particle, the device step
//          stamp), the radix sort of mpm_sort.hip, tile bounds and P2G work items
//   p2g    one workgroup per work item: fused affine matrix per particle, quantised 3^D
//          stencil contributions summed exactly with 64-bit LDS integer atomics, region -> slot
//   grid   one workgroup per node tile: sums the slots covering each node, grid_op
//          (src/nclr.h:263-310)
//   g2p    per sorted position j: particle perm[j] -> slot j of the other buffer (the state stays
//          physically sorted by tile)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mpm_device.hpp"
#include "mpm_step_common.hpp"

namespace nclr_dev {

// ---------------------------------------------------------------------------------------------
// keys + tile bounds + work items
// ---------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_keys(const float* __restrict__ st, const uint32_t* __restrict__ pid,
                                              StepParams p, uint32_t* __restrict__ stamp,
                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ status,
                                              uint32_t* __restrict__ tile_start, uint32_t* __restrict__ tile_end) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *stamp = *stamp + 1u;
    if (i <= int64_t(p.ntiles)) {
        tile_start[i] = 0;
        tile_end[i] = 0;
    }
    if (i >= p.n) return;
    const int s = p.nscenes == 1 ? 0 : scene_of_particle(p, pid[i]);
    const SceneParams sp = p.scenes[s];
    float x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = ldf(st, p.stride, Layout<D>::X + a, i);
    int base[D];
    float fx[D], w[D][3];
    if (stencil<D>(x, sp, base, fx, w)) {
        keys[i] = sp.tile_base + tile_key<D>(base, sp.nt);
    } else {
        keys[i] = p.ntiles;
        stop_scene(p, s, kStatusOutOfRange);
    }
}

__global__ __launch_bounds__(256) void k_tile_bounds(const uint32_t* __restrict__ keys, int64_t n,
                                                     uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t k = keys[j];
    if (j == 0 || keys[j - 1] != k) start[k] = uint32_t(j);
    if (j == n - 1 || keys[j + 1] != k) end[k] = uint32_t(j + 1);
}

// P2G work items: ceil(count / chunk) per tile, exclusive scan into item_offset[0..ntiles]
// (item_offset[ntiles] = total), and meta[0] = the particles on their grid (sorted first), which
// the incremental sort of the next step reads.  Up to kScanRound tiles: k_items, one workgroup;
// more: k_items_round_totals and k_items_rounds, one workgroup per round (same results).
struct ItemsVisit {
    uint32_t* item_offset;
    uint4* items;
    uint32_t chunk;
    NCLR_DI void operator()(uint32_t tile, uint32_t first_p, uint32_t c, uint32_t first, uint32_t its) const {
        item_offset[tile] = first;
        write_items(items, first, its, tile, first_p, c, chunk);
    }
};

__global__ __launch_bounds__(1024) void k_items(const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                uint32_t ntiles, uint32_t chunk, uint32_t* __restrict__ item_offset,
                                                uint4* __restrict__ items, uint32_t* __restrict__ meta) {
    const ItemsVisit visit{item_offset, items, chunk};
    auto count = [&](uint32_t tile) { return end[tile] - start[tile]; };
    uint32_t cp = 0, ci = 0;
    for (uint32_t b0 = 0; b0 < ntiles; b0 += kScanRound) {
        uint32_t tp, ti;
        scan_round(b0, ntiles, chunk, cp, ci, count, visit, tp, ti);
        cp += tp;
        ci += ti;
    }
    if (threadIdx.x == 0) {
        item_offset[ntiles] = ci;
        meta[0] = cp;
    }
}

__global__ __launch_bounds__(1024) void k_items_round_totals(const uint32_t* __restrict__ start,
                                                             const uint32_t* __restrict__ end, uint32_t ntiles,
                                                             uint32_t chunk, uint32_t* __restrict__ rtot) {
    uint32_t tp, ti;
    round_totals(blockIdx.x, ntiles, chunk, [&](uint32_t tile) { return end[tile] - start[tile]; }, tp, ti);
    if (threadIdx.x == 0) {
        rtot[2 * blockIdx.x] = tp;
        rtot[2 * blockIdx.x + 1] = ti;
    }
}

__global__ __launch_bounds__(1024) void k_items_rounds(const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ end, uint32_t ntiles,
                                                       uint32_t chunk, uint32_t* __restrict__ item_offset,
                                                       uint4* __restrict__ items,
                                                       const uint32_t* __restrict__ rtot, uint32_t* __restrict__ meta) {
    uint32_t cp, ci, tp, ti;
    round_carries(rtot, blockIdx.x, cp, ci);
    scan_round(blockIdx.x * kScanRound, ntiles, chunk, cp, ci, [&](uint32_t tile) { return end[tile] - start[tile]; },
               ItemsVisit{item_offset, items, chunk}, tp, ti);
    if (threadIdx.x == 0 && blockIdx.x + 1 == gridDim.x) {
        item_offset[ntiles] = ci + ti;
        meta[0] = cp + tp;
    }
}

// ---------------------------------------------------------------------------------------------
// P2G (src/nclr.h:104-165)
// ---------------------------------------------------------------------------------------------

// A work item of TPB * PPT particles, PPT per thread.
template <int D, int MAT, int PPT, int MODE, int TPB = 256>
__global__ __launch_bounds__(TPB) void k_p2g_tile(const float* __restrict__ st, const uint32_t* __restrict__ perm,
                                                  const uint32_t* __restrict__ item_offset, StepParams p,
                                                  float4* __restrict__ slots) {
    using TL = Tile<D>;
    using L = Layout<D>;
    constexpr int REG = TL::REGION;
    constexpr int NW = TPB / 64;
    // LDS accumulators: exact fixed-point sums (64-bit integer atomics are an order of magnitude
    // faster than float atomics in LDS on gfx950, and make the sum independent of lane order)
    __shared__ unsigned long long acc[4 * REG];
    __shared__ uint32_t nonfinite[REG];  // bit c: a non-finite contribution hit channel c
    __shared__ float s_bound[NW][4];     // per wave, per channel
    const uint32_t item = blockIdx.x;
    const WorkItem wi = work_item(p, item);  // loaded with the count below, used only if valid
    if (item >= item_offset[p.ntiles]) return;
    const uint32_t tile = wi.tile;
    const int scene = p.nscenes == 1 ? 0 : scene_of_tile(p, tile);
    if (scene_stopped(p, scene)) return;  // uniform over the workgroup
    for (int r = threadIdx.x; r < 4 * REG; r += blockDim.x) acc[r] = 0ull;
    for (int r = threadIdx.x; r < REG; r += blockDim.x) nonfinite[r] = 0u;
    __syncthreads();
    const SceneParams sp = p.scenes[scene];
    const int material = MAT == kMixed ? sp.material : MAT;
    int torg[D];
    tile_coords<D>(tile - sp.tile_base, sp.nt, torg);
#pragma unroll
    for (int a = 0; a < D; ++a) torg[a] *= TL::S;
    const uint32_t j0 = wi.j0, j1 = wi.j1;

    // pass 1: per particle the fused affine matrix, kept in registers, and a bound of the sum of
    // |contributions| per channel: sum_nodes |w| (|m v_c| + 2.5 dx sum_k |A_ck|)
    bool valid[PPT], finite[PPT];
    int rb[PPT];
    float fxs[PPT][D], mvs[PPT][D], ms[PPT];
    M<D> As[PPT];
    float bound[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const uint32_t j = j0 + uint32_t(k) * TPB + threadIdx.x;
        valid[k] = false;
        finite[k] = true;
        rb[k] = 0;
        ms[k] = 0.0f;
        if (j >= j1) continue;
        const int64_t i = perm[j];
        float x[D], v[D];
        M<D> F, C;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            x[a] = ldf(st, p.stride, L::X + a, i);
            v[a] = ldf(st, p.stride, L::V + a, i);
        }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                F.a[a][b] = ldf(st, p.stride, L::F + a * D + b, i);
                C.a[a][b] = ldf(st, p.stride, L::C + a * D + b, i);
            }
        const float Jp = ldf(st, p.stride, L::JP, i);
        const float mass = ldf(st, p.stride, L::MASS, i);
        const float vol = ldf(st, p.stride, L::VOL, i);
        int base[D];
        float fx[D], w[D][3];
        if (!stencil<D>(x, sp, base, fx, w)) continue;  // keyed into this tile, so never taken
        valid[k] = true;
        As[k] = fused_affine<D, MODE>(material, F, C, Jp, mass, vol, sp);
        ms[k] = mass;
        float wsum = 1.0f;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            fxs[k][a] = fx[a];
            mvs[k][a] = v[a] * mass;
            wsum *= fabsf(w[a][0]) + fabsf(w[a][1]) + fabsf(w[a][2]);
            rb[k] = rb[k] * TL::R + (base[a] - torg[a]);
        }
#pragma unroll
        for (int c = 0; c < D; ++c) {
            float s = 0.0f;
#pragma unroll
            for (int b = 0; b < D; ++b) s += fabsf(As[k].a[c][b]);
            const float bc = wsum * (fabsf(mvs[k][c]) + 2.5f * sp.dx * s);
            finite[k] = finite[k] && bc < 3.0e38f;
            if (bc < 3.0e38f) bound[c] = fmaxf(bound[c], bc);  // non-finite contributions take the flag path
        }
        const float bm = wsum * fabsf(mass);
        finite[k] = finite[k] && bm < 3.0e38f;
        if (bm < 3.0e38f) bound[3] = fmaxf(bound[3], bm);
    }
    // workgroup maximum of the particle bounds -> per channel scale 2^e with bound * 2^e < 2^50
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
    __syncthreads();
    float scale[4];
    double inv_scale[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float bmax;
        if constexpr (NW == 4) {
            bmax = fmaxf(fmaxf(s_bound[0][c], s_bound[1][c]), fmaxf(s_bound[2][c], s_bound[3][c]));
        } else {
            bmax = s_bound[0][c];
#pragma unroll
            for (int w = 1; w < NW; ++w) bmax = fmaxf(bmax, s_bound[w][c]);
        }
        fixed_point_scale(bmax, scale[c], inv_scale[c]);
    }

    // pass 2: quantised contributions -> 64-bit integer LDS atomics.  A particle whose bound is
    // finite has every |contribution * scale| below 2^50 and needs no per-contribution check.
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!valid[k]) continue;
        // an empty asm that takes the fractional offsets as operands: the per-node weight products
        // of the deposit cannot be computed before the barrier above.  Hoisted there, with one
        // particle per thread, they held the kernel at 256 VGPRs (one wave per SIMD); now 66-110.
#pragma unroll
        for (int a = 0; a < D; ++a) asm volatile("" : "+v"(fxs[k][a]));
        if (finite[k])
        {
            if constexpr (D == 3 && MODE == kSvdAccurate)
                p2g_deposit_fused<false>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale,
                                         tile_region<3>());
            else
                p2g_deposit<D, false>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale,
                                      tile_region<D>());
        }
        else if constexpr (D == 3 && MODE == kSvdAccurate)
            p2g_deposit_fused<true>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale,
                                    tile_region<3>());
        else
            p2g_deposit<D, true>(acc, nonfinite, rb[k], fxs[k], As[k], mvs[k], ms[k], sp.dx, scale, tile_region<D>());
    }
    __syncthreads();
    // every work item owns a slot: plain stores, no global atomics
    float4* dst = slots + int64_t(item) * REG;
    const float nan = __builtin_nanf("");
    for (int r = threadIdx.x; r < REG; r += blockDim.x) {
        float out[4];
        const uint32_t bad = nonfinite[r];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            out[c] = (bad >> c) & 1u ? nan : fixed_point_value(acc[c * REG + r], inv_scale[c]);
        dst[r] = make_float4(out[0], out[1], out[2], out[3]);
    }
}

// ---------------------------------------------------------------------------------------------
// grid update (src/nclr.h:263-310): one workgroup per node tile (256 threads; 1024 for a 2D launch
// with few work items).  In 3D a node sums its candidate tiles' work items in order; in 2D in four
// partial sums, partial w over items w, w + 4, ... of every candidate tile in order, added in
// partial order: four independent chains of additions behind the slot loads (the 256-scene 2D
// batch 3.6 % faster; in 3D it measured slower).  The grid adjoint re-sums in the same order.
// ---------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(D == 2 ? 1024 : 256) void k_grid_tiles(const float4* __restrict__ slots,
                                                    const uint32_t* __restrict__ item_offset, StepParams p,
                                                    float4* __restrict__ gv, float4* __restrict__ gpre,
                                                    uint32_t* __restrict__ stamps, uint32_t* __restrict__ active) {
    using TL = Tile<D>;
    constexpr int NC = 1 << D;
    __shared__ int s_first[NC];  // first P2G work item of candidate tile c (its slot), or -1
    __shared__ int s_count[NC];  // number of work items of candidate tile c
    __shared__ uint32_t s_active;
    const uint32_t nt_idx = xcd_block();
    if (nt_idx >= p.node_tiles) return;
    const int scene = p.nscenes == 1 ? 0 : scene_of_node_tile(p, nt_idx);
    if (scene_stopped(p, scene)) return;  // a stopped scene's grid keeps its last completed step
    const SceneParams sp = p.scenes[scene];
    int nto[D];
    tile_coords<D>(nt_idx - sp.node_tile_base, sp.ntn, nto);
    bool occupied = false;
    if (threadIdx.x == 0) s_active = 0;
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
    if (!__syncthreads_or(occupied)) {
        if (threadIdx.x == 0) active[nt_idx] = 0;
        return;
    }
    const uint32_t stamp = *p.stamp;
    if (threadIdx.x == 0) stamps[nt_idx] = stamp;
    int first[NC], count[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        first[c] = s_first[c];
        count[c] = s_count[c];
    }
    uint32_t massive_nodes = 0;
    auto add = [](float4& acc, const float4 v) {
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
    };
    if constexpr (D == 2) {
        // With 1024 threads (a launch with few work items) quarter w of the workgroup computes
        // partial sum w, which quarters the chain of slot loads a node waits on; with 256 threads
        // each thread computes all four.  Both give the same sums.
        static_assert(TL::NODES == 256, "one node per thread of a 2D node tile");
        __shared__ float4 s_part[3][TL::NODES];
        const int l = int(threadIdx.x) & (TL::NODES - 1);
        const int n[2] = {nto[0] * TL::S + (l >> TL::LOG), nto[1] * TL::S + (l & (TL::S - 1))};
        const bool split = blockDim.x == 4 * TL::NODES;
        const int w = split ? int(threadIdx.x) >> 8 : 0;
        float4 g0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g1 = g0, g2 = g0, g3 = g0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (first[c] < 0) continue;
            const int r0 = n[0] - (nto[0] - (c & 1)) * TL::S, r1 = n[1] - (nto[1] - ((c >> 1) & 1)) * TL::S;
            if (r0 >= TL::R || r1 >= TL::R) continue;
            const float4* q = slots + int64_t(first[c]) * TL::REGION + (r0 * TL::R + r1);
            if (split) {
                for (int it = w; it < count[c]; it += 4) add(g0, q[int64_t(it) * TL::REGION]);
            } else {
                int it = 0;
                for (; it + 4 <= count[c]; it += 4) {
                    const float4 s0 = q[int64_t(it) * TL::REGION], s1 = q[int64_t(it + 1) * TL::REGION],
                                 s2 = q[int64_t(it + 2) * TL::REGION], s3 = q[int64_t(it + 3) * TL::REGION];
                    add(g0, s0);
                    add(g1, s1);
                    add(g2, s2);
                    add(g3, s3);
                }
                if (it < count[c]) add(g0, q[int64_t(it) * TL::REGION]);
                if (it + 1 < count[c]) add(g1, q[int64_t(it + 1) * TL::REGION]);
                if (it + 2 < count[c]) add(g2, q[int64_t(it + 2) * TL::REGION]);
            }
        }
        if (split) {
            if (w > 0) s_part[w - 1][l] = g0;
            __syncthreads();
            if (w == 0) {
                g1 = s_part[0][l];
                g2 = s_part[1][l];
                g3 = s_part[2][l];
            }
        }
        if (w == 0) {
            float4 g = g0;
            add(g, g1);
            add(g, g2);
            add(g, g3);
            massive_nodes += g.w > 0.0f ? 1u : 0u;
            if (gpre) gpre[int64_t(nt_idx) * TL::NODES + l] = g;
            gv[int64_t(nt_idx) * TL::NODES + l] = grid_node<D>(g, n, sp);
        }
    } else
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
            // items in order; four loads in flight before their (ordered) sums
            const float4* q = slots + int64_t(first[c]) * TL::REGION + r;
            int it = 0;
            for (; it + 4 <= count[c]; it += 4, q += 4 * TL::REGION) {
                const float4 s0 = q[0], s1 = q[TL::REGION], s2 = q[2 * TL::REGION], s3 = q[3 * TL::REGION];
                add(g, s0);
                add(g, s1);
                add(g, s2);
                add(g, s3);
            }
            for (; it < count[c]; ++it, q += TL::REGION) add(g, q[0]);
        }
        massive_nodes += g.w > 0.0f ? 1u : 0u;
        if (gpre) gpre[int64_t(nt_idx) * TL::NODES + l] = g;
        gv[int64_t(nt_idx) * TL::NODES + l] = grid_node<D>(g, n, sp);
    }
    // nodes with post-P2G mass > 0, per node tile (summed on demand by mpm_active_nodes)
    if (massive_nodes) atomicAdd(&s_active, massive_nodes);
    __syncthreads();
    if (threadIdx.x == 0) active[nt_idx] = s_active;
}

// Sum of v[first .. first+n) into *out (one workgroup).
__global__ __launch_bounds__(1024) void k_sum_u32(const uint32_t* __restrict__ v, int64_t n,
                                                  unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[1024];
    unsigned long long s = 0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

// ---------------------------------------------------------------------------------------------
// G2P (src/nclr.h:167-261), reading slot perm[j] and writing slot j of the other buffer
// ---------------------------------------------------------------------------------------------
template <int D, int MAT, int MODE>
__global__ __launch_bounds__(256) void k_g2p_sorted(const float* __restrict__ st, float* __restrict__ out,
                                                    const uint32_t* __restrict__ pid_in,
                                                    uint32_t* __restrict__ pid_out, const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ keys, StepParams p,
                                                    const float4* __restrict__ gv) {
    using L = Layout<D>;
    const int64_t j = int64_t(xcd_block()) * blockDim.x + threadIdx.x;
    if (j >= p.n) return;
    const int64_t i = perm[j];
    const int64_t s = p.stride;
    const uint32_t id = pid_in[i];
    pid_out[j] = id;
    const int scene = p.nscenes == 1 ? 0 : scene_of_sorted(p, keys, j, id);
    const SceneParams sp = p.scenes[scene];
    const int material = MAT == kMixed ? sp.material : MAT;
    float x[D];
    M<D> F;
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = ldf(st, s, L::X + a, i);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) F.a[a][b] = ldf(st, s, L::F + a * D + b, i);
    float Jp = ldf(st, s, L::JP, i);
    out[int64_t(L::MASS) * s + j] = ldf(st, s, L::MASS, i);
    out[int64_t(L::VOL) * s + j] = ldf(st, s, L::VOL, i);
    int base[D];
    float fx[D], w[D][3];
    if (scene_stopped(p, scene) || !stencil<D>(x, sp, base, fx, w)) {
        // a stopped scene (the sort found one of its particles off the grid in this or an earlier
        // step): carried through unchanged; the step that stopped it is recorded (same value from
        // every particle of the scene)
        if (p.scene_frozen[scene] == 0u) p.scene_frozen[scene] = *p.stamp;
#pragma unroll
        for (int f = 0; f < L::JP + 1; ++f) out[int64_t(f) * s + j] = ldf(st, s, f, i);
        return;
    }
    float v[D];
    M<D> C;
    const StencilRows<D> node(gv, base, sp.ntn, sp.node_tile_base);
    g2p_particle<D, MODE>(material, sp, fx, w, node, x, v, C, F, Jp);
    out[int64_t(L::JP) * s + j] = Jp;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        out[int64_t(L::X + a) * s + j] = x[a];
        out[int64_t(L::V + a) * s + j] = v[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            out[int64_t(L::C + a * D + b) * s + j] = C.a[a][b];
            out[int64_t(L::F + a * D + b) * s + j] = F.a[a][b];
        }
}

// Dense export of one scene's updated grid in the reference's flat x-major order
// (src/nclr.h:87); nodes of node tiles the last step did not touch are empty cells (zero).
// A stopped scene shows the grid of its last completed step.
template <int D>
__global__ __launch_bounds__(256) void k_export_grid(const float4* __restrict__ gv, const uint32_t* __restrict__ stamps,
                                                     const uint32_t* __restrict__ stamp,
                                                     const uint32_t* __restrict__ frozen, int res,
                                                     uint32_t node_tile_base, int ntn, float* __restrict__ vel,
                                                     float* __restrict__ mass) {
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t side = res + 1;
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
    const int64_t node = gv_index<D>(n, ntn, node_tile_base);
    const uint32_t now = *frozen != 0u ? *frozen - 1u : *stamp;
    const bool live = now != 0 && stamps[node / Tile<D>::NODES] == now;
    const float4 g = live ? gv[node] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    vel[idx * D + 0] = g.x;
    vel[idx * D + 1] = g.y;
    if constexpr (D == 3) vel[idx * D + 2] = g.z;
    mass[idx] = g.w;
}

// The step's SVD on a batch of row-major matrices.
template <int D, int MODE>
__global__ __launch_bounds__(256) void k_svd(const float* __restrict__ A, float* __restrict__ U,
                                             float* __restrict__ S, float* __restrict__ V, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    M<D> a, u, v;
    float s[D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) a.a[r][c] = A[(i * D + r) * D + c];
    step_svd<MODE, D>(a, u, s, v);
#pragma unroll
    for (int r = 0; r < D; ++r) {
        S[i * D + r] = s[r];
#pragma unroll
        for (int c = 0; c < D; ++c) {
            U[(i * D + r) * D + c] = u.a[r][c];
            V[(i * D + r) * D + c] = v.a[r][c];
        }
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t value) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = uint32_t(i);
}

__global__ void k_fill_f4(float4* __restrict__ p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace nclr_dev

// ---------------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------------
namespace nclr_dev {

__global__ void k_bump_u32(uint32_t* __restrict__ p) { *p = *p + 1u; }

}  // namespace nclr_dev

namespace nclr_host {

using namespace nclr_dev;

static inline unsigned blocks_for(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

static int key_bits(uint32_t max_key) {
    int b = 0;
    while (b < 32 && (uint64_t(1) << b) <= max_key) ++b;
    return b;
}

template <int D>
static hipError_t sort_dispatch(const LaunchArgs& a) {
    const auto& p = a.p;
    const int64_t nb = int64_t(p.ntiles) + 1;
    hipStream_t s = a.stream;
    k_keys<D><<<blocks_for(std::max<int64_t>(std::max<int64_t>(p.n, nb), 1), 256), 256, 0, s>>>(
        a.state_in, a.pid_in, p, a.stamp, a.keys, a.status, a.tile_start, a.tile_end);
    if (p.n > 0) {
        hipError_t e = radix_sort_pairs(a.keys, a.perm, a.keys_tmp, a.perm_tmp, a.hist, p.n, key_bits(p.ntiles),
                                        true, s);
        if (e != hipSuccess) return e;
        k_tile_bounds<<<blocks_for(p.n, 256), 256, 0, s>>>(a.keys, p.n, a.tile_start, a.tile_end);
    }
    const uint32_t rounds = (p.ntiles + kScanRound - 1) / kScanRound;
    if (rounds <= 1) {
        k_items<<<1, 1024, 0, s>>>(a.tile_start, a.tile_end, p.ntiles, uint32_t(p.chunk), a.item_offset, p.items,
                                   a.meta);
    } else {
        k_items_round_totals<<<rounds, 1024, 0, s>>>(a.tile_start, a.tile_end, p.ntiles, uint32_t(p.chunk), a.rtot);
        k_items_rounds<<<rounds, 1024, 0, s>>>(a.tile_start, a.tile_end, p.ntiles, uint32_t(p.chunk), a.item_offset,
                                               p.items, a.rtot, a.meta);
    }
    return hipGetLastError();
}

// 256 threads, chunk / 256 particles each; in 2D, when the particles fill fewer work items than
// the chip has CUs, one particle per thread (a workgroup of chunk threads), which halves the
// latency of a work item (the results are the same: the sums are exact and the scale is the
// item's maximum).  Measured: the 19,600-particle 2D jelly step's P2G faster, the 1.75M-particle
// 2D batch's slower with it.
template <int D, int MAT, int PPT>
static hipError_t p2g_one(const LaunchArgs& a) {
    const int64_t items = (a.p.n + 256 * PPT - 1) / (256 * PPT) + a.p.ntiles;
    const unsigned g = unsigned(items);
    if constexpr (D == 2) {
        if (a.p.n < int64_t(kFewItems) * 256 * PPT)
            k_p2g_tile<D, MAT, 1, kSvdExact, 256 * PPT><<<g, 256 * PPT, 0, a.stream>>>(a.state_in, a.perm,
                                                                                     a.item_offset, a.p, a.slots);
        else
            k_p2g_tile<D, MAT, PPT, kSvdExact><<<g, 256, 0, a.stream>>>(a.state_in, a.perm, a.item_offset, a.p,
                                                                       a.slots);
    } else if (a.svd_mode == kSvdFast) {
        k_p2g_tile<D, MAT, PPT, kSvdFast><<<g, 256, 0, a.stream>>>(a.state_in, a.perm, a.item_offset, a.p, a.slots);
    } else if (a.svd_mode == kSvdAccurate) {
        k_p2g_tile<D, MAT, PPT, kSvdAccurate><<<g, 256, 0, a.stream>>>(a.state_in, a.perm, a.item_offset, a.p,
                                                                      a.slots);
    } else {
        k_p2g_tile<D, MAT, PPT, kSvdExact><<<g, 256, 0, a.stream>>>(a.state_in, a.perm, a.item_offset, a.p, a.slots);
    }
    return hipGetLastError();
}

template <int D, int MAT>
static hipError_t p2g_chunk(const LaunchArgs& a) {
    switch (a.p.chunk) {
        case 256: return p2g_one<D, MAT, 1>(a);
        case 1024: return p2g_one<D, MAT, 4>(a);
        default: return p2g_one<D, MAT, 2>(a);
    }
}

template <int D>
static hipError_t p2g_dispatch(const LaunchArgs& a) {
    if (a.p.n == 0) return hipSuccess;
    switch (a.material) {
        case MPM_SNOW: return p2g_chunk<D, MPM_SNOW>(a);
        case MPM_JELLY: return p2g_chunk<D, MPM_JELLY>(a);
        case MPM_LIQUID: return p2g_chunk<D, MPM_LIQUID>(a);
        default: return p2g_chunk<D, kMixed>(a);
    }
}

template <int D, int MAT>
static hipError_t g2p_one(const LaunchArgs& a) {
    const int threads = a.g2p_threads > 0 ? a.g2p_threads : 256;
    const unsigned g = xcd_grid(blocks_for(a.p.n, threads));
    if (D == 3 && a.svd_mode == kSvdFast)
        k_g2p_sorted<D, MAT, kSvdFast><<<g, threads, 0, a.stream>>>(a.state_in, a.state_out, a.pid_in, a.pid_out,
                                                                   a.perm, a.keys, a.p, a.gv);
    else if (D == 3 && a.svd_mode == kSvdAccurate)
        k_g2p_sorted<D, MAT, kSvdAccurate><<<g, threads, 0, a.stream>>>(a.state_in, a.state_out, a.pid_in,
                                                                       a.pid_out, a.perm, a.keys, a.p, a.gv);
    else
        k_g2p_sorted<D, MAT, kSvdExact><<<g, threads, 0, a.stream>>>(a.state_in, a.state_out, a.pid_in, a.pid_out,
                                                                    a.perm, a.keys, a.p, a.gv);
    return hipGetLastError();
}

template <int D>
static hipError_t g2p_dispatch(const LaunchArgs& a) {
    if (a.p.n == 0) return hipSuccess;
    switch (a.material) {
        case MPM_SNOW: return g2p_one<D, MPM_SNOW>(a);
        case MPM_JELLY: return g2p_one<D, MPM_JELLY>(a);
        case MPM_LIQUID: return g2p_one<D, MPM_LIQUID>(a);
        default: return g2p_one<D, kMixed>(a);
    }
}

hipError_t launch_sort(const LaunchArgs& a) { return a.dim == 2 ? sort_dispatch<2>(a) : sort_dispatch<3>(a); }
hipError_t launch_p2g(const LaunchArgs& a) { return a.dim == 2 ? p2g_dispatch<2>(a) : p2g_dispatch<3>(a); }
hipError_t launch_g2p(const LaunchArgs& a) { return a.dim == 2 ? g2p_dispatch<2>(a) : g2p_dispatch<3>(a); }

hipError_t launch_grid(const LaunchArgs& a) {
    const unsigned tiles = a.p.node_tiles;
    if (tiles == 0) return hipSuccess;
    if (a.dim == 2)  // few work items: four threads per node (the same sums)
        k_grid_tiles<2><<<xcd_grid(tiles), a.p.n < int64_t(kFewItems) * a.p.chunk ? 1024 : 256, 0, a.stream>>>(
            a.slots, a.item_offset, a.p, a.gv, a.gpre, a.stamps, a.active);
    else
        k_grid_tiles<3><<<xcd_grid(tiles), 256, 0, a.stream>>>(a.slots, a.item_offset, a.p, a.gv, a.gpre, a.stamps,
                                                                a.active);
    return hipGetLastError();
}

hipError_t launch_active_sum(const LaunchArgs& a, uint32_t first_tile, uint32_t tiles, unsigned long long* out) {
    k_sum_u32<<<1, 1024, 0, a.stream>>>(a.active + first_tile, int64_t(tiles), out);
    return hipGetLastError();
}

hipError_t launch_export_grid(const LaunchArgs& a, int scene, int res, uint32_t node_tile_base, int ntn, float* vel,
                              float* mass, int64_t total) {
    const uint32_t* frozen = a.p.scene_frozen + scene;
    if (a.dim == 2)
        k_export_grid<2><<<blocks_for(total, 256), 256, 0, a.stream>>>(a.gv, a.stamps, a.stamp, frozen, res,
                                                                       node_tile_base, ntn, vel, mass);
    else
        k_export_grid<3><<<blocks_for(total, 256), 256, 0, a.stream>>>(a.gv, a.stamps, a.stamp, frozen, res,
                                                                       node_tile_base, ntn, vel, mass);
    return hipGetLastError();
}

hipError_t launch_svd(int dim, int mode, const float* A, float* U, float* S, float* V, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned b = blocks_for(n, 256);
    if (dim == 2) k_svd<2, kSvdExact><<<b, 256, 0, s>>>(A, U, S, V, n);
    else if (mode == kSvdFast) k_svd<3, kSvdFast><<<b, 256, 0, s>>>(A, U, S, V, n);
    else if (mode == kSvdAccurate) k_svd<3, kSvdAccurate><<<b, 256, 0, s>>>(A, U, S, V, n);
    else k_svd<3, kSvdExact><<<b, 256, 0, s>>>(A, U, S, V, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<blocks_for(n, 256), 256, 0, s>>>(p, n, value);
    return hipGetLastError();
}

hipError_t launch_bump_u32(uint32_t* p, hipStream_t s) {
    k_bump_u32<<<1, 1, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_iota_u32<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_zero_f4(float4* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_fill_f4<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

int64_t max_items(int64_t n, uint32_t ntiles) { return (n + kMinChunk - 1) / kMinChunk + int64_t(ntiles); }
int num_fields(int dim) { return dim == 2 ? Layout<2>::NF : Layout<3>::NF; }
int region_nodes(int dim) { return dim == 2 ? Tile<2>::REGION : Tile<3>::REGION; }
int tile_nodes(int dim) { return dim == 2 ? Tile<2>::NODES : Tile<3>::NODES; }
int tile_log(int dim) { return dim == 2 ? Tile<2>::LOG : Tile<3>::LOG; }

}  // namespace nclr_host
