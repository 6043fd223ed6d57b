// Incremental stable counting sort of particles by tile, for every step after the first.
//
// After a step the state is stored in the order of that step's sort, so the particles of each old
// tile are contiguous, in the previous step's P2G work items.  A particle moves less than one cell
// per step (the grid velocity is clamped to 0.9 dx / dt per component, src/nclr.h:285), so its new
// tile is one of the 3^D neighbours of its old tile.  The stable sort by new tile then needs no
// radix passes:
//
//   count    one workgroup per old work item: each particle's new key and neighbour offset
//            o = new - old (or the off-grid bucket), its rank among the item's particles with the
//            same offset (in slot order), and the item's count per offset
//   prefix   one wave per new tile: its <= 3^D source tiles in increasing order, and their items
//            in order, give each (item, offset) the number of earlier particles bound for the tile
//   scan     one workgroup: bucket starts (exclusive scan of the tile totals) and the P2G work
//            items of the new order
//   scatter  per slot: position = start[new tile] + prefix[item][offset] + rank -> perm, keys
//
// Positions follow slot order within every tile's bucket, i.e. for particles on their grid the
// result is the stable sort by key of the current order, identical to the radix sort's
// (mpm_sort.hip) bit for bit.  Particles off their grid fill the last bucket in any order.
#include <hip/hip_runtime.h>

#include "mpm_device.hpp"

namespace nclr_dev {

template <int D>
struct Buckets {
    static constexpr int NEIGH = D == 2 ? 9 : 27;
    static constexpr int NO = NEIGH + 1;  // neighbour offsets + the off-grid bucket
    static constexpr int OFF = NEIGH;     // index of the off-grid bucket
    static constexpr int SELF = NEIGH / 2;
};

// Offset index of a move by delta[a] in {-1, 0, 1}: sum (delta_a + 1) 3^(D-1-a).
template <int D>
NCLR_DI int offset_index(const int (&delta)[D]) {
    int o = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) o = o * 3 + (delta[a] + 1);
    return o;
}

template <int D>
__global__ __launch_bounds__(256) void k_bucket_count(const float* __restrict__ st, StepParams p,
                                                      uint32_t* __restrict__ stamp,
                                                      const uint32_t* __restrict__ item_offset,
                                                      uint32_t* __restrict__ newkey, uint32_t* __restrict__ rankpk,
                                                      uint32_t* __restrict__ uidx, uint32_t* __restrict__ ucnt,
                                                      uint32_t* __restrict__ status) {
    using B = Buckets<D>;
    __shared__ uint32_t running[B::NO];
    __shared__ uint32_t wave_count[4][B::NO];
    const uint32_t item = blockIdx.x;
    if (item == 0 && threadIdx.x == 0) *stamp = *stamp + 1u;
    const WorkItem wi = work_item(p, item);  // loaded with the count below, used only if valid
    if (item >= item_offset[p.ntiles]) return;
    for (int o = threadIdx.x; o < B::NO; o += blockDim.x) {
        running[o] = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) wave_count[w][o] = 0;
    }
    __syncthreads();
    const uint32_t tile = wi.tile;
    const int scene = p.nscenes == 1 ? 0 : scene_of_tile(p, tile);
    const SceneParams sp = p.scenes[scene];
    int tc[D];
    tile_coords<D>(tile - sp.tile_base, sp.nt, tc);
    const uint32_t j0 = wi.j0, j1 = wi.j1;
    const int wave = threadIdx.x >> 6;
    // new keys and offsets of all rounds first (the loads overlap), then the ranking rounds
    constexpr int kRounds = 4;  // work items hold at most 1024 particles
    int o_r[kRounds];
    uint32_t key_r[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t j = j0 + uint32_t(r) * 256 + threadIdx.x;
        o_r[r] = B::OFF;
        key_r[r] = p.ntiles;
        if (j >= j1) continue;
        float x[D];
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = ldf(st, p.stride, Layout<D>::X + a, j);
        int base[D];
        float fx[D], w[D][3];
        if (stencil<D>(x, sp, base, fx, w)) {
            int delta[D];
            bool near = true;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                delta[a] = (base[a] >> Tile<D>::LOG) - tc[a];
                near = near && delta[a] >= -1 && delta[a] <= 1;
            }
            if (near) {
                o_r[r] = offset_index<D>(delta);
                key_r[r] = sp.tile_base + tile_key<D>(base, sp.nt);
            } else {
                stop_scene(p, scene, kStatusMovedTooFar);
            }
        } else {
            stop_scene(p, scene, kStatusOutOfRange);
        }
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t r0 = j0 + uint32_t(r) * 256;
        if (r0 >= j1) break;  // uniform over the workgroup
        const uint32_t j = r0 + threadIdx.x;
        const bool valid = j < j1;
        const int o = o_r[r];
        const uint32_t key = key_r[r];
        // stable rank among this item's particles with the same offset: wave match, then waves
        // in order (per-wave counts in LDS), then earlier rounds (running)
        uint64_t todo = __ballot(valid);
        uint64_t peers = 0;
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int lo = __shfl(o, leader);
            const uint64_t same = __ballot(valid && o == lo);
            if (valid && o == lo) peers = same;
            todo &= ~same;
        }
        const int lane = threadIdx.x & 63;
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const bool leader = valid && (peers & below) == 0;
        if (leader) wave_count[wave][o] = uint32_t(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t rank = running[o] + uint32_t(__popcll(peers & below));
            for (int w = 0; w < wave; ++w) rank += wave_count[w][o];
            newkey[j] = key;
            rankpk[j] = (uint32_t(o) << 16) | rank;
            uidx[j] = item;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&running[o], wave_count[wave][o]);
            wave_count[wave][o] = 0;
        }
        __syncthreads();
    }
    for (int o = threadIdx.x; o < B::NO; o += blockDim.x) ucnt[int64_t(item) * B::NO + o] = running[o];
}

// One wave per new tile k, lane s < 3^D for the source tile kc + d(s) - 1 (d(s) the base-3 digits
// of s, so that increasing s is increasing source tile): each lane sums its source's items bound
// for k, a wave scan orders the sources, and each lane then writes its items' prefixes in order.
template <int D>
__global__ __launch_bounds__(256) void k_bucket_prefix(StepParams p, const uint32_t* __restrict__ item_offset,
                                                       const uint32_t* __restrict__ ucnt,
                                                       uint32_t* __restrict__ upref, uint32_t* __restrict__ ttotal) {
    using B = Buckets<D>;
    const int64_t k = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int s = threadIdx.x & 63;
    if (k >= int64_t(p.ntiles)) return;  // whole waves
    const SceneParams sp = p.scenes[p.nscenes == 1 ? 0 : scene_of_tile(p, uint32_t(k))];
    int kc[D];
    tile_coords<D>(uint32_t(k) - sp.tile_base, sp.nt, kc);
    bool inside = s < B::NEIGH;
    uint32_t t = 0;
    {
        int sc[D];
        int rem = s;
#pragma unroll
        for (int a = D - 1; a >= 0; --a) {
            sc[a] = kc[a] + (rem % 3) - 1;
            rem /= 3;
            inside = inside && sc[a] >= 0 && sc[a] < sp.nt;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) t = t * uint32_t(sp.nt) + uint32_t(inside ? sc[a] : 0);
        t += sp.tile_base;
    }
    const int o = B::NEIGH - 1 - s;  // the move from source tile to k
    const uint32_t u0 = inside ? item_offset[t] : 0u, u1 = inside ? item_offset[t + 1] : 0u;
    // the source items' counts, four loads in flight (integer sums: the order is free)
    const uint32_t* cnt = ucnt + o;
    uint32_t mine = 0;
    uint32_t u = u0;
    for (; u + 4 <= u1; u += 4) {
        const uint32_t c0 = cnt[int64_t(u) * B::NO], c1 = cnt[int64_t(u + 1) * B::NO],
                       c2 = cnt[int64_t(u + 2) * B::NO], c3 = cnt[int64_t(u + 3) * B::NO];
        mine += (c0 + c1) + (c2 + c3);
    }
    for (; u < u1; ++u) mine += cnt[int64_t(u) * B::NO];
    uint32_t x = mine;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {  // 3^D <= 27 lanes take part
        const uint32_t y = __shfl_up(x, off);
        if (s >= off) x += y;
    }
    uint32_t run = x - mine;
    u = u0;
    for (; u + 4 <= u1; u += 4) {
        const uint32_t c0 = cnt[int64_t(u) * B::NO], c1 = cnt[int64_t(u + 1) * B::NO],
                       c2 = cnt[int64_t(u + 2) * B::NO], c3 = cnt[int64_t(u + 3) * B::NO];
        upref[int64_t(u) * B::NO + o] = run;
        upref[int64_t(u + 1) * B::NO + o] = run + c0;
        upref[int64_t(u + 2) * B::NO + o] = run + c0 + c1;
        upref[int64_t(u + 3) * B::NO + o] = run + c0 + c1 + c2;
        run += (c0 + c1) + (c2 + c3);
    }
    for (; u < u1; ++u) {
        upref[int64_t(u) * B::NO + o] = run;
        run += cnt[int64_t(u) * B::NO];
    }
    const uint32_t total = __shfl(x, B::NEIGH - 1);
    if (s == 0) ttotal[k] = total;
}

// Bucket starts (exclusive scan of the tile totals) and the P2G work items of the new order.  On
// entry prev_meta[0] holds the particles on their grid in the old order (left there by the
// previous sort; prev_meta is meta itself unless the step sorts into another block).  On exit
// meta[0] = those of the new order (the start of the off-grid bucket), meta[1] = 0 (the off-grid
// bucket's fill counter), meta[2] = the old count.  Particles off their grid take
// no part in P2G and G2P copies them unchanged, so their order within the last bucket cannot
// affect any result; they are placed by an atomic counter.
//
// Up to kScanRound tiles: k_bucket_scan, one workgroup.  More: k_bucket_round_totals and
// k_bucket_rounds, one workgroup per round of kScanRound tiles (totals, then each round with the
// sum of the earlier rounds' totals as its carry); the results are the same.
struct BucketVisit {
    uint32_t *tile_start, *tile_end, *item_offset;
    uint4* items;
    uint32_t chunk;
    NCLR_DI void operator()(uint32_t tile, uint32_t first_p, uint32_t c, uint32_t first_i, uint32_t its) const {
        tile_start[tile] = first_p;
        tile_end[tile] = first_p + c;
        item_offset[tile] = first_i;
        write_items(items, first_i, its, tile, first_p, c, chunk);
    }
};

template <int D>
__global__ __launch_bounds__(1024) void k_bucket_scan(StepParams p, uint32_t chunk,
                                                      uint32_t* __restrict__ item_offset,
                                                      uint32_t* __restrict__ tile_start,
                                                      uint32_t* __restrict__ tile_end,
                                                      const uint32_t* __restrict__ ttotal,
                                                      const uint32_t* prev_meta, uint32_t* meta) {
    __shared__ uint32_t s_seen;
    if (threadIdx.x == 0) s_seen = prev_meta[0];
    __syncthreads();  // meta (possibly the same buffer as prev_meta) is rewritten below
    const BucketVisit visit{tile_start, tile_end, item_offset, p.items, chunk};
    auto count = [&](uint32_t tile) { return ttotal[tile]; };
    uint32_t cp = 0, ci = 0;
    for (uint32_t b0 = 0; b0 < p.ntiles; b0 += kScanRound) {
        uint32_t tp, ti;
        scan_round(b0, p.ntiles, chunk, cp, ci, count, visit, tp, ti);
        cp += tp;
        ci += ti;
    }
    if (threadIdx.x == 0) {
        item_offset[p.ntiles] = ci;  // the total
        meta[0] = cp;
        meta[1] = 0;
        meta[2] = s_seen;
    }
}

__global__ __launch_bounds__(1024) void k_bucket_round_totals(uint32_t ntiles, uint32_t chunk,
                                                              const uint32_t* __restrict__ ttotal,
                                                              uint32_t* __restrict__ rtot, const uint32_t* prev_meta,
                                                              uint32_t* meta) {
    uint32_t tp, ti;
    round_totals(blockIdx.x, ntiles, chunk, [&](uint32_t tile) { return ttotal[tile]; }, tp, ti);
    if (threadIdx.x == 0) {
        rtot[2 * blockIdx.x] = tp;
        rtot[2 * blockIdx.x + 1] = ti;
        if (blockIdx.x == 0) meta[2] = prev_meta[0];
    }
}

template <int D>
__global__ __launch_bounds__(1024) void k_bucket_rounds(StepParams p, uint32_t chunk, uint32_t* __restrict__ item_offset,
                                                        uint32_t* __restrict__ tile_start,
                                                        uint32_t* __restrict__ tile_end,
                                                        const uint32_t* __restrict__ ttotal,
                                                        const uint32_t* __restrict__ rtot, uint32_t* __restrict__ meta) {
    uint32_t cp, ci, tp, ti;
    round_carries(rtot, blockIdx.x, cp, ci);
    const BucketVisit visit{tile_start, tile_end, item_offset, p.items, chunk};
    scan_round(blockIdx.x * kScanRound, p.ntiles, chunk, cp, ci, [&](uint32_t tile) { return ttotal[tile]; }, visit,
               tp, ti);
    if (threadIdx.x == 0 && blockIdx.x + 1 == gridDim.x) {
        item_offset[p.ntiles] = ci + ti;
        meta[0] = cp + tp;
        meta[1] = 0;
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_bucket_scatter(StepParams p, const uint32_t* __restrict__ newkey,
                                                        const uint32_t* __restrict__ rankpk,
                                                        const uint32_t* __restrict__ uidx,
                                                        const uint32_t* __restrict__ upref,
                                                        const uint32_t* __restrict__ tile_start,
                                                        const uint32_t* __restrict__ pid,
                                                        uint32_t* __restrict__ meta, uint32_t* __restrict__ perm,
                                                        uint32_t* __restrict__ keys, uint32_t* __restrict__ status) {
    using B = Buckets<D>;
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= p.n) return;
    const uint32_t seen = meta[2];
    uint32_t pos, key;
    const uint32_t pk = uint32_t(j) < seen ? rankpk[j] : uint32_t(B::OFF) << 16;
    const uint32_t o = pk >> 16;
    if (o != uint32_t(B::OFF)) {
        key = newkey[j];
        pos = tile_start[key] + upref[int64_t(uidx[j]) * B::NO + o] + (pk & 0xffffu);
    } else {  // off the grid (now or already before the step)
        key = p.ntiles;
        pos = meta[0] + atomicAdd(&meta[1], 1u);
        // already off the grid before the step: its scene is reported again (every advance() of the
        // reference throws for that scene again)
        if (uint32_t(j) >= seen) stop_scene(p, p.nscenes == 1 ? 0 : scene_of_particle(p, pid[j]), kStatusOutOfRange);
    }
    if (pos < p.n) {
        perm[pos] = uint32_t(j);
        keys[pos] = key;
    } else {
        atomicOr(status, kStatusMovedTooFar);
    }
}

}  // namespace nclr_dev

namespace nclr_host {

using namespace nclr_dev;

template <int D>
static hipError_t bucket_dispatch(const LaunchArgs& a) {
    const auto& p = a.p;
    hipStream_t s = a.stream;
    const int64_t items = max_items(p.n, p.ntiles);
    // the old order's work items (the previous sort's block) are read by the count and prefix
    // kernels; the scans write the new ones (into the same block or another)
    StepParams prev = p;
    prev.items = const_cast<uint4*>(a.prev_items);
    k_bucket_count<D><<<unsigned(std::max<int64_t>(items, 1)), 256, 0, s>>>(
        a.state_in, prev, a.stamp, a.prev_item_offset, a.keys_tmp, a.perm_tmp, a.uidx, a.ucnt, a.status);
    if (p.ntiles > 0)
        k_bucket_prefix<D><<<unsigned((int64_t(p.ntiles) + 3) / 4), 256, 0, s>>>(p, a.prev_item_offset, a.ucnt,
                                                                                a.upref, a.ttotal);
    const uint32_t rounds = (p.ntiles + kScanRound - 1) / kScanRound;
    if (rounds <= 1) {
        k_bucket_scan<D><<<1, 1024, 0, s>>>(p, uint32_t(p.chunk), a.item_offset, a.tile_start, a.tile_end, a.ttotal,
                                            a.prev_meta, a.meta);
    } else {
        k_bucket_round_totals<<<rounds, 1024, 0, s>>>(p.ntiles, uint32_t(p.chunk), a.ttotal, a.rtot, a.prev_meta,
                                                      a.meta);
        k_bucket_rounds<D><<<rounds, 1024, 0, s>>>(p, uint32_t(p.chunk), a.item_offset, a.tile_start, a.tile_end,
                                                   a.ttotal, a.rtot, a.meta);
    }
    if (p.n > 0)
        k_bucket_scatter<D><<<unsigned((p.n + 255) / 256), 256, 0, s>>>(p, a.keys_tmp, a.perm_tmp, a.uidx, a.upref,
                                                                        a.tile_start, a.pid_in, a.meta, a.perm,
                                                                        a.keys, a.status);
    return hipGetLastError();
}

hipError_t launch_sort_incremental(const LaunchArgs& a) {
    return a.dim == 2 ? bucket_dispatch<2>(a) : bucket_dispatch<3>(a);
}

int bucket_offsets(int dim) { return dim == 2 ? Buckets<2>::NO : Buckets<3>::NO; }

}  // namespace nclr_host
