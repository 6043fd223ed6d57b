// Stable LSD counting sort (radix 256) of 32-bit keys with 32-bit values on gfx950.
//
// Each pass sorts by one 8-bit digit and keeps the previous order among equal digits, so
// the composition over ceil(bits/8) passes equals a stable sort by the whole key
// (std::stable_sort by key, bit for bit).  A pass is three kernels:
//   hist     one workgroup per tile of kTileKeys keys: digit counts -> hist[digit][tile]
//   scan     one workgroup: exclusive scan of hist in digit-major order
//   scatter  each tile re-ranks its keys in input order (wave match + per-wave digit
//            counts in LDS, waves in order) and writes key/value to its final slot.
// Keys that are equal across a wavefront (the common case: particles arrive nearly sorted
// by the previous step) cost one ballot per distinct digit, so both the histogram and the
// ranking are nearly free of LDS atomic contention.
#include <hip/hip_runtime.h>

#include "mpm_internal.hpp"

namespace nclr_dev {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;
constexpr int kTileKeys = kSortThreads * kSortItems;  // 4096

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of the current wave whose `d` equals this lane's (for valid lanes).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t todo = __ballot(valid);
    uint64_t mine = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t ld = __shfl(d, leader);
        const uint64_t same = __ballot(valid && d == ld);
        if (valid && d == ld) mine = same;
        todo &= ~same;
    }
    return mine;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n,
                                                             int shift, uint32_t dmask, uint32_t* __restrict__ hist,
                                                             int64_t ntiles) {
    __shared__ uint32_t counts[256];
    for (int i = threadIdx.x; i < 256; i += kSortThreads) counts[i] = 0;
    const int64_t base = int64_t(blockIdx.x) * kTileKeys;
    // all loads first, so that their latencies overlap
    uint32_t d[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t idx = base + int64_t(r) * kSortThreads + threadIdx.x;
        d[r] = idx < n ? (keys[idx] >> shift) & dmask : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const bool valid = base + int64_t(r) * kSortThreads + threadIdx.x < n;
        const uint64_t peers = match_digit(d[r], valid);
        // the lowest lane of each group adds the group's size
        if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&counts[d[r]], uint32_t(__popcll(peers)));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += kSortThreads) hist[int64_t(i) * ntiles + blockIdx.x] = counts[i];
}

// Inclusive scan of one value per thread over a workgroup of N threads (Hillis-Steele in LDS).
template <int N>
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* part) {
    const int t = threadIdx.x;
    part[t] = v;
    __syncthreads();
#pragma unroll
    for (int off = 1; off < N; off <<= 1) {
        const uint32_t u = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    return part[t];
}

// Per digit d (one workgroup each): exclusive scan of hist[d][0..tiles) in place, total to
// digit_total[d].  Coalesced rounds of 256 tiles.
__global__ __launch_bounds__(256) void k_radix_digit_scan(uint32_t* __restrict__ hist, int64_t tiles,
                                                          uint32_t* __restrict__ digit_total) {
    __shared__ uint32_t part[256];
    uint32_t* row = hist + int64_t(blockIdx.x) * tiles;
    uint32_t carry = 0;
    for (int64_t b = 0; b < tiles; b += 256) {
        const int64_t i = b + threadIdx.x;
        const uint32_t v = i < tiles ? row[i] : 0u;
        const uint32_t inc = block_inclusive_scan<256>(v, part);
        if (i < tiles) row[i] = carry + inc - v;
        carry += part[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out, int64_t n,
                                                                int shift, uint32_t dmask,
                                                                const uint32_t* __restrict__ offsets,
                                                                const uint32_t* __restrict__ digit_total,
                                                                int64_t ntiles, int identity_values) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wave_count[kSortWaves][256];
    __shared__ uint32_t part[256];
    const int t = threadIdx.x;
    const int wave = t >> 6;
    // start of digit t = keys with a smaller digit + keys with digit t in earlier tiles
    const uint32_t tot = digit_total[t];
    const uint32_t digit_base = block_inclusive_scan<kSortThreads>(tot, part) - tot;
    running[t] = digit_base + offsets[int64_t(t) * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) wave_count[w][t] = 0;
    __syncthreads();
    const int64_t base = int64_t(blockIdx.x) * kTileKeys;
    // all loads first, so that their latencies overlap; then the ranking rounds in input order
    uint32_t keys_r[kSortItems], vals_r[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t idx = base + int64_t(r) * kSortThreads + t;
        const bool valid = idx < n;
        keys_r[r] = valid ? keys_in[idx] : 0u;
        vals_r[r] = valid ? (identity_values ? uint32_t(idx) : vals_in[idx]) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t idx = base + int64_t(r) * kSortThreads + t;
        const bool valid = idx < n;
        const uint32_t key = keys_r[r];
        const uint32_t val = vals_r[r];
        const uint32_t d = (key >> shift) & dmask;
        const uint64_t peers = match_digit(d, valid);
        const uint32_t rank = uint32_t(__popcll(peers & lanemask_lt()));
        const bool leader = valid && (peers & lanemask_lt()) == 0;
        if (leader) wave_count[wave][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wave_count[w][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&running[d], wave_count[wave][d]);
            wave_count[wave][d] = 0;
        }
        __syncthreads();
    }
}

}  // namespace nclr_dev

namespace nclr_host {

using namespace nclr_dev;

int64_t sort_tiles(int64_t n) { return (n + kTileKeys - 1) / kTileKeys; }

// Sorts (keys, values) stably by the low `bits` bits of the keys.  keys/vals hold the input and,
// on return, the output; keys_tmp/vals_tmp are scratch of n entries; hist holds 256*tiles + 256
// counters.  When identity_values is set the input values are taken to be 0..n-1 (vals is
// then output only).
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                            uint32_t* hist, int64_t n, int bits, bool identity_values, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = sort_tiles(n);
    const int passes = bits <= 0 ? 1 : (bits + 7) / 8;
    uint32_t *kin = keys, *vin = vals, *kout = keys_tmp, *vout = vals_tmp;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        const int dbits = bits - shift < 8 ? (bits - shift > 0 ? bits - shift : 0) : 8;
        const uint32_t dmask = (1u << dbits) - 1u;  // digits use only the requested low bits
        uint32_t* digit_total = hist + 256 * tiles;
        k_radix_hist<<<unsigned(tiles), kSortThreads, 0, s>>>(kin, n, shift, dmask, hist, tiles);
        k_radix_digit_scan<<<256, 256, 0, s>>>(hist, tiles, digit_total);
        k_radix_scatter<<<unsigned(tiles), kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift, dmask, hist,
                                                                 digit_total, tiles,
                                                                 (identity_values && p == 0) ? 1 : 0);
        uint32_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (kin != keys) {
        e = hipMemcpyAsync(keys, kin, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(vals, vin, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    }
    return e;
}

}  // namespace nclr_host
