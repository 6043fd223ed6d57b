// Scene generators and the viewer projection on the device: the reference's scene helpers of
// src/nclr_math.h (cube 100-129, nc_rand_int / nc_rand / randvec 22-39, gyroid 131-139, sample
// 141-147, pt_3d_to_2d 149-167), producing particle state directly in HBM.
//
// Output format ("state image"): float channels x (D), v (D), F (D*D, row-major), C (D*D), Jp, each
// a row of n floats, particle i at column i; mpm_set_state_device takes exactly this image.  Every
// generator writes all channels: v from the shape, F = diag<D>(1) (diag(1, 1, 0) in 3D for the
// reference identity mode, I for the physical one), C = 0, Jp = 1.
//
// Random draws reproduce the reference's global xorshift128 stream (Marsaglia 2003, the
// transition of src/nclr_math.h:22-29; uniform = u32 * 2^-32 in float, src/nclr_math.h:31) in
// parallel: every thread produces kDrawsPerThread consecutive draws from its own start state,
// which the host computes by jumping the 128-bit state ahead (the transition is linear over
// GF(2), so T^k is a 128 x 128 bit matrix).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "mpm_internal.hpp"

namespace nclr_dev {

constexpr int kDrawsPerThread = 256;

// Eigen's LinSpaced(n, low, high)(i) for float (low + i * step, the last value exactly high;
// built from the high end when |high| < |low|), in FP32 without contraction.
__device__ __forceinline__ float linspaced_at(int i, int n, float lo, float hi) {
    if (n == 1) return hi;
    const int size1 = n - 1;
    const float step = (hi - lo) / float(size1);
    if (fabsf(hi) < fabsf(lo)) return i == 0 ? lo : hi - float(size1 - i) * step;
    return i == size1 ? hi : lo + float(i) * step;
}

// v, F = diag<D>(1) (or I), C = 0, Jp = 1 of particle i; x is written by the caller.
template <int D>
__device__ __forceinline__ void store_defaults(float* __restrict__ img, int64_t n, int64_t i, const float* v,
                                               int identity_mode) {
#pragma unroll
    for (int a = 0; a < D; ++a) img[int64_t(D + a) * n + i] = v ? v[a] : 0.0f;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            const float f = (a == b && (a < 2 || identity_mode == 1)) ? 1.0f : 0.0f;
            img[int64_t(2 * D + a * D + b) * n + i] = f;
            img[int64_t(2 * D + D * D + a * D + b) * n + i] = 0.0f;
        }
    img[int64_t(2 * D + 2 * D * D) * n + i] = 1.0f;
}

// Lattice coordinate of point l of a cube, last axis fastest (src/nclr_math.h:107-122).
template <int D>
__device__ __forceinline__ void lattice_point(const mpm_cube& c, int64_t l, float (&x)[D]) {
    int64_t rem = l;
    int idx[D];
#pragma unroll
    for (int a = D - 1; a >= 0; --a) {
        idx[a] = int(rem % c.res[a]);
        rem /= c.res[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = linspaced_at(idx[a], c.res[a], c.lo[a], c.hi[a]);
}

template <int D>
__global__ void k_gen_cubes(const mpm_cube* __restrict__ cubes, const int64_t* __restrict__ offsets, int ncubes,
                            int identity_mode, int64_t n, float* __restrict__ img) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = ncubes - 1;  // last cube with offsets[c] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const mpm_cube c = cubes[lo];
    float x[D];
    lattice_point<D>(c, i - offsets[lo], x);
#pragma unroll
    for (int a = 0; a < D; ++a) img[int64_t(a) * n + i] = x[a];
    store_defaults<D>(img, n, i, c.v, identity_mode);
}

struct XorShift {
    uint32_t x, y, z, w;
    __device__ __forceinline__ uint32_t next() {
        const uint32_t t = x ^ (x << 11);
        x = y;
        y = z;
        z = w;
        w = (w ^ (w >> 19)) ^ (t ^ (t >> 8));
        return w;
    }
};

__device__ __forceinline__ float nc_rand(XorShift& r) { return float(r.next()) * (1.0f / 4294967296.0f); }

// draws [t * K, min((t + 1) * K, count)) of the stream as floats
__global__ void k_gen_rand(const uint4* __restrict__ starts, int64_t count, float* __restrict__ out) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t k0 = t * kDrawsPerThread;
    if (k0 >= count) return;
    const uint4 s = starts[t];
    XorShift r{s.x, s.y, s.z, s.w};
    const int64_t k1 = min(count, k0 + kDrawsPerThread);
    for (int64_t k = k0; k < k1; ++k) out[k] = nc_rand(r);
}

// sample<D>(center) (src/nclr_math.h:141-147) for every centre: draw k sets axis k % D of
// particle k / D to (u * 2 - 1) * half + centre, in FP32 in that order.
template <int D>
__global__ void k_gen_sample(const uint4* __restrict__ starts, int64_t count, const float* __restrict__ centers,
                             int per_center, float half, int64_t n, float* __restrict__ img) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t k0 = t * kDrawsPerThread;
    if (k0 >= count) return;
    const uint4 s = starts[t];
    XorShift r{s.x, s.y, s.z, s.w};
    const int64_t k1 = min(count, k0 + kDrawsPerThread);
    for (int64_t k = k0; k < k1; ++k) {
        const int64_t p = k / D;
        const int a = int(k - p * D);
        const float u = nc_rand(r);
        const float c = centers[(p / per_center) * D + a];
        img[int64_t(a) * n + p] = (u * 2.0f - 1.0f) * half + c;
    }
}

template <int D>
__global__ void k_fill_defaults(int64_t n, int identity_mode, float* __restrict__ img) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) store_defaults<D>(img, n, i, nullptr, identity_mode);
}

// gyroid(k, t, pos) (src/nclr_math.h:131-139): sin/cos and the sums in double, rounded to float.
__device__ __forceinline__ float gyroid_at(double two_pi_k, float t, float x, float y, float z) {
    const double gx = two_pi_k * double(x), gy = two_pi_k * double(y), gz = two_pi_k * double(z);
    const double g = sin(gx) * cos(gy) + sin(gy) * cos(gz) + sin(gz) * cos(gx) - double(t);
    return float(g);
}

__global__ void k_eval_gyroid(int64_t n, const float* __restrict__ pos, double two_pi_k, float t,
                              float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = gyroid_at(two_pi_k, t, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

// Lattice point l of the box is kept where gyroid <= 0 (thickness == 0), else |gyroid| <= thickness.
__device__ __forceinline__ bool gyroid_keep(const mpm_cube& box, int64_t l, double two_pi_k, float t, float thickness,
                                            float (&x)[3]) {
    lattice_point<3>(box, l, x);
    const float g = gyroid_at(two_pi_k, t, x[0], x[1], x[2]);
    return thickness > 0.0f ? fabsf(g) <= thickness : g <= 0.0f;
}

// Block-wide exclusive prefix of `keep` (256 threads = 4 waves) and the block total.
__device__ __forceinline__ uint32_t block_prefix(bool keep, uint32_t& total) {
    __shared__ uint32_t s_wave[4];
    const uint64_t ballot = __ballot(keep);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t below = uint32_t(__popcll(ballot & ((uint64_t(1) << lane) - 1)));
    if (lane == 0) s_wave[wave] = uint32_t(__popcll(ballot));
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if (w < wave) before += s_wave[w];
        total += s_wave[w];
    }
    return before + below;
}

__global__ __launch_bounds__(256) void k_gyroid_count(mpm_cube box, int64_t points, double two_pi_k, float t,
                                                      float thickness, uint32_t* __restrict__ counts) {
    const int64_t l = int64_t(blockIdx.x) * 256 + threadIdx.x;
    float x[3];
    const bool keep = l < points && gyroid_keep(box, l, two_pi_k, t, thickness, x);
    uint32_t total;
    (void)block_prefix(keep, total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Exclusive scan of counts[0..m) into offsets[0..m], offsets[m] = the sum (one workgroup).
__global__ __launch_bounds__(1024) void k_scan_counts(const uint32_t* __restrict__ counts, int64_t m,
                                                      int64_t* __restrict__ offsets) {
    __shared__ int64_t s_wave[16];
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t base = 0; base < m; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < m ? int64_t(counts[i]) : 0;
        int64_t incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int64_t o = __shfl_up(incl, off);
            if (lane >= off) incl += o;
        }
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        int64_t before = s_carry;
        for (int w = 0; w < wave; ++w) before += s_wave[w];
        if (i < m) offsets[i] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[m] = s_carry;
}

__global__ __launch_bounds__(256) void k_gyroid_scatter(mpm_cube box, int64_t points, double two_pi_k, float t,
                                                        float thickness, const int64_t* __restrict__ offsets,
                                                        int identity_mode, int64_t n, float* __restrict__ img) {
    const int64_t l = int64_t(blockIdx.x) * 256 + threadIdx.x;
    float x[3];
    const bool keep = l < points && gyroid_keep(box, l, two_pi_k, t, thickness, x);
    uint32_t total;
    const uint32_t pre = block_prefix(keep, total);
    if (!keep) return;
    const int64_t i = offsets[blockIdx.x] + pre;
    if (i >= n) return;  // the host checks that n is the counted total
#pragma unroll
    for (int a = 0; a < 3; ++a) img[int64_t(a) * n + i] = x[a];
    store_defaults<3>(img, n, i, box.v, identity_mode);
}

// pt_3d_to_2d (src/nclr_math.h:154-167): (x0 c + x2 s + 0.5, x1 C + (x2 c - x0 s) S + 0.5) with
// c, s = cos, sin 28 deg and C, S = cos, sin 32 deg; `centered` applies it to x - 0.5 (the
// reference computes that difference and does not use it).
__global__ void k_project(int64_t n, const float* __restrict__ pos, float c, float s, float C, float S, int centered,
                          float* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p0 = pos[3 * i], p1 = pos[3 * i + 1], p2 = pos[3 * i + 2];
    if (centered) {
        p0 = p0 - 0.5f;
        p1 = p1 - 0.5f;
        p2 = p2 - 0.5f;
    }
    const float x = p0 * c + p2 * s;
    const float z = p2 * c - p0 * s;
    const float v = p1 * C + z * S;
    out[2 * i] = x + 0.5f;
    out[2 * i + 1] = v + 0.5f;
}

}  // namespace nclr_dev

namespace nclr_host {

namespace {

unsigned blocks(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

// 128-bit xorshift128 state as 4 words {x, y, z, w}; the transition is linear over GF(2).
struct State128 {
    uint32_t w[4];
};

State128 step(const State128& s) {
    const uint32_t t = s.w[0] ^ (s.w[0] << 11);
    return {{s.w[1], s.w[2], s.w[3], (s.w[3] ^ (s.w[3] >> 19)) ^ (t ^ (t >> 8))}};
}

// A linear map on states, by its images of the 128 basis vectors.
struct Map128 {
    State128 col[128];
    State128 apply(const State128& s) const {
        State128 r{{0, 0, 0, 0}};
        for (int b = 0; b < 128; ++b)
            if ((s.w[b >> 5] >> (b & 31)) & 1u)
                for (int k = 0; k < 4; ++k) r.w[k] ^= col[b].w[k];
        return r;
    }
};

Map128 transition() {
    Map128 m;
    for (int b = 0; b < 128; ++b) {
        State128 e{{0, 0, 0, 0}};
        e.w[b >> 5] = 1u << (b & 31);
        m.col[b] = step(e);
    }
    return m;
}

Map128 compose(const Map128& a, const Map128& b) {  // a after b
    Map128 m;
    for (int k = 0; k < 128; ++k) m.col[k] = a.apply(b.col[k]);
    return m;
}

}  // namespace

void rand_jump(uint32_t state[4], uint64_t k) {
    State128 s{{state[0], state[1], state[2], state[3]}};
    if (k < 512) {
        for (uint64_t i = 0; i < k; ++i) s = step(s);
    } else {
        Map128 p = transition();  // T^(2^b)
        for (; k; k >>= 1) {
            if (k & 1u) s = p.apply(s);
            if (k > 1) p = compose(p, p);
        }
    }
    for (int i = 0; i < 4; ++i) state[i] = s.w[i];
}

// Start states of the threads of a generator over `count` draws from `state`.
std::vector<uint4> rand_starts(const uint32_t state[4], int64_t count) {
    const int64_t threads = (count + nclr_dev::kDrawsPerThread - 1) / nclr_dev::kDrawsPerThread;
    std::vector<uint4> out(size_t(std::max<int64_t>(threads, 0)));
    Map128 j = transition();
    for (int b = 0; (1 << b) < nclr_dev::kDrawsPerThread; ++b) j = compose(j, j);  // T^K
    State128 s{{state[0], state[1], state[2], state[3]}};
    for (int64_t t = 0; t < threads; ++t) {
        out[size_t(t)] = make_uint4(s.w[0], s.w[1], s.w[2], s.w[3]);
        if (t + 1 < threads) s = j.apply(s);
    }
    return out;
}

hipError_t launch_gen_cubes(int dim, const mpm_cube* cubes, const int64_t* offsets, int ncubes, int identity_mode,
                            int64_t n, float* img, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (dim == 2) nclr_dev::k_gen_cubes<2><<<blocks(n, 256), 256, 0, s>>>(cubes, offsets, ncubes, identity_mode, n, img);
    else nclr_dev::k_gen_cubes<3><<<blocks(n, 256), 256, 0, s>>>(cubes, offsets, ncubes, identity_mode, n, img);
    return hipGetLastError();
}

hipError_t launch_gen_rand(const uint4* starts, int64_t count, float* out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int64_t threads = (count + nclr_dev::kDrawsPerThread - 1) / nclr_dev::kDrawsPerThread;
    nclr_dev::k_gen_rand<<<blocks(threads, 256), 256, 0, s>>>(starts, count, out);
    return hipGetLastError();
}

hipError_t launch_gen_sample(int dim, const uint4* starts, const float* centers, int per_center, float half,
                             int identity_mode, int64_t n, float* img, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t count = n * dim;
    const int64_t threads = (count + nclr_dev::kDrawsPerThread - 1) / nclr_dev::kDrawsPerThread;
    if (dim == 2) {
        nclr_dev::k_fill_defaults<2><<<blocks(n, 256), 256, 0, s>>>(n, identity_mode, img);
        nclr_dev::k_gen_sample<2><<<blocks(threads, 256), 256, 0, s>>>(starts, count, centers, per_center, half, n, img);
    } else {
        nclr_dev::k_fill_defaults<3><<<blocks(n, 256), 256, 0, s>>>(n, identity_mode, img);
        nclr_dev::k_gen_sample<3><<<blocks(threads, 256), 256, 0, s>>>(starts, count, centers, per_center, half, n, img);
    }
    return hipGetLastError();
}

hipError_t launch_eval_gyroid(int64_t n, const float* pos, float k, float t, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    nclr_dev::k_eval_gyroid<<<blocks(n, 256), 256, 0, s>>>(n, pos, (2.0 * M_PI) / double(k), t, out);
    return hipGetLastError();
}

int64_t gyroid_blocks(const mpm_cube& box) {
    return (int64_t(box.res[0]) * box.res[1] * box.res[2] + 255) / 256;
}

hipError_t launch_gyroid_count(const mpm_cube& box, float k, float t, float thickness, uint32_t* counts,
                               int64_t* offsets, hipStream_t s) {
    const int64_t points = int64_t(box.res[0]) * box.res[1] * box.res[2];
    const int64_t nb = gyroid_blocks(box);
    if (nb == 0) return hipSuccess;
    nclr_dev::k_gyroid_count<<<unsigned(nb), 256, 0, s>>>(box, points, (2.0 * M_PI) / double(k), t, thickness, counts);
    nclr_dev::k_scan_counts<<<1, 1024, 0, s>>>(counts, nb, offsets);
    return hipGetLastError();
}

hipError_t launch_gyroid_scatter(const mpm_cube& box, float k, float t, float thickness, const int64_t* offsets,
                                 int identity_mode, int64_t n, float* img, hipStream_t s) {
    const int64_t points = int64_t(box.res[0]) * box.res[1] * box.res[2];
    const int64_t nb = gyroid_blocks(box);
    if (nb == 0 || n == 0) return hipSuccess;
    nclr_dev::k_gyroid_scatter<<<unsigned(nb), 256, 0, s>>>(box, points, (2.0 * M_PI) / double(k), t, thickness,
                                                            offsets, identity_mode, n, img);
    return hipGetLastError();
}

hipError_t launch_project(int64_t n, const float* pos, int centered, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    // to_radians (src/nclr_math.h:149-152) and the float cos / sin of the reference, on the host
    constexpr float kDegToRad = float(M_PI / 180.0);
    const float phi = 28.0f * kDegToRad, theta = 32.0f * kDegToRad;
    nclr_dev::k_project<<<blocks(n, 256), 256, 0, s>>>(n, pos, std::cos(phi), std::sin(phi), std::cos(theta),
                                                       std::sin(theta), centered, out);
    return hipGetLastError();
}

}  // namespace nclr_host
