// The compact step: many steps of one small scene in one cooperative launch.
//
// A scene whose particles stay inside a box of a few thousand grid nodes does not need the tile
// sort: every workgroup takes a fixed range of particles (in slot order), deposits them into a
// fixed-point copy of the box's nodes in its own LDS (only its sub-box: the nodes its particles'
// stencils reach), and stores that sub-box as float4 {momentum, mass} in its slot.  After one
// grid-wide barrier every workgroup sums, for each node of its own sub-box, the slots of the
// workgroups whose sub-boxes hold the node, in workgroup order, applies grid_op
// (src/nclr.h:263-310) into LDS and runs G2P (src/nclr.h:312-360) on its particles in place from
// there.  A node's sum has the same terms in the same order in every workgroup that forms it, so
// the step is deterministic.  One barrier per step, one launch for all steps: each workgroup
// publishes its sub-box and flags in a table entry stamped with the barrier's count, and the wait for
// every entry is also their read.  Slots alternate by step parity and the tables rotate over three,
// so a workgroup may deposit step s + 1 while another still reads step s.
//
// The box of step s: the stencil bases of the particles before step s lie within one cell of
// those before step s - 1 (a particle moves less than one cell per step: the grid velocity is
// clamped to 0.9 dx / dt per component, src/nclr.h:285, and G2P's weights are non-negative and
// sum to one), so the box of step s is the union of the sub-boxes of step s - 1 (every workgroup
// reads them after the barrier) widened by one node; the first step's box is the exact range, a
// grid-wide reduction before the first step.  When the box outgrows the LDS, every workgroup stops
// at the same step (they all hold the same union) and the host continues with the tiled step;
// when the whole grid fits the LDS a particle outside the box (never expected) retries its step
// over the whole grid instead.  A particle off its grid stops the scene as the tiled step does: the
// step that finds it leaves every particle as it was, the grid keeps the last completed step and
// the stamp still counts every step of the launch.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <climits>

#include "mpm_device.hpp"
#include "mpm_step_common.hpp"

namespace nclr_dev {

// Four fixed-point channel planes per node (momentum x, y, z, mass; z unused in 2D, as in the
// tiles) and a flag word: 36 bytes per node, so a box holds up to 4266 nodes (65^2 in 2D, 16^3
// in 3D).  After the deposits the same LDS holds the box's updated grid (16 bytes per node).
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

// Grid-wide barrier of the compact step: an arrival counter and a generation word (zeroed before
// the launch); the k-th barrier's last arrival publishes generation k.  Thread 0 of each
// workgroup fences (release before its arrival, acquire after the wait): the XCDs' L2 caches are
// not coherent with each other.  The wait is bounded: past ~2 s it raises kStatusBarrierTimeout
// and returns false, and the caller leaves the step loop (so a launch always ends).
constexpr uint32_t kBarrierSpinLimit = 1u << 26;

NCLR_DI bool compact_barrier(const CompactArgs& a, uint32_t& k, uint32_t* s_flag) {
    ++k;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t arrived = atomicAdd(&a.bar[0], 1u) + 1u;
        uint32_t ok = 1u;
        if (arrived == k * gridDim.x) {
            __hip_atomic_store(&a.bar[1], k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&a.bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < k) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kBarrierSpinLimit) {
                    atomicOr(a.status, kStatusBarrierTimeout);
                    ok = 0u;
                    break;
                }
            }
        }
        __threadfence();
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0u;
}

__global__ void k_compact_init(int* __restrict__ box, uint32_t* __restrict__ bar, int* __restrict__ info,
                               unsigned long long* __restrict__ subbox, int dim) {
    const int t = int(threadIdx.x);
    // stamp 0 in every table word: no step of the launch matches it (its stamps count from 1)
    for (int r = t; r < kCompactTableWords; r += blockDim.x) subbox[r] = 0ull;
    if (t < kInfoWords) info[t] = 0;
    if (t < 2) bar[t] = 0u;
    if (t < dim) {  // the base range before the first step
        box[t] = INT_MAX;
        box[dim + t] = INT_MIN;
    }
}

// Nodes [lo, hi] (absolute, per axis) of a sub-box; empty when lo > hi on an axis.
template <int D>
NCLR_DI bool sub_box_has(const int* sb, const int (&n)[D]) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) in = in && n[d] >= sb[d] && n[d] <= sb[3 + d];
    return in;
}

template <int D>
NCLR_DI int sub_box_nodes(const int* sb) {
    int c = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) c *= max(0, sb[3 + d] - sb[d] + 1);
    return c;
}

// The idx-th node of a sub-box (last axis fastest): its absolute coordinates and its index in the
// box (origin lo, extent ext).
template <int D>
NCLR_DI int sub_box_node(const int* sb, int idx, const int (&lo)[D], const int (&ext)[D], int (&n)[D]) {
    int rem = idx;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        const int e = sb[3 + d] - sb[d] + 1;
        n[d] = sb[d] + rem % e;
        rem /= e;
    }
    int r = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) r = r * ext[d] + (n[d] - lo[d]);
    return r;
}

// Particle of thread t in round k of workgroup chunk p0 (256 per round).  The state is in Hilbert
// order of the particles' cells, so neighbouring slots share cells; consecutive lanes take slots
// four apart, so a wave's 64 particles spread over the whole round and its LDS atomics meet fewer
// equal addresses.
NCLR_DI int64_t compact_particle(int64_t p0, int k, int t) { return p0 + int64_t(k) * 256 + ((t & 63) * 4 + (t >> 6)); }

// One particle's state.
template <int D>
struct PState {
    float x[D], v[D];
    M<D> F, C;
    float Jp, mass, vol;
};

template <int D>
NCLR_DI void load_pstate(PState<D>& q, const float* __restrict__ st, int64_t s, int64_t i) {
    using L = Layout<D>;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        q.x[d] = ldf(st, s, L::X + d, i);
        q.v[d] = ldf(st, s, L::V + d, i);
    }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            q.F.a[r][c] = ldf(st, s, L::F + r * D + c, i);
            q.C.a[r][c] = ldf(st, s, L::C + r * D + c, i);
        }
    q.Jp = ldf(st, s, L::JP, i);
    q.mass = ldf(st, s, L::MASS, i);
    q.vol = ldf(st, s, L::VOL, i);
}

// The fields a step changes (x, v, F, C, Jp).
template <int D>
NCLR_DI void store_pstate(const PState<D>& q, float* __restrict__ st, int64_t s, int64_t i) {
    using L = Layout<D>;
    st[int64_t(L::JP) * s + i] = q.Jp;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        st[int64_t(L::X + d) * s + i] = q.x[d];
        st[int64_t(L::V + d) * s + i] = q.v[d];
    }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            st[int64_t(L::C + r * D + c) * s + i] = q.C.a[r][c];
            st[int64_t(L::F + r * D + c) * s + i] = q.F.a[r][c];
        }
}

// In 2D every thread keeps its particles' state in registers for the whole launch (loaded before
// the first step, stored after the last); in 3D (25 floats a particle) it loads and stores them
// every step.
template <int D, int MAT, int MODE, int PPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_compact_steps(CompactArgs a) {
    using CP = Compact<D>;
    constexpr bool RESIDENT = D == 2;
    extern __shared__ unsigned long long lds[];
    __shared__ float s_bound[4][4];
    __shared__ int s_red[4][6];                      // per wave: a range of bases or nodes
    __shared__ int s_sb[7];                          // this workgroup's sub-box and flags
    __shared__ int s_tab[kCompactMaxWorkgroups][6];  // sub-boxes of the overlapping (or all) workgroups
    __shared__ int s_list[kCompactMaxWorkgroups];    // their workgroup indices, ascending
    __shared__ int s_wcount[4];
    __shared__ int s_wflags[4];                      // per wave: off the grid (1), outside the box (2)
    __shared__ int s_box[2 * D];
    __shared__ uint32_t s_flag;
    const SceneParams sp = a.sp;
    const int material = MAT == kMixed ? sp.material : MAT;
    const int nwg = int(gridDim.x), wg = int(blockIdx.x);
    const int64_t s = a.stride;
    const int64_t p0 = int64_t(wg) * 256 * PPT;
    float* st = a.st;
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    uint32_t k_bar = 0;
    // the step barriers of this launch (the stamp of their table entries) and the table of the last
    // completed step
    uint32_t kb = 0;
    int last_tab = 0;
    // a scene stopped before the launch (read before the first barrier: no step has run yet)
    const bool pre_stopped = (a.step.scene_flags[0] | a.step.scene_frozen[0]) != 0u;
    int64_t whole_nodes = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) whole_nodes *= sp.res + 1;
    const bool whole_fits = whole_nodes <= CP::CAP;
    PState<D> ps[RESIDENT ? PPT : 1];

    // the exact base range of the particles before the first step (a grid-wide reduction)
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
            const int64_t i = compact_particle(p0, k, int(threadIdx.x));
            if (i >= a.n) continue;
            PState<D> q;
            load_pstate<D>(q, st, s, i);
            if constexpr (RESIDENT) ps[k] = q;
            float fx[D], w[D][3];
            int base[D];
            if (!stencil<D>(q.x, sp, base, fx, w)) continue;  // found again by the first step
#pragma unroll
            for (int d = 0; d < D; ++d) {
                lo[d] = min(lo[d], base[d]);
                hi[d] = max(hi[d], base[d]);
            }
        }
        box_accumulate<D>(a.box, s_box, lo, hi);
    }
    bool alive = compact_barrier(a, k_bar, &s_flag);
    // node range of the next step's box (empty: no particle on the grid)
    int nlo[D], nhi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        nlo[d] = a.box[d];
        nhi[d] = a.box[d] <= a.box[D + d] ? a.box[D + d] + 2 : a.box[D + d];
    }

    int done = 0, reason = kCompactDone;
    bool whole = false;
    int lo[D], ext[D];
    int glo[D], gext[D], gnodes = 0;  // box of the last completed step
    uint32_t gstamp = 0;               // its stamp (workgroup 0, thread 0)
#pragma unroll
    for (int d = 0; d < D; ++d) {
        glo[d] = 0;
        gext[d] = 1;
    }
    for (int step = 0; alive && step < a.nsteps;) {
        const int par = step & 1;
        // ---- the box of this step --------------------------------------------------------------
        int nodes = 1;
        bool empty = false;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            lo[d] = whole ? 0 : max(0, nlo[d]);
            const int hi = whole ? sp.res : min(sp.res, nhi[d]);
            empty = empty || hi < lo[d];
            ext[d] = hi >= lo[d] ? hi - lo[d] + 1 : 1;
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
        float4* slots = a.slots + (int64_t(par) * a.slot_sets + wg) * CP::CAP;

        // ---- P2G: the particles' affine momenta, bounds and base range ------------------------
        bool valid[PPT], finite[PPT];
        int rb[PPT];
        float fxs[PPT][D], mvs[PPT][D], ms[PPT];
        M<D> As[PPT];
        float bound[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int blo[D], bhi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            blo[d] = INT_MAX;
            bhi[d] = INT_MIN;
        }
        bool off_grid = false, outside = false;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int64_t i = compact_particle(p0, k, int(threadIdx.x));
            valid[k] = false;
            finite[k] = true;
            rb[k] = 0;
            ms[k] = 0.0f;
            if (i >= a.n) continue;
            PState<D> q;
            if constexpr (RESIDENT) q = ps[k];
            else load_pstate<D>(q, st, s, i);
            float fx[D], w[D][3];
            int base[D];
            if (!stencil<D>(q.x, sp, base, fx, w)) {
                off_grid = true;
                continue;
            }
            bool inside = true;
            int r = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int qd = base[d] - lo[d];
                inside = inside && qd >= 0 && qd + 2 < ext[d];
                r = r * ext[d] + qd;
            }
            if (!inside) {
                outside = true;
                continue;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                blo[d] = min(blo[d], base[d]);
                bhi[d] = max(bhi[d], base[d]);
            }
            valid[k] = true;
            rb[k] = r;
            As[k] = fused_affine<D, MODE>(material, q.F, q.C, q.Jp, q.mass, q.vol, sp);
            ms[k] = q.mass;
            float wsum = 1.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                fxs[k][d] = fx[d];
                mvs[k][d] = q.v[d] * q.mass;
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
            const float bm = wsum * fabsf(q.mass);
            finite[k] = finite[k] && bm < 3.0e38f;
            if (bm < 3.0e38f) bound[3] = fmaxf(bound[3], bm);
        }
        if (off_grid) stop_scene(a.step, 0, kStatusOutOfRange);  // for the host; the table carries it too
        if (outside) atomicOr(&a.info[kInfoOutside], 1);
        {
            const unsigned long long mo = __ballot(off_grid), mu = __ballot(outside);
            if (lane == 0) s_wflags[wave] = (mo != 0ull ? 1 : 0) | (mu != 0ull ? 2 : 0);
        }
        // workgroup maxima of the bounds and the base range (this workgroup's sub-box)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float b = bound[c];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) b = fmaxf(b, __shfl_xor(b, off));
            bound[c] = b;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                blo[d] = min(blo[d], __shfl_xor(blo[d], off));
                bhi[d] = max(bhi[d], __shfl_xor(bhi[d], off));
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) s_bound[wave][c] = bound[c];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                s_red[wave][d] = blo[d];
                s_red[wave][3 + d] = bhi[d];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int d = 0; d < 3; ++d) {
                int l = 0, h = 0;  // axes beyond D: one node
                if (d < D) {
                    l = min(min(s_red[0][d], s_red[1][d]), min(s_red[2][d], s_red[3][d]));
                    h = max(max(s_red[0][3 + d], s_red[1][3 + d]), max(s_red[2][3 + d], s_red[3][3 + d]));
                    if (l <= h) {
                        h += 2;  // the stencil's three nodes
                    } else {     // no particle: the empty range [1, 0]
                        l = 1;
                        h = 0;
                    }
                }
                s_sb[d] = l;
                s_sb[3 + d] = h;
            }
            s_sb[6] = s_wflags[0] | s_wflags[1] | s_wflags[2] | s_wflags[3];
        }
        float scale[4];
        double inv_scale[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            fixed_point_scale(fmaxf(fmaxf(s_bound[0][c], s_bound[1][c]), fmaxf(s_bound[2][c], s_bound[3][c])), scale[c],
                              inv_scale[c]);
        __syncthreads();
        int sb[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) sb[e] = s_sb[e];
        const int sb_nodes = sub_box_nodes<D>(sb);

        // ---- zero the sub-box, deposit, store it in this workgroup's slot -----------------------
        for (int idx = threadIdx.x; idx < sb_nodes; idx += 256) {
            int n[D];
            const int r = sub_box_node<D>(sb, idx, lo, ext, n);
#pragma unroll
            for (int c = 0; c < CP::CH; ++c) acc[c * nodes + r] = 0ull;
            nonfinite[r] = 0u;
        }
        __syncthreads();
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
            const float nan = __builtin_nanf("");
            for (int idx = threadIdx.x; idx < sb_nodes; idx += 256) {
                int n[D];
                const int r = sub_box_node<D>(sb, idx, lo, ext, n);
                float out[4];
                const uint32_t bad = nonfinite[r];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    out[c] = (D == 2 && c == 2) ? 0.0f
                                                : ((bad >> c) & 1u ? nan : fixed_point_value(acc[c * nodes + r], inv_scale[c]));
                slots[r] = make_float4(out[0], out[1], out[2], out[3]);
            }
        }

        // ---- the step barrier: publish this workgroup's table entry, wait for every entry -------
        // Each workgroup writes its sub-box and flags into the table of this barrier's parity, every
        // 32-bit half of the entry stamped with the barrier's count (after a release fence, so its
        // slots are visible first), and threads t < nwg poll entry t until all its halves carry the
        // stamp.  The wait and the read of the entries are one round trip; a half with another
        // stamp is an older barrier's (or the launch's zeroed table).  Three tables rotate with the
        // barrier count: a workgroup one barrier ahead never overwrites an entry another still
        // polls, and the last completed step's table survives the (at most two) barriers that can
        // follow it, a step found outside its box and its retry, for the export after the loop.
        // Bounded like compact_barrier: past ~2 s it raises kStatusBarrierTimeout.
        ++kb;
        const int tab = int(kb % 3u);
        const uint32_t stamp16 = kb & 0xffffu;
        __syncthreads();  // every slot store of the workgroup is issued
        if (threadIdx.x == 0) {
            __threadfence();
            unsigned long long* e = a.subbox + (int64_t(tab) * kCompactMaxWorkgroups + wg) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t v0 = uint32_t(s_sb[2 * q]) & 0xffffu;
                const uint32_t v1 = 2 * q + 1 < 7 ? uint32_t(s_sb[2 * q + 1]) & 0xffffu : 0u;
                const unsigned long long word = (static_cast<unsigned long long>((stamp16 << 16) | v1) << 32) |
                                                ((stamp16 << 16) | v0);
                __hip_atomic_store(&e[q], word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const int t = int(threadIdx.x);
        int tb[7] = {1, 1, 1, 0, 0, 0, 0};  // empty
        bool polled = true;
        if (t < nwg) {
            const unsigned long long* e = a.subbox + (int64_t(tab) * kCompactMaxWorkgroups + t) * 4;
            uint32_t spins = 0;
            for (;;) {
                unsigned long long v[4];
                bool all = true;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    v[q] = __hip_atomic_load(&e[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    all = all && (uint32_t(v[q]) >> 16) == stamp16 && (uint32_t(v[q] >> 32) >> 16) == stamp16;
                }
                if (all) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        tb[2 * q] = int(uint32_t(v[q]) & 0xffffu);
                        if (2 * q + 1 < 7) tb[2 * q + 1] = int(uint32_t(v[q] >> 32) & 0xffffu);
                    }
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kBarrierSpinLimit) {
                    atomicOr(a.status, kStatusBarrierTimeout);
                    polled = false;
                    break;
                }
            }
            __threadfence();  // acquire: workgroup t's slots after its entry
        }
        alive = __syncthreads_and(polled ? 1 : 0) != 0;
        if (threadIdx.x == 0) __threadfence();
        __syncthreads();
        if (!alive) break;

        // ---- every workgroup's sub-box: the overlap list (ascending: the order of the sums) and
        //      their union, the stencil nodes before this step; the next step's box is the union
        //      widened by one node (a particle moves less than one cell per step); and the flags of
        //      all workgroups ---------------------------------------------------------------------
        int flags = 0;
        {
            bool hit = false;
            int ul[3] = {INT_MAX, INT_MAX, INT_MAX}, uh[3] = {INT_MIN, INT_MIN, INT_MIN};
            if (t < nwg) {
                bool nonempty = true;
#pragma unroll
                for (int d = 0; d < 3; ++d) nonempty = nonempty && tb[d] <= tb[3 + d];
                hit = nonempty;
#pragma unroll
                for (int d = 0; d < 3; ++d) hit = hit && tb[d] <= sb[3 + d] && sb[d] <= tb[3 + d];
                if (nonempty)
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        ul[d] = tb[d];
                        uh[d] = tb[3 + d];
                    }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    ul[d] = min(ul[d], __shfl_xor(ul[d], off));
                    uh[d] = max(uh[d], __shfl_xor(uh[d], off));
                }
            }
            const unsigned long long m = __ballot(hit);
            const unsigned long long mo = __ballot((tb[6] & 1) != 0), mu = __ballot((tb[6] & 2) != 0);
            if (lane == 0) {
                s_wcount[wave] = __popcll(m);
                s_wflags[wave] = (mo != 0ull ? 1 : 0) | (mu != 0ull ? 2 : 0);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    s_red[wave][d] = ul[d];
                    s_red[wave][3 + d] = uh[d];
                }
            }
            __syncthreads();
            flags = s_wflags[0] | s_wflags[1] | s_wflags[2] | s_wflags[3];
            int pos = __popcll(m & ((1ull << lane) - 1ull));
            for (int w = 0; w < wave; ++w) pos += s_wcount[w];
            if (hit) {
                s_list[pos] = t;
#pragma unroll
                for (int e = 0; e < 6; ++e) s_tab[pos][e] = tb[e];
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int l = min(min(s_red[0][d], s_red[1][d]), min(s_red[2][d], s_red[3][d]));
                const int h = max(max(s_red[0][3 + d], s_red[1][3 + d]), max(s_red[2][3 + d], s_red[3][3 + d]));
                nlo[d] = l <= h ? l - 1 : 1;
                nhi[d] = l <= h ? h + 1 : 0;
            }
            __syncthreads();
        }

        // ---- stop or retry (the same decision in every workgroup: the same entries) -------------
        if (pre_stopped || (flags & 1)) {
            // a particle off its grid: the scene stops with every particle as before this step;
            // the stamp counts this and the remaining steps, as the tiled step's does
            if (wg == 0 && threadIdx.x == 0) {
                if (a.step.scene_frozen[0] == 0u) a.step.scene_frozen[0] = *a.stamp + 1u;
                *a.stamp = *a.stamp + uint32_t(a.nsteps - step);
            }
            reason = kCompactStopped;
            break;
        }
        if (!whole && (flags & 2)) {
            // a particle outside the box
            if (!whole_fits) {
                reason = kCompactOverflow;
                break;
            }
            whole = true;
            continue;
        }
        if (wg == 0 && threadIdx.x == 0) {
            gstamp = *a.stamp + 1u;
            *a.stamp = gstamp;
        }
        const int nlist = s_wcount[0] + s_wcount[1] + s_wcount[2] + s_wcount[3];

        // ---- grid update of this workgroup's sub-box, into LDS ------------------------------------
        float4* lgrid = reinterpret_cast<float4*>(lds);  // the box's nodes (the accumulators are done)
        const float4* slot0 = a.slots + int64_t(par) * a.slot_sets * CP::CAP;
        for (int idx = threadIdx.x; idx < sb_nodes; idx += 256) {
            int n[D];
            const int r = sub_box_node<D>(sb, idx, lo, ext, n);
            float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (int j0 = 0; j0 < nlist; j0 += 8) {
                // eight loads in flight, then their sums in order (terms of workgroups whose
                // sub-box misses the node are absent)
                float4 t[8];
                bool in[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = j0 + u;
                    in[u] = j < nlist && sub_box_has<D>(s_tab[j], n);
                    t[u] = in[u] ? slot0[int64_t(s_list[j]) * CP::CAP + r] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (in[u]) {
                        g.x += t[u].x;
                        g.y += t[u].y;
                        g.z += t[u].z;
                        g.w += t[u].w;
                    }
                }
            }
            lgrid[r] = grid_node<D>(g, n, sp);
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < D; ++d) {
            glo[d] = lo[d];
            gext[d] = ext[d];
        }
        gnodes = nodes;
        last_tab = tab;

        // ---- G2P from LDS ---------------------------------------------------------------------------
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int64_t i = compact_particle(p0, k, int(threadIdx.x));
            if (i >= a.n || !valid[k]) continue;  // every particle of the scene is valid here
            PState<D> q;
            if constexpr (RESIDENT) q = ps[k];
            else load_pstate<D>(q, st, s, i);
            float fx[D], w[D][3];
            int base[D];
            (void)stencil<D>(q.x, sp, base, fx, w);  // on the grid: P2G found it there
            const float4* g0 = lgrid + rb[k];
            auto node = [&](const int (&off)[D]) {
                const int o = D == 3 ? off[0] * rs.s0 + off[1] * rs.s1 + off[D - 1] : off[0] * rs.s0 + off[1];
                return g0[o];
            };
            g2p_particle<D, MODE>(material, sp, fx, w, node, q.x, q.v, q.C, q.F, q.Jp);
            if constexpr (RESIDENT) ps[k] = q;
            else store_pstate<D>(q, st, s, i);
        }
        ++step;
        done = step;
    }
    if constexpr (RESIDENT) {
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int64_t i = compact_particle(p0, k, int(threadIdx.x));
            if (i < a.n) store_pstate<D>(ps[k], st, s, i);
        }
    }

    // ---- the grid of the last completed step over its whole box, for export ----------------------
    if (done > 0 && alive && compact_barrier(a, k_bar, &s_flag)) {
        const int par = (done - 1) & 1;
        if (int(threadIdx.x) < nwg) {
            const unsigned long long* e = a.subbox + (int64_t(last_tab) * kCompactMaxWorkgroups + threadIdx.x) * 4;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const unsigned long long v = e[q];
                s_tab[threadIdx.x][2 * q] = int(uint32_t(v) & 0xffffu);
                s_tab[threadIdx.x][2 * q + 1] = int(uint32_t(v >> 32) & 0xffffu);
            }
        }
        __syncthreads();
        const float4* slot0 = a.slots + int64_t(par) * a.slot_sets * CP::CAP;
        for (int r = wg * 256 + int(threadIdx.x); r < gnodes; r += nwg * 256) {
            int n[D];
            {
                int rem = r;
#pragma unroll
                for (int d = D - 1; d >= 0; --d) {
                    n[d] = glo[d] + rem % gext[d];
                    rem /= gext[d];
                }
            }
            float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (int q = 0; q < nwg; ++q) {
                if (!sub_box_has<D>(s_tab[q], n)) continue;
                const float4 t = slot0[int64_t(q) * CP::CAP + r];
                g.x += t.x;
                g.y += t.y;
                g.z += t.z;
                g.w += t.w;
            }
            a.massive[r] = g.w > 0.0f ? uint8_t(1) : uint8_t(0);
            a.grid[r] = grid_node<D>(g, n, sp);
        }
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

// Hilbert order of the particles' cells (the slot order of a compact launch): cells next to each
// other on the curve are neighbours on the grid, so a run of slots covers a compact patch of
// cells.  The index follows J. Skilling, "Programming the Hilbert curve" (AIP Conf. Proc. 707,
// 2004): the axes are transformed in place into the curve's "transpose" form, whose bits
// interleaved (axis 0 most significant within each level) are the index.  Cells are clamped to
// [0, 2^bits); positions off the grid (or NaN) take the last key.
template <int D>
NCLR_DI uint32_t hilbert_index(uint32_t (&X)[D], int bits) {
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1u; Q >>= 1) {  // inverse undo
        const uint32_t P = Q - 1u;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
#pragma unroll
    for (int i = 1; i < D; ++i) X[i] ^= X[i - 1];  // Gray encode
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1u; Q >>= 1)
        if (X[D - 1] & Q) t ^= Q - 1u;
#pragma unroll
    for (int i = 0; i < D; ++i) X[i] ^= t;
    uint32_t key = 0;
    for (int b = bits - 1; b >= 0; --b)
#pragma unroll
        for (int i = 0; i < D; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
    return key;
}

template <int D>
__global__ __launch_bounds__(256) void k_cell_keys(const float* __restrict__ st, int64_t stride, int64_t n,
                                                   float inv_dx, int bits_per_axis, uint32_t* __restrict__ keys) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float top = float((1 << bits_per_axis) - 1);
    uint32_t X[D];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float c = floorf(st[int64_t(Layout<D>::X + d) * stride + i] * inv_dx);
        ok = ok && c >= 0.0f && c <= top;  // false for NaN
        X[d] = ok ? uint32_t(c) : 0u;
    }
    keys[i] = ok ? hilbert_index<D>(X, bits_per_axis) : 0xffffffffu >> (32 - D * bits_per_axis);
}

// out[f][j] = in[f][perm[j]] for every field, and the slot -> particle map likewise.
__global__ __launch_bounds__(256) void k_gather_state(const float* __restrict__ in, const uint32_t* __restrict__ pid_in,
                                                      const uint32_t* __restrict__ perm, float* __restrict__ out,
                                                      uint32_t* __restrict__ pid_out, int fields, int64_t stride,
                                                      int64_t n) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t i = perm[j];
    pid_out[j] = pid_in[i];
    for (int f = 0; f < fields; ++f) out[int64_t(f) * stride + j] = in[int64_t(f) * stride + i];
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
    k_compact_init<<<1, 256, 0, s>>>(a.box, a.bar, a.info, a.subbox, D);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    CompactArgs args = a;
    void* params[] = {&args};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(unsigned(nwg)), dim3(256), params,
                                      unsigned(kCompactLdsBytes), s);
}

template <int D, int MAT, int MODE>
static hipError_t compact_ppt(const CompactArgs& a, int ppt, int nwg, hipStream_t s) {
    // 3D: 1 or 2 particles per thread (4 would spill, and so would 1 in the accurate mode: that
    // launch runs 2 per thread over the same workgroups, the second half of them without particles)
    if constexpr (!(D == 3 && MODE == kSvdAccurate))
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
    if (!ppt_ok || nwg < 1 || nwg > kCompactMaxWorkgroups || int64_t(nwg) * 256 * ppt < a.n || a.n <= 0 ||
        a.sp.res > kCompactMaxRes)
        return hipErrorInvalidValue;
    return dim == 2 ? compact_dim<2>(a, material, mode, ppt, nwg, s) : compact_dim<3>(a, material, mode, ppt, nwg, s);
}

hipError_t launch_compact_order(int dim, const float* st, const uint32_t* pid, float* st_out, uint32_t* pid_out,
                                int64_t n, int64_t stride, int fields, const SceneParams& sp, uint32_t* keys,
                                uint32_t* perm, uint32_t* keys_tmp, uint32_t* perm_tmp, uint32_t* hist,
                                hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int bpa = 1;
    while ((1 << bpa) < sp.res + 1 && bpa < 16) ++bpa;
    bpa = dim == 3 ? (bpa < 10 ? bpa : 10) : bpa;  // at most 32 key bits
    const unsigned b = unsigned((n + 255) / 256);
    if (dim == 2) k_cell_keys<2><<<b, 256, 0, s>>>(st, stride, n, sp.inv_dx, bpa, keys);
    else k_cell_keys<3><<<b, 256, 0, s>>>(st, stride, n, sp.inv_dx, bpa, keys);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(keys, perm, keys_tmp, perm_tmp, hist, n, dim * bpa, true, s);
    if (e != hipSuccess) return e;
    k_gather_state<<<b, 256, 0, s>>>(st, pid, perm, st_out, pid_out, fields, stride, n);
    return hipGetLastError();
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
