// C ABI of the MLS-MPM engine (declarations and contracts: include/nclr_mpm.h).
//
// One handle = one batch of scenes (a single simulation is a batch of one) = one HIP stream.
// The handle owns two ping-pong buffers of structure-of-arrays particle state (G2P writes the
// state in tile order into the other buffer), the original particle index of every slot, the
// sort buffers, the per-work-item P2G slots, the tile-blocked updated grid, the device step
// stamp, and a device status word that kernels set when a particle leaves the grid.  The status
// is read back (and the error reported) only by synchronising calls, so mpm_advance() never
// waits on the device.  With opts.graph each step is one replay of a captured hipGraph (one
// graph per ping-pong parity; every other kernel argument is step-invariant).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mpm_internal.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(MPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RC_TRY(expr)                       \
    do {                                   \
        const int rc_ = (expr);            \
        if (rc_ != MPM_OK) return rc_;     \
    } while (0)

enum KernelId {
    kSort = 0,
    kP2G,
    kGrid,
    kG2P,
    kStep,
    kG2PAdj,
    kGridAdj,
    kP2GAdj,
    kBackward,
    kNumKernels
};
const char* const kKernelNames[kNumKernels] = {"sort",        "p2g",          "grid",        "g2p",     "step",
                                               "g2p_adjoint", "grid_adjoint", "p2g_adjoint", "backward"};

struct PendingEvent {
    int kernel;
    hipEvent_t start, stop;
};

struct KernelTimer {
    double total_ms = 0.0;
    int64_t launches = 0;
};

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

bool valid_chunk(int c) { return c == 256 || c == 512 || c == 1024; }
bool valid_threads(int t) { return t == 64 || t == 128 || t == 256; }

}  // namespace

namespace nclr_host {
int set_last_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace nclr_host

struct nclr_mpm_sim {
    int dim = 0;
    int material = 0;  // common material of all scenes or nclr_dev::kMixed
    mpm_options opts{};
    std::vector<nclr_dev::SceneParams> scenes;
    nclr_dev::SceneParams* d_scenes = nullptr;
    nclr_dev::StepParams p{};
    hipStream_t stream = nullptr;
    float* state[2] = {nullptr, nullptr};
    uint32_t* pid[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* stamp = nullptr;  // device step stamp
    uint32_t *keys = nullptr, *perm = nullptr, *keys_tmp = nullptr, *perm_tmp = nullptr, *hist = nullptr;
    uint32_t *tile_start = nullptr, *tile_end = nullptr, *item_offset = nullptr;
    float4* slots = nullptr;
    float4* gv = nullptr;
    uint32_t* stamps = nullptr;
    uint32_t* status = nullptr;
    uint32_t* active = nullptr;
    unsigned long long* active_sum = nullptr;
    // incremental sort: valid while the state is stored in the order of the last sort and the
    // tile bounds / work items of that sort are intact (cleared by uploads, loads, re-tuning)
    uint32_t *uidx = nullptr, *ucnt = nullptr, *upref = nullptr, *ttotal = nullptr, *meta = nullptr;
    uint32_t* rtot = nullptr;
    uint32_t *tile_scene = nullptr, *node_tile_scene = nullptr;  // scene maps (batches only)
    // per scene (device): status bits raised since the last read, and the stamp of the step that
    // stopped the scene (StepParams::scene_flags / scene_frozen); host: every bit reported since
    // the state was last set
    uint32_t *scene_flags = nullptr, *scene_frozen = nullptr;
    std::vector<uint32_t> scene_seen;
    int64_t stamp_host = 0;  // mirror of the device step stamp (one increment per step and per adjoint step)
    bool sorted_valid = false;
    // adjoint buffers, allocated by the first backward call
    float4* slots_b = nullptr;
    float4* gbar = nullptr;
    float* epart = nullptr;
    int64_t* erange = nullptr;  // adjoint: per scene, its sorted range
    double* epartial = nullptr; // adjoint: per scene, partial sums of dL/dE
    int echunks = 1;
    float* gtmp[2] = {nullptr, nullptr};
    char* tape = nullptr;  // saved states of a backward segment
    // The sort's outputs (keys, perm, tile bounds, item offsets, the item table, meta) live in one
    // block at fixed offsets, so a backward segment can run each recomputed step's sort in a block
    // of its own (a sort tape entry) and the adjoint kernels read it there.
    char* sort_block = nullptr;
    size_t sort_block_bytes = 0;
    size_t so_keys = 0, so_perm = 0, so_ts = 0, so_te = 0, so_io = 0, so_items = 0, so_meta = 0;
    // the block the next incremental sort reads the old order's work items and meta from, when it
    // sorts into another block (nullptr: the current one)
    char* prev_sort = nullptr;
    char* sort_tape = nullptr;  // per recomputed step of a backward segment, its sort block
    int64_t tape_states = 0;
    // per step of a backward segment, the forward grid (gv) and its pre-update sums (gpre): the
    // adjoint steps then skip P2G and the grid update.  Dense (node tiles x NODES float4 each), so
    // allocated only while device memory allows; without it each adjoint step recomputes them.
    float4 *tape_gv = nullptr, *tape_gpre = nullptr;
    int64_t tape_grid_states = 0;
    float4* gpre = nullptr;  // the grid kernel's pre-update output (nullptr: not kept)
    int g2p_threads = 256;
    // compact step (single scene; mpm_compact.hip): particles per thread, 0 = off; its buffers are
    // allocated by the first compact launch
    int compact_ppt = 0;
    bool compact_overflowed = false;  // a launch's box outgrew the LDS since the state was last set
    int compact_wg_limit = 0;         // co-resident workgroups of the compact step
    float4 *cslots = nullptr, *cgrid = nullptr;
    uint8_t* cmassive = nullptr;
    int *cbox = nullptr, *cinfo = nullptr;
    unsigned long long* csubbox = nullptr;
    uint32_t* cbar = nullptr;
    int64_t steps = 0;
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> pool;
    KernelTimer timers[kNumKernels];

    nclr_host::LaunchArgs args(int parity) const {
        nclr_host::LaunchArgs a{};
        a.dim = dim;
        a.material = material;
        a.g2p_threads = g2p_threads;
        a.svd_mode = opts.fast_svd;
        a.p = p;
        a.stamp = stamp;
        a.state_in = state[parity];
        a.state_out = state[parity ^ 1];
        a.pid_in = pid[parity];
        a.pid_out = pid[parity ^ 1];
        a.keys = keys;
        a.perm = perm;
        a.keys_tmp = keys_tmp;
        a.perm_tmp = perm_tmp;
        a.hist = hist;
        a.tile_start = tile_start;
        a.tile_end = tile_end;
        a.item_offset = item_offset;
        a.slots = slots;
        a.gv = gv;
        a.gpre = gpre;
        a.stamps = stamps;
        a.status = status;
        a.active = active;
        a.uidx = uidx;
        a.ucnt = ucnt;
        a.upref = upref;
        a.ttotal = ttotal;
        a.meta = meta;
        a.rtot = rtot;
        a.prev_item_offset = prev_sort ? reinterpret_cast<const uint32_t*>(prev_sort + so_io) : item_offset;
        a.prev_items = prev_sort ? reinterpret_cast<const uint4*>(prev_sort + so_items) : p.items;
        a.prev_meta = prev_sort ? reinterpret_cast<const uint32_t*>(prev_sort + so_meta) : meta;
        a.stream = stream;
        return a;
    }

    int nf() const { return nclr_host::num_fields(dim); }
    int64_t scene_nodes(int s) const {
        int64_t n = 1;
        for (int a = 0; a < dim; ++a) n *= (scenes[size_t(s)].res + 1);
        return n;
    }

    hipEvent_t take_event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }

    // Resolves recorded event pairs into the per-kernel accumulators.
    int resolve_events() {
        for (auto& pe : pending) {
            HIP_TRY(hipEventSynchronize(pe.stop));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, pe.start, pe.stop));
            timers[pe.kernel].total_ms += ms;
            timers[pe.kernel].launches += 1;
            pool.push_back(pe.start);
            pool.push_back(pe.stop);
        }
        pending.clear();
        return MPM_OK;
    }

    int record_start(PendingEvent& pe, int kernel) {
        if (pending.size() > 8192) RC_TRY(resolve_events());
        pe = PendingEvent{kernel, take_event(), take_event()};
        if (!pe.start || !pe.stop) return fail(MPM_ERR_HIP, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(pe.start, stream));
        return MPM_OK;
    }
    int record_stop(PendingEvent& pe) {
        HIP_TRY(hipEventRecord(pe.stop, stream));
        pending.push_back(pe);
        return MPM_OK;
    }

    template <typename Fn>
    int timed(int kernel, Fn&& launch) {
        if (!opts.timing) {
            HIP_TRY(launch());
            return MPM_OK;
        }
        PendingEvent pe{};
        RC_TRY(record_start(pe, kernel));
        HIP_TRY(launch());
        return record_stop(pe);
    }

    int launch_step_kernels(const nclr_host::LaunchArgs& a) {
        if (sorted_valid) RC_TRY(timed(kSort, [&] { return nclr_host::launch_sort_incremental(a); }));
        else RC_TRY(timed(kSort, [&] { return nclr_host::launch_sort(a); }));
        RC_TRY(timed(kP2G, [&] { return nclr_host::launch_p2g(a); }));
        RC_TRY(timed(kGrid, [&] { return nclr_host::launch_grid(a); }));
        RC_TRY(timed(kG2P, [&] { return nclr_host::launch_g2p(a); }));
        return MPM_OK;
    }

    void drop_graphs() {
        for (auto& g : graph_exec)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    int capture(int parity) {
        const nclr_host::LaunchArgs a = args(parity);
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = nclr_host::launch_sort_incremental(a);
        if (e == hipSuccess) e = nclr_host::launch_p2g(a);
        if (e == hipSuccess) e = nclr_host::launch_grid(a);
        if (e == hipSuccess) e = nclr_host::launch_g2p(a);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream, &graph);
        if (e != hipSuccess || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(MPM_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
        }
        e = hipGraphInstantiate(&graph_exec[parity], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return MPM_OK;
    }

    int step() {
        // graphs hold the incremental sort; a step after an upload or load runs eagerly with the
        // full radix sort, which also rebuilds the tile bounds and work items
        const bool use_graph = opts.graph && !opts.timing && sorted_valid;
        if (use_graph) {
            if (!graph_exec[cur]) RC_TRY(capture(cur));
            HIP_TRY(hipGraphLaunch(graph_exec[cur], stream));
        } else {
            PendingEvent pe{};
            if (opts.timing) RC_TRY(record_start(pe, kStep));
            RC_TRY(launch_step_kernels(args(cur)));
            if (opts.timing) RC_TRY(record_stop(pe));
        }
        if (p.n > 0) cur ^= 1;
        sorted_valid = true;
        steps += 1;
        stamp_host += 1;
        return MPM_OK;
    }

    // ---- the compact step --------------------------------------------------------------------
    int compact_workgroups(int ppt) const { return int((p.n + 256 * int64_t(ppt) - 1) / (256 * int64_t(ppt))); }
    bool compact_whole_fits() const {
        int64_t nodes = 1;
        for (int a = 0; a < dim; ++a) nodes *= scenes[0].res + 1;
        return nodes <= nclr_host::compact_capacity(dim);
    }
    // co-resident workgroups of the compact step: one per CU (its LDS box takes 150 of the CU's
    // 160 KiB), at most kCompactMaxWorkgroups
    int compact_limit_checked(int& limit) {
        if (!compact_wg_limit) {
            int cus = 0;
            HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, opts.device));
            compact_wg_limit = std::max(1, std::min(cus, nclr_dev::kCompactMaxWorkgroups));
        }
        limit = compact_wg_limit;
        return MPM_OK;
    }
    static bool valid_compact(int dim, int ppt) { return ppt == 0 || ppt == 1 || ppt == 2 || (dim == 2 && ppt == 4); }
    int set_compact(int ppt) {
        if (!valid_compact(dim, ppt))
            return fail(MPM_ERR_INVALID, dim == 2 ? "compact must be 0, 1, 2 or 4" : "compact must be 0, 1 or 2 in 3D");
        if (ppt && scenes.size() != 1) return fail(MPM_ERR_UNSUPPORTED, "the compact step runs single-scene engines");
        if (ppt && scenes[0].res > nclr_dev::kCompactMaxRes)
            return fail(MPM_ERR_UNSUPPORTED,
                        "the compact step runs grids up to res " + std::to_string(nclr_dev::kCompactMaxRes));
        if (ppt) {
            int limit = 0;
            RC_TRY(compact_limit_checked(limit));
            if (compact_workgroups(ppt) > limit)
                return fail(MPM_ERR_UNSUPPORTED, "the compact step with " + std::to_string(ppt) +
                                                     " particles per thread holds at most " +
                                                     std::to_string(int64_t(limit) * 256 * ppt) + " particles");
        }
        compact_ppt = ppt;
        return MPM_OK;
    }
    bool compact_usable() const { return compact_ppt > 0 && !compact_overflowed && !opts.timing && p.n > 0; }
    int ensure_compact() {
        if (cslots) return MPM_OK;
        const size_t cap = size_t(nclr_host::compact_capacity(dim));
        HIP_TRY(hipStreamSynchronize(stream));
        int limit = 0;
        RC_TRY(compact_limit_checked(limit));
        if (dmalloc(&cslots, 2 * size_t(limit) * cap) != hipSuccess || dmalloc(&cgrid, cap) != hipSuccess ||
            dmalloc(&cmassive, cap) != hipSuccess || dmalloc(&cbox, size_t(2 * dim)) != hipSuccess ||
            dmalloc(&cinfo, size_t(nclr_dev::kInfoWords)) != hipSuccess ||
            dmalloc(&csubbox, size_t(nclr_dev::kCompactTableWords)) != hipSuccess ||
            dmalloc(&cbar, size_t(2)) != hipSuccess)
            return fail(MPM_ERR_HIP, "compact step: hipMalloc failed");
        return MPM_OK;
    }
    // Up to nsteps steps in one launch; `done` = the steps it took (fewer only when its box outgrew
    // the LDS, which needs a read of the launch's info: when the whole grid fits, none is needed and
    // the call does not wait).
    int compact_steps(int nsteps, int& done) {
        RC_TRY(ensure_compact());
        // slots in Hilbert order of the cells, so each workgroup's particles are neighbours
        HIP_TRY(nclr_host::launch_compact_order(dim, state[cur], pid[cur], state[cur ^ 1], pid[cur ^ 1], p.n, p.stride,
                                                nf(), scenes[0], keys, perm, keys_tmp, perm_tmp, hist, stream));
        cur ^= 1;
        nclr_dev::CompactArgs a{};
        a.st = state[cur];
        a.stride = p.stride;
        a.n = p.n;
        a.nsteps = nsteps;
        RC_TRY(compact_limit_checked(a.slot_sets));
        a.sp = scenes[0];
        a.step = p;
        a.stamp = stamp;
        a.status = status;
        a.slots = cslots;
        a.subbox = csubbox;
        a.bar = cbar;
        a.grid = cgrid;
        a.massive = cmassive;
        a.box = cbox;
        a.info = cinfo;
        HIP_TRY(nclr_host::launch_compact_steps(dim, material, opts.fast_svd, compact_ppt,
                                                compact_workgroups(compact_ppt), a, stream));
        HIP_TRY(nclr_host::launch_compact_to_tiles(dim, a, gv, stamps, active, stream));
        sorted_valid = false;  // particles moved many steps since the last sort
        done = nsteps;
        if (!compact_whole_fits()) {
            int info[nclr_dev::kInfoWords];
            HIP_TRY(hipMemcpyAsync(info, cinfo, sizeof(info), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            if (info[nclr_dev::kInfoReason] == nclr_dev::kCompactOverflow) {
                done = info[nclr_dev::kInfoDone];
                compact_overflowed = true;
            }
        }
        steps += done;
        stamp_host += done;
        return MPM_OK;
    }
    // mpm_advance: the compact step while it is on and its box fits, the tiled step otherwise.
    int advance(int nsteps) {
        int left = nsteps;
        if (left > 0 && compact_usable()) {
            int done = 0;
            RC_TRY(compact_steps(left, done));
            left -= done;
        }
        for (int s = 0; s < left; ++s) RC_TRY(step());
        return MPM_OK;
    }

    // Waits for the stream, checks the engine-wide status word and reads the per-scene words;
    // the scenes' new status bits move into scene_seen (and are cleared on the device).
    int sync_read(std::vector<uint32_t>& flags, std::vector<uint32_t>& frozen) {
        HIP_TRY(hipSetDevice(opts.device));
        HIP_TRY(hipStreamSynchronize(stream));
        const size_t ns = scenes.size();
        uint32_t st = 0;
        flags.assign(ns, 0u);
        frozen.assign(ns, 0u);
        HIP_TRY(hipMemcpy(&st, status, sizeof(st), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(flags.data(), scene_flags, ns * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(frozen.data(), scene_frozen, ns * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (opts.timing) RC_TRY(resolve_events());
        bool any = false;
        for (size_t s = 0; s < ns; ++s) {
            scene_seen[s] |= flags[s];
            any = any || flags[s] != 0u;
        }
        if (any) HIP_TRY(nclr_host::launch_fill_u32(scene_flags, int64_t(ns), 0u, stream));
        if (st != 0) HIP_TRY(nclr_host::launch_fill_u32(status, 1, 0u, stream));
        if (any || st != 0) HIP_TRY(hipStreamSynchronize(stream));
        if (st & nclr_dev::kStatusBarrierTimeout)
            return fail(MPM_ERR_HIP, "the compact step's grid-wide barrier timed out (workgroups not co-resident?); "
                                     "the state is undefined");
        if (st & nclr_dev::kStatusMovedTooFar)
            return fail(MPM_ERR_OUT_OF_RANGE,
                        "the tile sort placed a particle outside the particle range (the state was not produced "
                        "by the step; set it with mpm_set_particles)");
        if (st != 0) return fail(MPM_ERR_HIP, "device status " + std::to_string(st));
        return MPM_OK;
    }

    // Engine step (1-based count of steps taken) that ran with device stamp `stamp`, -1 for 0.
    int64_t step_of_stamp(uint32_t stamp_value) const {
        return stamp_value ? steps - (stamp_host - int64_t(stamp_value)) : -1;
    }

    // mpm_synchronize: MPM_ERR_OUT_OF_RANGE naming every scene the steps since the last check
    // found a particle of off its grid (a stopped scene is found again by every later step).
    int sync() {
        std::vector<uint32_t> flags, frozen;
        RC_TRY(sync_read(flags, frozen));
        std::string list;
        int count = 0;
        uint32_t bits = 0;
        for (size_t s = 0; s < flags.size(); ++s) {
            if (!flags[s]) continue;
            bits |= flags[s];
            if (++count <= 8)
                list += (count > 1 ? ", " : "") + std::to_string(s) + " (step " +
                        std::to_string(step_of_stamp(frozen[s])) + ")";
        }
        if (!count) return MPM_OK;
        if (count > 8) list += ", ... (" + std::to_string(count) + " scenes)";
        const std::string why = (bits & nclr_dev::kStatusMovedTooFar)
                                    ? "a particle moved more than one tile in one step (the state was not produced by "
                                      "the step; set it with mpm_set_particles)"
                                    : "a particle's interpolation stencil left the grid (position not finite or outside "
                                      "[0, 1 - 1.5 dx) on some axis)";
        if (scenes.size() == 1)
            return fail(MPM_ERR_OUT_OF_RANGE, "out_of_range: " + why + " in step " +
                                                  std::to_string(step_of_stamp(frozen[0])) +
                                                  "; the simulation stays at its state before that step");
        return fail(MPM_ERR_OUT_OF_RANGE, "out_of_range in scene " + list + ": " + why +
                                              "; those scenes stay at their state before that step, the others "
                                              "keep stepping");
    }

    // Clears the per-scene words (the state was replaced: every scene runs again).
    int clear_scene_state() {
        HIP_TRY(nclr_host::launch_fill_u32(scene_flags, int64_t(scenes.size()), 0u, stream));
        HIP_TRY(nclr_host::launch_fill_u32(scene_frozen, int64_t(scenes.size()), 0u, stream));
        std::fill(scene_seen.begin(), scene_seen.end(), 0u);
        compact_overflowed = false;
        return MPM_OK;
    }

    // An adjoint needs every scene live (a stopped scene's step is not the reference's step).
    int require_live() {
        std::vector<uint32_t> flags, frozen;
        RC_TRY(sync_read(flags, frozen));
        for (size_t s = 0; s < frozen.size(); ++s)
            if (flags[s] || frozen[s])
                return fail(MPM_ERR_OUT_OF_RANGE, "scene " + std::to_string(s) +
                                                      " is stopped (out_of_range); the adjoint needs every scene live: "
                                                      "set or load a state first");
        return MPM_OK;
    }

    // Device state -> host SoA image (physical slot order) and the slot -> particle map.
    int download(std::vector<float>& host, std::vector<uint32_t>& ids) {
        host.assign(size_t(nf()) * size_t(p.stride), 0.0f);
        ids.assign(size_t(p.stride), 0u);
        HIP_TRY(hipMemcpyAsync(host.data(), state[cur], host.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(ids.data(), pid[cur], ids.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    // Host particle-major arrays (NULL = keep) -> host SoA image in original particle order ->
    // device; resets the slot order to the particle order.
    int upload(const float* x, const float* v, const float* F, const float* C, const float* Jp, const float* mass,
               const float* volume, std::vector<float>& host) {
        const int D = dim;
        const int64_t n = p.n, stride = p.stride;
        const int fX = 0, fV = D, fF = 2 * D, fC = 2 * D + D * D, fJ = 2 * D + 2 * D * D;
        for (int64_t i = 0; i < n; ++i) {
            for (int a = 0; a < D; ++a) {
                if (x) host[(fX + a) * stride + i] = x[i * D + a];
                if (v) host[(fV + a) * stride + i] = v[i * D + a];
            }
            for (int k = 0; k < D * D; ++k) {
                if (F) host[(fF + k) * stride + i] = F[i * D * D + k];
                if (C) host[(fC + k) * stride + i] = C[i * D * D + k];
            }
            if (Jp) host[fJ * stride + i] = Jp[i];
            if (mass) host[(fJ + 1) * stride + i] = mass[i];
            if (volume) host[(fJ + 2) * stride + i] = volume[i];
        }
        cur = 0;
        sorted_valid = false;
        RC_TRY(clear_scene_state());
        HIP_TRY(hipMemcpyAsync(state[0], host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_TRY(nclr_host::launch_iota_u32(pid[0], p.stride, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    // Host SoA image in slot order -> the same image in original particle order.
    void unpermute(std::vector<float>& host, const std::vector<uint32_t>& ids) const {
        std::vector<float> out(host.size(), 0.0f);
        const int64_t n = p.n, stride = p.stride;
        for (int f = 0; f < nf(); ++f)
            for (int64_t s = 0; s < n; ++s) out[f * stride + ids[s]] = host[f * stride + s];
        host.swap(out);
    }

    int set_tuning(int chunk, int threads) {
        chunk = chunk ? chunk : 512;
        threads = threads ? threads : 256;
        if (!valid_chunk(chunk)) return fail(MPM_ERR_INVALID, "p2g_chunk must be 256, 512 or 1024");
        if (!valid_threads(threads)) return fail(MPM_ERR_INVALID, "g2p_threads must be 64, 128 or 256");
        if (chunk != p.chunk || threads != g2p_threads) {
            HIP_TRY(hipStreamSynchronize(stream));
            drop_graphs();
        }
        if (chunk != p.chunk) sorted_valid = false;  // the stored work items use the old chunk
        p.chunk = chunk;
        g2p_threads = threads;
        return MPM_OK;
    }

    int autotune(int trial_steps) {
        if (trial_steps < 1) return fail(MPM_ERR_INVALID, "steps_per_trial must be >= 1");
        RC_TRY(sync());
        const size_t state_bytes = size_t(nf()) * size_t(p.stride) * sizeof(float);
        const size_t pid_bytes = size_t(p.stride) * sizeof(uint32_t);
        float* saved = nullptr;
        uint32_t* saved_pid = nullptr;
        HIP_TRY(dmalloc(&saved, state_bytes / sizeof(float)));
        if (dmalloc(&saved_pid, size_t(p.stride)) != hipSuccess) {
            (void)hipFree(saved);
            return fail(MPM_ERR_HIP, "hipMalloc failed");
        }
        const int cur0 = cur;
        const int64_t steps0 = steps;
        const int timing0 = opts.timing;
        opts.timing = 0;
        hipEvent_t e0 = take_event(), e1 = take_event();
        int rc = (e0 && e1) ? MPM_OK : fail(MPM_ERR_HIP, "hipEventCreate failed");
        auto restore = [&]() -> int {
            cur = cur0;
            sorted_valid = false;
            RC_TRY(clear_scene_state());
            HIP_TRY(hipMemcpyAsync(state[cur], saved, state_bytes, hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipMemcpyAsync(pid[cur], saved_pid, pid_bytes, hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            return MPM_OK;
        };
        if (rc == MPM_OK && hipMemcpyAsync(saved, state[cur], state_bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
            rc = fail(MPM_ERR_HIP, "autotune: state copy failed");
        if (rc == MPM_OK && hipMemcpyAsync(saved_pid, pid[cur], pid_bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
            rc = fail(MPM_ERR_HIP, "autotune: state copy failed");
        int best_chunk = p.chunk, best_threads = g2p_threads, best_compact = 0;
        float best_ms = 3.0e38f;
        const int chunks[3] = {256, 512, 1024};
        const int threads[3] = {64, 128, 256};
        // one trial: a warm-up step (and graph capture), then trial_steps timed steps
        auto trial = [&](float& ms) -> int {
            RC_TRY(advance(1));
            HIP_TRY(hipEventRecord(e0, stream));
            RC_TRY(advance(trial_steps));
            HIP_TRY(hipEventRecord(e1, stream));
            HIP_TRY(hipEventSynchronize(e1));
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            return MPM_OK;
        };
        compact_ppt = 0;
        for (int ci = 0; ci < 3 && rc == MPM_OK; ++ci)
            for (int ti = 0; ti < 3 && rc == MPM_OK; ++ti) {
                rc = restore();
                if (rc == MPM_OK) rc = set_tuning(chunks[ci], threads[ti]);
                float ms = 0.0f;
                if (rc == MPM_OK) rc = trial(ms);
                if (rc == MPM_OK && ms < best_ms) {
                    best_ms = ms;
                    best_chunk = chunks[ci];
                    best_threads = threads[ti];
                }
            }
        // the compact step (single scenes small enough for one co-resident launch)
        if (rc == MPM_OK) rc = set_tuning(best_chunk, best_threads);
        for (int ppt : {1, 2, 4}) {
            if (rc != MPM_OK || scenes.size() != 1 || !valid_compact(dim, ppt) ||
                scenes[0].res > nclr_dev::kCompactMaxRes)
                break;
            int limit = 0;
            rc = compact_limit_checked(limit);
            if (rc != MPM_OK || compact_workgroups(ppt) > limit) continue;
            rc = restore();
            compact_ppt = ppt;
            float ms = 0.0f;
            if (rc == MPM_OK) rc = trial(ms);
            if (rc == MPM_OK && ms < best_ms) {
                best_ms = ms;
                best_compact = ppt;
            }
            compact_ppt = 0;
        }
        if (rc == MPM_OK) rc = restore();
        if (rc == MPM_OK) rc = set_tuning(best_chunk, best_threads);
        if (rc == MPM_OK) compact_ppt = best_compact;
        // trial steps may have stopped scenes whose particles leave the grid; the real steps will
        // find them again
        if (rc == MPM_OK) rc = nclr_host::launch_fill_u32(status, 1, 0u, stream) == hipSuccess ? MPM_OK : fail(MPM_ERR_HIP, "status reset");
        if (rc == MPM_OK) rc = clear_scene_state();
        steps = steps0;
        opts.timing = timing0;
        if (e0) pool.push_back(e0);
        if (e1) pool.push_back(e1);
        (void)hipStreamSynchronize(stream);
        (void)hipFree(saved);
        (void)hipFree(saved_pid);
        return rc;
    }

    // ---- state snapshots and the adjoint -------------------------------------------------------
    int grad_channels() const { return 2 * dim + 2 * dim * dim + 1; }
    size_t state_floats() const { return size_t(nf()) * size_t(p.stride); }
    size_t state_bytes() const { return state_floats() * sizeof(float) + size_t(p.stride) * sizeof(uint32_t); }

    // Snapshot layout: the SoA state (nf fields x stride floats), then the slot -> particle map.
    int save_state(void* dst) {
        char* d = static_cast<char*>(dst);
        HIP_TRY(hipMemcpyAsync(d, state[cur], state_floats() * sizeof(float), hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d + state_floats() * sizeof(float), pid[cur], size_t(p.stride) * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, stream));
        return MPM_OK;
    }
    int load_state(const void* src) {
        const char* s = static_cast<const char*>(src);
        sorted_valid = false;
        RC_TRY(clear_scene_state());
        HIP_TRY(hipMemcpyAsync(state[cur], s, state_floats() * sizeof(float), hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(pid[cur], s + state_floats() * sizeof(float), size_t(p.stride) * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, stream));
        return MPM_OK;
    }

    int echunks_for_scenes() {
        int64_t largest = 0;
        for (const auto& sc : scenes) largest = std::max(largest, sc.n);
        echunks = nclr_host::e_chunks(largest);
        return echunks;
    }

    // keys, perm (n each), tile_start, tile_end, item_offset (ntiles + 1 each), the item table and
    // meta, each 256-byte aligned inside one allocation
    hipError_t alloc_sort_block(size_t nn, size_t nbins, size_t max_items) {
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off += (bytes + 255) / 256 * 256;
            return o;
        };
        const size_t o_keys = take(nn * sizeof(uint32_t)), o_perm = take(nn * sizeof(uint32_t));
        const size_t o_ts = take(nbins * sizeof(uint32_t)), o_te = take(nbins * sizeof(uint32_t));
        const size_t o_io = take(nbins * sizeof(uint32_t)), o_items = take(max_items * sizeof(uint4));
        const size_t o_meta = take(4 * sizeof(uint32_t));
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&sort_block), off);
        if (e != hipSuccess) return e;
        sort_block_bytes = off;
        so_keys = o_keys;
        so_perm = o_perm;
        so_ts = o_ts;
        so_te = o_te;
        so_io = o_io;
        so_items = o_items;
        so_meta = o_meta;
        point_sort_block(sort_block);
        return hipSuccess;
    }

    // The sort's pointers into the block at `base` (the engine's own or a sort tape entry).
    void point_sort_block(char* base) {
        keys = reinterpret_cast<uint32_t*>(base + so_keys);
        perm = reinterpret_cast<uint32_t*>(base + so_perm);
        tile_start = reinterpret_cast<uint32_t*>(base + so_ts);
        tile_end = reinterpret_cast<uint32_t*>(base + so_te);
        item_offset = reinterpret_cast<uint32_t*>(base + so_io);
        p.items = reinterpret_cast<uint4*>(base + so_items);
        meta = reinterpret_cast<uint32_t*>(base + so_meta);
    }

    int ensure_adjoint() {
        if (slots_b) return MPM_OK;
        const size_t slot_nodes = size_t(nclr_host::max_items(p.n, p.ntiles)) * nclr_host::region_nodes(dim);
        const size_t gv_nodes = size_t(p.node_tiles) * nclr_host::tile_nodes(dim);
        const size_t gfl = size_t(grad_channels()) * size_t(std::max<int64_t>(p.n, 1));
        if (dmalloc(&slots_b, slot_nodes) != hipSuccess || dmalloc(&gbar, gv_nodes) != hipSuccess ||
            dmalloc(&epart, size_t(std::max<int64_t>(p.n, 1))) != hipSuccess ||
            dmalloc(&erange, 2 * scenes.size()) != hipSuccess ||
            dmalloc(&epartial, scenes.size() * size_t(echunks_for_scenes())) != hipSuccess ||
            dmalloc(&gtmp[0], gfl) != hipSuccess || dmalloc(&gtmp[1], gfl) != hipSuccess)
            return fail(MPM_ERR_HIP, "hipMalloc of the adjoint buffers failed (out of device memory?)");
        return MPM_OK;
    }


    // One adjoint step of the current state: recompute sort, P2G and the grid update, then run the
    // adjoint kernels.  The state and the step counter are left unchanged.
    // presorted: the sort block already holds this state's sort (a backward segment restored it);
    // only the device step stamp advances.  grid_kept: gv and gpre already hold this state's
    // forward grid (a backward segment's recomputed step wrote them): only the adjoint kernels run.
    int backward_step(const float* gout, float* gin, double* grad_E, bool presorted = false, bool grid_kept = false) {
        RC_TRY(ensure_adjoint());
        if (p.n == 0) return MPM_OK;
        const nclr_host::LaunchArgs a = args(cur);
        const nclr_host::AdjointArgs b{gout, gin, p.n, slots_b, gbar, epart, grad_E, erange, epartial, echunks};
        PendingEvent pe{};
        if (opts.timing) RC_TRY(record_start(pe, kBackward));
        if (presorted && grid_kept) {
            RC_TRY(timed(kG2PAdj, [&] { return nclr_host::launch_g2p_adjoint(a, b); }));
            RC_TRY(timed(kGridAdj, [&] { return nclr_host::launch_grid_adjoint(a, b); }));
            RC_TRY(timed(kP2GAdj, [&] { return nclr_host::launch_p2g_adjoint(a, b); }));
            if (opts.timing) RC_TRY(record_stop(pe));
            return MPM_OK;
        }
        if (presorted) {
            HIP_TRY(nclr_host::launch_bump_u32(stamp, stream));
        } else {
            // the radix sort rewrites the sort block (keys, tile bounds, work items) for the current
            // state, whose slots are still in the previous step's order: the next step must sort
            // from scratch, not incrementally from that block
            RC_TRY(timed(kSort, [&] { return nclr_host::launch_sort(a); }));
            sorted_valid = false;
        }
        RC_TRY(timed(kP2G, [&] { return nclr_host::launch_p2g(a); }));
        RC_TRY(timed(kGrid, [&] { return nclr_host::launch_grid(a); }));
        RC_TRY(timed(kG2PAdj, [&] { return nclr_host::launch_g2p_adjoint(a, b); }));
        RC_TRY(timed(kGridAdj, [&] { return nclr_host::launch_grid_adjoint(a, b); }));
        RC_TRY(timed(kP2GAdj, [&] { return nclr_host::launch_p2g_adjoint(a, b); }));
        if (opts.timing) RC_TRY(record_stop(pe));
        stamp_host += 1;
        return MPM_OK;
    }

    // Where a segment's steps live on a tape: state t (t >= 1) at states + (t - 1) * state_bytes(),
    // laid out as save_state writes it (state 0 stays in the engine's own buffer); step t's sort
    // block (t < nsteps - 1) at sorts + t * sort_block_bytes; with a grid tape (gv non-null), step
    // t's forward grid and pre-update sums at gv + t * gv_nodes and gpre + t * gv_nodes (nsteps
    // entries each: the last step's adjoint recomputes its grid into its entry).
    struct SegTape {
        char* states;
        char* sorts;
        float4* gv;
        float4* gpre;
    };

    size_t gv_nodes() const { return size_t(p.node_tiles) * nclr_host::tile_nodes(dim); }

    // The engine's own buffers and pointers for the length of a segment (graph replay off: graphs
    // hold the engine's own pointers); the destructor puts them back.  Unless `advanced` is set,
    // the step counter and the current state go back to the saved ones (the segment's first state);
    // with it, the segment's last step stands: its state is in the own buffer state[cur ^ 1] saved.
    struct SegGuard {
        nclr_mpm_sim& e;
        float* state[2];
        uint32_t* pid[2];
        int cur;
        int64_t steps;
        int graph;
        float4* gv;
        bool advanced = false;
        explicit SegGuard(nclr_mpm_sim& s)
            : e(s), state{s.state[0], s.state[1]}, pid{s.pid[0], s.pid[1]}, cur(s.cur), steps(s.steps),
              graph(s.opts.graph), gv(s.gv) {
            s.opts.graph = 0;
        }
        ~SegGuard() {
            for (int k = 0; k < 2; ++k) {
                e.state[k] = state[k];
                e.pid[k] = pid[k];
            }
            if (advanced) {
                e.cur = cur ^ 1;
            } else {
                e.cur = cur;
                e.steps = steps;
            }
            e.opts.graph = graph;
            e.sorted_valid = advanced;  // the own sort block holds the last step's sort
            e.gv = gv;
            e.gpre = nullptr;
            e.point_sort_block(e.sort_block);
        }
    };

    // slot <- state t of the segment on tape T
    void point_state(int slot, int t, const SegTape& T, const SegGuard& g) {
        if (t == 0) {
            state[slot] = g.state[g.cur];
            pid[slot] = g.pid[g.cur];
        } else {
            char* entry = T.states + state_bytes() * size_t(t - 1);
            state[slot] = reinterpret_cast<float*>(entry);
            pid[slot] = reinterpret_cast<uint32_t*>(entry + state_floats() * sizeof(float));
        }
    }

    // Steps 0 .. nsteps - 2 of a segment from the current state, each into its tape entries: step t
    // writes state t + 1, sorts state t into its own sort block (the sort pointers are redirected,
    // and the incremental sort reads the old order's work items and meta from the step before's
    // block) and, with a grid tape, writes its grid and pre-update sums there.  The state pointers
    // are left on state nsteps - 1 and the sort pointers on its block.
    int record_steps(int nsteps, const SegTape& T, const SegGuard& g) {
        const size_t gn = gv_nodes();
        struct PrevReset {
            char*& prev;
            ~PrevReset() { prev = nullptr; }
        } reset{prev_sort};
        for (int t = 0; t + 1 < nsteps; ++t) {
            point_state(cur ^ 1, t + 1, T, g);
            gv = T.gv ? T.gv + size_t(t) * gn : g.gv;
            gpre = T.gv ? T.gpre + size_t(t) * gn : nullptr;
            char* entry = T.sorts + sort_block_bytes * size_t(t);
            prev_sort = t == 0 ? sort_block : T.sorts + sort_block_bytes * size_t(t - 1);
            point_sort_block(entry);
            RC_TRY(step());
        }
        return MPM_OK;
    }

    // The adjoint steps of a segment recorded on tape T, from state nsteps - 1 back to state 0, on
    // the tape entries in place.  Steps with a recorded sort skip the sort, and with a recorded grid
    // they run only the adjoint kernels.  The last state was never stepped, so its adjoint step
    // sorts and recomputes its grid; the first one (t = 0) recomputes its grid into the engine's own
    // grid, so the grid export and active-node count afterwards show a step from the segment's start.
    int adjoint_steps(int nsteps, const SegTape& T, const SegGuard& g, const float* gout, float* gin,
                      double* grad_E) {
        const size_t gn = gv_nodes();
        const size_t gbytes = size_t(grad_channels()) * size_t(p.n) * sizeof(float);
        HIP_TRY(hipMemcpyAsync(gtmp[0], gout, gbytes, hipMemcpyDeviceToDevice, stream));
        int c = 0;
        for (int t = nsteps - 1; t >= 0; --t) {
            point_state(cur, t, T, g);
            sorted_valid = false;
            const bool kept = t + 1 < nsteps;
            point_sort_block(kept ? T.sorts + sort_block_bytes * size_t(t) : sort_block);
            gv = T.gv && t != 0 ? T.gv + size_t(t) * gn : g.gv;
            gpre = T.gv ? T.gpre + size_t(t) * gn : nullptr;
            RC_TRY(backward_step(gtmp[c], gtmp[c ^ 1], grad_E, kept, kept && T.gv && t > 0));
            c ^= 1;
        }
        HIP_TRY(hipMemcpyAsync(gin, gtmp[c], gbytes, hipMemcpyDeviceToDevice, stream));
        return MPM_OK;
    }

    // Adjoint of `nsteps` steps from the current state (a checkpoint): the segment's states are
    // recomputed straight into the engine's tape (record_steps) and the adjoint steps run on it
    // (adjoint_steps); no state copies.  The engine's own buffers, pointers and step counter are
    // restored at the end, so the state is the checkpoint again.
    int backward_segment(int nsteps, const float* gout, float* gin, double* grad_E) {
        RC_TRY(ensure_adjoint());
        if (nsteps < 1) return fail(MPM_ERR_INVALID, "nsteps must be >= 1");
        if (p.n == 0) return MPM_OK;
        const size_t sb = state_bytes();
        if (tape_states < nsteps - 1) {
            HIP_TRY(hipStreamSynchronize(stream));
            if (tape) (void)hipFree(tape);
            if (sort_tape) (void)hipFree(sort_tape);
            tape = sort_tape = nullptr;
            tape_states = 0;
            HIP_TRY(hipMalloc(reinterpret_cast<void**>(&tape), sb * size_t(nsteps - 1)));
            HIP_TRY(hipMalloc(reinterpret_cast<void**>(&sort_tape), sort_block_bytes * size_t(nsteps - 1)));
            tape_states = nsteps - 1;
        }
        const size_t gn = gv_nodes();
        if (tape_grid_states < nsteps) {
            // the grid tape: every step of the segment, if it leaves a quarter of the free device memory
            HIP_TRY(hipStreamSynchronize(stream));
            if (tape_gv) (void)hipFree(tape_gv);
            if (tape_gpre) (void)hipFree(tape_gpre);
            tape_gv = tape_gpre = nullptr;
            tape_grid_states = 0;
            size_t free_b = 0, total_b = 0;
            HIP_TRY(hipMemGetInfo(&free_b, &total_b));
            const size_t need = 2 * size_t(nsteps) * gn * sizeof(float4);
            if (need <= free_b / 4 * 3 && dmalloc(&tape_gv, size_t(nsteps) * gn) == hipSuccess) {
                if (dmalloc(&tape_gpre, size_t(nsteps) * gn) == hipSuccess) {
                    tape_grid_states = nsteps;
                } else {
                    (void)hipFree(tape_gv);
                    tape_gv = nullptr;
                    (void)hipGetLastError();
                }
            } else {
                (void)hipGetLastError();
            }
        }
        const bool grid_tape = tape_grid_states >= nsteps;
        const SegTape T{tape, sort_tape, grid_tape ? tape_gv : nullptr, grid_tape ? tape_gpre : nullptr};
        SegGuard g(*this);
        RC_TRY(record_steps(nsteps, T, g));
        return adjoint_steps(nsteps, T, g, gout, gin, grad_E);
    }

    // A recorded segment's tape inside caller memory: states, sort blocks, then (with_grid) the
    // forward grids and the pre-update sums, each part 256-byte aligned.
    size_t record_layout(int nsteps, bool with_grid, char* base, SegTape* T) const {
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off += (bytes + 255) / 256 * 256;
            return o;
        };
        const size_t o_states = take(state_bytes() * size_t(nsteps - 1));
        const size_t o_sorts = take(sort_block_bytes * size_t(nsteps - 1));
        const size_t grid_bytes = with_grid ? size_t(nsteps) * gv_nodes() * sizeof(float4) : 0;
        const size_t o_gv = take(grid_bytes), o_gpre = take(grid_bytes);
        if (T)
            *T = SegTape{base + o_states, base + o_sorts,
                         with_grid ? reinterpret_cast<float4*>(base + o_gv) : nullptr,
                         with_grid ? reinterpret_cast<float4*>(base + o_gpre) : nullptr};
        return std::max<size_t>(off, 256);  // a one-step segment without grids keeps nothing
    }

    struct RecordInfo {
        const void* base;
        int nsteps;
        bool with_grid;
        uint64_t params;  // params_digest() at the recording
    };
    std::vector<RecordInfo> records;  // segments recorded by advance_recorded, most recent last
    static constexpr size_t kMaxRecords = 4096;

    // FNV-1a over the P2G chunk and every scene's parameters (elasticity, grid, time step), the
    // inputs a recorded segment's steps depend on besides the state
    uint64_t params_digest() const {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](const void* data, size_t bytes) {
            const unsigned char* b = static_cast<const unsigned char*>(data);
            for (size_t k = 0; k < bytes; ++k) h = (h ^ b[k]) * 1099511628211ull;
        };
        mix(&p.chunk, sizeof(p.chunk));
        mix(scenes.data(), scenes.size() * sizeof(scenes[0]));
        return h;
    }

    // Advances `nsteps` steps from the current state and records on the caller's tape `rec`
    // (record_layout bytes) what backward_segment would recompute from this state: the steps run
    // as record_steps does, and the last one into the engine's own buffers, so the engine is left
    // at the segment's last state like advance().  The sorts run on the tiled step's path.
    int advance_recorded(int nsteps, char* rec, bool with_grid) {
        if (nsteps < 1) return fail(MPM_ERR_INVALID, "nsteps must be >= 1");
        if (!rec) return fail(MPM_ERR_INVALID, "NULL tape");
        SegTape T{};
        (void)record_layout(nsteps, with_grid, rec, &T);
        if (p.n == 0) {
            RC_TRY(advance(nsteps));
        } else {
            SegGuard g(*this);
            RC_TRY(record_steps(nsteps, T, g));
            const int own = g.cur ^ 1;
            state[cur ^ 1] = g.state[own];
            pid[cur ^ 1] = g.pid[own];
            gv = g.gv;
            gpre = nullptr;
            // the last step sorts into the engine's own block, from the step before's
            prev_sort = nsteps > 1 ? T.sorts + sort_block_bytes * size_t(nsteps - 2) : nullptr;
            point_sort_block(sort_block);
            const int rc = step();
            prev_sort = nullptr;
            RC_TRY(rc);
            g.advanced = true;
        }
        records.erase(std::remove_if(records.begin(), records.end(), [&](const RecordInfo& r) { return r.base == rec; }),
                      records.end());
        if (records.size() >= kMaxRecords) records.erase(records.begin());
        records.push_back(RecordInfo{rec, nsteps, with_grid, params_digest()});
        return MPM_OK;
    }

    // backward_segment over a segment recorded by advance_recorded from the current state (the
    // caller restores the segment's first state): no recomputation of the recorded steps.
    int backward_recorded(int nsteps, char* rec, bool with_grid, const float* gout, float* gin, double* grad_E) {
        RC_TRY(ensure_adjoint());
        if (nsteps < 1) return fail(MPM_ERR_INVALID, "nsteps must be >= 1");
        auto it = std::find_if(records.begin(), records.end(), [&](const RecordInfo& r) { return r.base == rec; });
        if (it == records.end() || it->nsteps != nsteps || it->with_grid != with_grid)
            return fail(MPM_ERR_INVALID, "the tape was not recorded by mpm_advance_recorded with these nsteps and with_grid");
        if (it->params != params_digest())
            return fail(MPM_ERR_INVALID, "the engine's parameters (tuning or elasticity) changed since the tape was recorded");
        if (p.n == 0) return MPM_OK;
        SegTape T{};
        (void)record_layout(nsteps, with_grid, rec, &T);
        SegGuard g(*this);
        return adjoint_steps(nsteps, T, g, gout, gin, grad_E);
    }

    int set_elasticity(int s, float E, float nu) {
        if (!(std::isfinite(E) && std::isfinite(nu)) || nu <= -1.0f || nu >= 0.5f)
            return fail(MPM_ERR_INVALID, "E must be finite and nu in (-1, 0.5)");
        HIP_TRY(hipStreamSynchronize(stream));
        auto& sp = scenes[size_t(s)];
        set_scene_elasticity(sp, E, nu);
        HIP_TRY(hipMemcpyAsync(d_scenes + s, &sp, sizeof(sp), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    static void set_scene_elasticity(nclr_dev::SceneParams& sp, float E, float nu) {
        // src/nclr.h:77-78, float arithmetic
        sp.mu0 = E / (2.0f * (1.0f + nu));
        sp.lambda0 = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
        sp.dmu_dE = float(1.0 / (2.0 * (1.0 + double(nu))));
        sp.dlambda_dE = float(double(nu) / ((1.0 + double(nu)) * (1.0 - 2.0 * double(nu))));
    }

    void release() {
        if (opts.timing) (void)resolve_events();
        drop_graphs();
        for (auto e : pool) (void)hipEventDestroy(e);
        pool.clear();
        void* bufs[] = {state[0], state[1], pid[0],  pid[1],   sort_block, keys_tmp, perm_tmp,
                        hist,     slots, gv, stamps, status,
                        active,   active_sum, stamp,    d_scenes, slots_b, gbar, epart, erange, epartial, gtmp[0],
                        gtmp[1],  tape,       sort_tape, uidx,     ucnt,        upref, ttotal, rtot,
                        tile_scene, node_tile_scene, scene_flags, scene_frozen, cslots, cgrid, cmassive, cbox, cinfo,
                        csubbox, cbar};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        uidx = ucnt = upref = ttotal = meta = rtot = nullptr;
        p.items = nullptr;
        tile_scene = node_tile_scene = nullptr;
        scene_flags = scene_frozen = nullptr;
        cslots = cgrid = nullptr;
        cmassive = nullptr;
        cbox = cinfo = nullptr;
        csubbox = nullptr;
        cbar = nullptr;
        p.tile_scene = p.node_tile_scene = nullptr;
        p.scene_flags = p.scene_frozen = nullptr;
        slots_b = gbar = nullptr;
        epart = gtmp[0] = gtmp[1] = nullptr;
        erange = nullptr;
        epartial = nullptr;
        tape = nullptr;
        sort_block = sort_tape = nullptr;
        tape_states = 0;
        if (tape_gv) (void)hipFree(tape_gv);
        if (tape_gpre) (void)hipFree(tape_gpre);
        tape_gv = tape_gpre = nullptr;
        tape_grid_states = 0;
        state[0] = state[1] = nullptr;
        pid[0] = pid[1] = nullptr;
        keys = perm = keys_tmp = perm_tmp = hist = tile_start = tile_end = item_offset = nullptr;
        slots = gv = nullptr;
        stamps = status = active = stamp = nullptr;
        active_sum = nullptr;
        d_scenes = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

int create_impl(int dim, int nscenes, const mpm_scene* sc, const float* x, const float* v, const float* F,
                const float* C, const float* Jp, const float* mass, const float* volume, const mpm_options* opts,
                mpm_handle* out) {
    if (!out) return fail(MPM_ERR_INVALID, "out handle pointer is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(MPM_ERR_INVALID, "dim must be 2 or 3");
    if (nscenes < 1 || !sc) return fail(MPM_ERR_INVALID, "at least one scene is required");
    int64_t n = 0;
    int common = sc[0].material;
    for (int s = 0; s < nscenes; ++s) {
        const mpm_scene& d = sc[s];
        if (d.material < MPM_SNOW || d.material > MPM_LIQUID) return fail(MPM_ERR_INVALID, "unknown material model");
        if (d.res < 3 || d.res > (dim == 3 ? 512 : 8192)) return fail(MPM_ERR_INVALID, "res out of range");
        if (d.n < 0) return fail(MPM_ERR_INVALID, "negative particle count");
        if (!(d.dt > 0.0f) || !std::isfinite(d.dt)) return fail(MPM_ERR_INVALID, "dt must be positive and finite");
        if (d.material != common) common = nclr_dev::kMixed;
        n += d.n;
    }
    if (n > (int64_t(1) << 31) - 64) return fail(MPM_ERR_INVALID, "particle count out of range");
    if (n > 0 && !x) return fail(MPM_ERR_INVALID, "x must not be NULL");

    if (opts && (opts->fast_svd < 0 || opts->fast_svd > 2))
        return fail(MPM_ERR_INVALID, "fast_svd must be 0 (exact), 1 (fast) or 2 (accurate)");
    auto* h = new nclr_mpm_sim();
    if (opts)
        h->opts = *opts;
    else
        mpm_default_options(&h->opts);
    h->dim = dim;
    h->material = common;
    auto& p = h->p;
    const int lg = nclr_host::tile_log(dim);
    uint64_t tiles = 0, node_tiles = 0;
    int64_t base = 0;
    for (int s = 0; s < nscenes; ++s) {
        const mpm_scene& d = sc[s];
        nclr_dev::SceneParams sp{};
        sp.res = d.res;
        sp.nt = ((d.res - 2) >> lg) + 1;  // stencil bases lie in [0, res-2]
        sp.ntn = (d.res >> lg) + 1;       // nodes lie in [0, res]
        sp.material = d.material;
        sp.identity_mode = h->opts.identity_mode;
        sp.tile_base = uint32_t(tiles);
        sp.node_tile_base = uint32_t(node_tiles);
        sp.dt = d.dt;
        sp.dx = float(1.0 / d.res);
        sp.inv_dx = 1.0f / sp.dx;
        sp.gravity = d.gravity;
        nclr_mpm_sim::set_scene_elasticity(sp, d.E, d.nu);
        sp.allowed = float(double(sp.dx) * 0.9 / double(d.dt));
        sp.particle_base = base;
        sp.n = d.n;
        uint64_t t = 1, nt = 1;
        for (int a = 0; a < dim; ++a) {
            t *= uint64_t(sp.nt);
            nt *= uint64_t(sp.ntn);
        }
        tiles += t;
        node_tiles += nt;
        base += d.n;
        h->scenes.push_back(sp);
    }
    if (tiles >= (uint64_t(1) << 31) || node_tiles >= (uint64_t(1) << 31)) {
        delete h;
        return fail(MPM_ERR_INVALID, "too many tiles");
    }
    p.nscenes = nscenes;
    p.ntiles = uint32_t(tiles);
    p.node_tiles = uint32_t(node_tiles);
    p.chunk = 512;
    p.n = n;
    p.stride = std::max<int64_t>(64, (n + 63) / 64 * 64);

    auto cleanup = [&](int rc) {
        h->release();
        delete h;
        return rc;
    };
    {
        const int rc = h->set_tuning(h->opts.p2g_chunk, h->opts.g2p_threads);
        if (rc != MPM_OK) {
            delete h;
            return rc;
        }
    }
    if (hipSetDevice(h->opts.device) != hipSuccess) return cleanup(fail(MPM_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "hipStreamCreate failed"));
    const size_t state_floats = size_t(h->nf()) * size_t(p.stride);
    const size_t nbins = size_t(tiles) + 1;
    const size_t gv_nodes = size_t(node_tiles) * nclr_host::tile_nodes(dim);
    const size_t slot_nodes = size_t(nclr_host::max_items(n, uint32_t(tiles))) * nclr_host::region_nodes(dim);
    const size_t nn = size_t(std::max<int64_t>(n, 1));
    const size_t bucket_cells =
        size_t(nclr_host::max_items(n, uint32_t(tiles))) * size_t(nclr_host::bucket_offsets(dim));
    if (dmalloc(&h->state[0], state_floats) != hipSuccess || dmalloc(&h->state[1], state_floats) != hipSuccess ||
        dmalloc(&h->pid[0], size_t(p.stride)) != hipSuccess || dmalloc(&h->pid[1], size_t(p.stride)) != hipSuccess ||
        h->alloc_sort_block(nn, nbins, size_t(nclr_host::max_items(n, uint32_t(tiles)))) != hipSuccess ||
        dmalloc(&h->keys_tmp, nn) != hipSuccess || dmalloc(&h->perm_tmp, nn) != hipSuccess ||
        dmalloc(&h->hist, size_t(256) * size_t(nclr_host::sort_tiles(int64_t(nn))) + 256) != hipSuccess ||
        dmalloc(&h->slots, slot_nodes) != hipSuccess ||
        dmalloc(&h->gv, gv_nodes) != hipSuccess || dmalloc(&h->stamps, size_t(node_tiles)) != hipSuccess ||
        dmalloc(&h->status, 1) != hipSuccess || dmalloc(&h->active, size_t(node_tiles)) != hipSuccess ||
        dmalloc(&h->active_sum, 1) != hipSuccess || dmalloc(&h->stamp, 1) != hipSuccess ||
        dmalloc(&h->d_scenes, size_t(nscenes)) != hipSuccess || dmalloc(&h->uidx, nn) != hipSuccess ||
        dmalloc(&h->ucnt, bucket_cells) != hipSuccess || dmalloc(&h->upref, bucket_cells) != hipSuccess ||
        dmalloc(&h->ttotal, nbins) != hipSuccess ||
        dmalloc(&h->rtot, 2 * (size_t(tiles) / nclr_dev::kScanRound + 2)) != hipSuccess ||
        dmalloc(&h->scene_flags, size_t(nscenes)) != hipSuccess || dmalloc(&h->scene_frozen, size_t(nscenes)) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "hipMalloc failed (out of device memory?)"));
    p.scenes = h->d_scenes;
    p.stamp = h->stamp;
    p.scene_flags = h->scene_flags;
    p.scene_frozen = h->scene_frozen;
    h->scene_seen.assign(size_t(nscenes), 0u);
    if (nscenes > 1) {
        std::vector<uint32_t> ts(static_cast<size_t>(tiles)), nts(static_cast<size_t>(node_tiles));
        for (int s = 0; s < nscenes; ++s) {
            const auto& sp = h->scenes[size_t(s)];
            const uint64_t t1 = s + 1 < nscenes ? h->scenes[size_t(s) + 1].tile_base : tiles;
            const uint64_t n1 = s + 1 < nscenes ? h->scenes[size_t(s) + 1].node_tile_base : node_tiles;
            std::fill(ts.begin() + sp.tile_base, ts.begin() + int64_t(t1), uint32_t(s));
            std::fill(nts.begin() + sp.node_tile_base, nts.begin() + int64_t(n1), uint32_t(s));
        }
        if (dmalloc(&h->tile_scene, ts.size()) != hipSuccess || dmalloc(&h->node_tile_scene, nts.size()) != hipSuccess ||
            hipMemcpy(h->tile_scene, ts.data(), ts.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->node_tile_scene, nts.data(), nts.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
                hipSuccess)
            return cleanup(fail(MPM_ERR_HIP, "scene maps: hipMalloc or copy failed"));
        p.tile_scene = h->tile_scene;
        p.node_tile_scene = h->node_tile_scene;
    }
    if (hipMemcpyAsync(h->d_scenes, h->scenes.data(), h->scenes.size() * sizeof(nclr_dev::SceneParams),
                       hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        nclr_host::launch_zero_f4(h->gv, int64_t(gv_nodes), h->stream) != hipSuccess ||
        nclr_host::launch_zero_f4(h->slots, int64_t(slot_nodes), h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->stamps, int64_t(node_tiles), 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->stamp, 1, 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->status, 1, 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->active, int64_t(node_tiles), 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->pid[1], p.stride, 0u, h->stream) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "device initialisation failed"));

    // Defaults of the Particle constructor (src/nclr.h:46-47): v = 0, F = diag<dim>(1), C = 0,
    // Jp = 1, mass = volume = 1.
    std::vector<float> host(state_floats, 0.0f);
    const int D = dim;
    const int fF = 2 * D, fJ = 2 * D + 2 * D * D;
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a)
            if (a < 2 || h->opts.identity_mode == 1) host[size_t(fF + a * D + a) * p.stride + i] = 1.0f;
        host[size_t(fJ) * p.stride + i] = 1.0f;
        host[size_t(fJ + 1) * p.stride + i] = 1.0f;
        host[size_t(fJ + 2) * p.stride + i] = 1.0f;
    }
    int rc = h->upload(x, v, F, C, Jp, mass, volume, host);
    if (rc != MPM_OK) return cleanup(rc);
    if (hipMemcpyAsync(h->state[1], h->state[0], state_floats * sizeof(float), hipMemcpyDeviceToDevice, h->stream) !=
            hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "device initialisation failed"));
    if (h->opts.compact) {
        rc = h->set_compact(h->opts.compact);
        if (rc != MPM_OK) return cleanup(rc);
    }
    if (h->opts.autotune && n > 0) {
        rc = h->autotune(20);
        if (rc != MPM_OK) return cleanup(rc);
    }
    *out = h;
    return MPM_OK;
}

}  // namespace

extern "C" {

void mpm_default_options(mpm_options* opts) {
    if (!opts) return;
    std::memset(opts, 0, sizeof(*opts));
}

int mpm_create(int dim, int material, int res, float dt, float E, float nu, float gravity, int64_t n, const float* x,
               const float* v, const float* F, const float* C, const float* Jp, const float* mass,
               const float* volume, const mpm_options* opts, mpm_handle* out) {
    if (n < 0) return fail(MPM_ERR_INVALID, "particle count out of range");
    const mpm_scene sc{material, res, dt, E, nu, gravity, n};
    return create_impl(dim, 1, &sc, x, v, F, C, Jp, mass, volume, opts, out);
}

int mpm_create_batch(int dim, int nscenes, const mpm_scene* scenes, const float* x, const float* v, const float* F,
                     const float* C, const float* Jp, const float* mass, const float* volume,
                     const mpm_options* opts, mpm_handle* out) {
    return create_impl(dim, nscenes, scenes, x, v, F, C, Jp, mass, volume, opts, out);
}

int mpm_advance(mpm_handle h, int nsteps) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (nsteps < 0) return fail(MPM_ERR_INVALID, "nsteps must be >= 0");
    HIP_TRY(hipSetDevice(h->opts.device));
    return h->advance(nsteps);
}

int mpm_synchronize(mpm_handle h) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    return h->sync();
}

int mpm_scene_status(mpm_handle h, uint32_t* flags, int64_t* failed_step) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    std::vector<uint32_t> f, frozen;
    RC_TRY(h->sync_read(f, frozen));
    for (size_t s = 0; s < frozen.size(); ++s) {
        if (flags) flags[s] = h->scene_seen[s];
        if (failed_step) failed_step[s] = h->step_of_stamp(frozen[s]);
    }
    return MPM_OK;
}

int mpm_get_particles(mpm_handle h, float* x, float* v, float* F, float* C, float* Jp, float* mass,
                      float* volume) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    RC_TRY(h->sync());
    std::vector<float> host;
    std::vector<uint32_t> ids;
    RC_TRY(h->download(host, ids));
    h->unpermute(host, ids);
    const int D = h->dim;
    const int64_t n = h->p.n, stride = h->p.stride;
    const int fX = 0, fV = D, fF = 2 * D, fC = 2 * D + D * D, fJ = 2 * D + 2 * D * D;
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) {
            if (x) x[i * D + a] = host[(fX + a) * stride + i];
            if (v) v[i * D + a] = host[(fV + a) * stride + i];
        }
        for (int k = 0; k < D * D; ++k) {
            if (F) F[i * D * D + k] = host[(fF + k) * stride + i];
            if (C) C[i * D * D + k] = host[(fC + k) * stride + i];
        }
        if (Jp) Jp[i] = host[fJ * stride + i];
        if (mass) mass[i] = host[(fJ + 1) * stride + i];
        if (volume) volume[i] = host[(fJ + 2) * stride + i];
    }
    return MPM_OK;
}

int mpm_set_particles(mpm_handle h, const float* x, const float* v, const float* F, const float* C,
                      const float* Jp, const float* mass, const float* volume) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    RC_TRY(h->sync());
    std::vector<float> host;
    std::vector<uint32_t> ids;
    RC_TRY(h->download(host, ids));
    h->unpermute(host, ids);
    return h->upload(x, v, F, C, Jp, mass, volume, host);
}

int mpm_get_grid_scene(mpm_handle h, int scene, float* velocity, float* mass) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (scene < 0 || scene >= int(h->scenes.size())) return fail(MPM_ERR_INVALID, "scene index out of range");
    RC_TRY(h->sync());
    const auto& sp = h->scenes[size_t(scene)];
    const int64_t total = h->scene_nodes(scene);
    float* dvel = nullptr;
    float* dmass = nullptr;
    HIP_TRY(dmalloc(&dvel, size_t(total) * h->dim));
    if (dmalloc(&dmass, size_t(total)) != hipSuccess) {
        (void)hipFree(dvel);
        return fail(MPM_ERR_HIP, "hipMalloc failed");
    }
    hipError_t e = nclr_host::launch_export_grid(h->args(h->cur), scene, sp.res, sp.node_tile_base, sp.ntn, dvel,
                                                 dmass, total);
    if (e == hipSuccess && velocity)
        e = hipMemcpyAsync(velocity, dvel, size_t(total) * h->dim * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && mass)
        e = hipMemcpyAsync(mass, dmass, size_t(total) * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dvel);
    (void)hipFree(dmass);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("grid export: ") + hipGetErrorString(e));
    return MPM_OK;
}

int mpm_get_grid(mpm_handle h, float* velocity, float* mass) { return mpm_get_grid_scene(h, 0, velocity, mass); }

int mpm_get_scene_params(mpm_handle h, int scene, float* mu_0, float* lambda_0, float* dx, float* inv_dx) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (scene < 0 || scene >= int(h->scenes.size())) return fail(MPM_ERR_INVALID, "scene index out of range");
    const auto& sp = h->scenes[size_t(scene)];
    if (mu_0) *mu_0 = sp.mu0;
    if (lambda_0) *lambda_0 = sp.lambda0;
    if (dx) *dx = sp.dx;
    if (inv_dx) *inv_dx = sp.inv_dx;
    return MPM_OK;
}

int mpm_get_params(mpm_handle h, float* mu_0, float* lambda_0, float* dx, float* inv_dx) {
    return mpm_get_scene_params(h, 0, mu_0, lambda_0, dx, inv_dx);
}

int64_t mpm_num_particles(mpm_handle h) { return h ? h->p.n : -1; }
int64_t mpm_num_nodes(mpm_handle h) {
    if (!h) return -1;
    int64_t t = 0;
    for (size_t s = 0; s < h->scenes.size(); ++s) t += h->scene_nodes(int(s));
    return t;
}
int64_t mpm_steps_taken(mpm_handle h) { return h ? h->steps : -1; }
int mpm_num_scenes(mpm_handle h) { return h ? int(h->scenes.size()) : -1; }

int mpm_kernel_time(mpm_handle h, const char* name, double* avg_ms, int64_t* launches) {
    if (!h || !name) return fail(MPM_ERR_INVALID, "NULL argument");
    int k = -1;
    for (int i = 0; i < kNumKernels; ++i)
        if (std::strcmp(name, kKernelNames[i]) == 0) k = i;
    if (k < 0) return fail(MPM_ERR_INVALID, std::string("unknown kernel name: ") + name);
    HIP_TRY(hipStreamSynchronize(h->stream));
    RC_TRY(h->resolve_events());
    const auto& t = h->timers[k];
    if (avg_ms) *avg_ms = t.launches ? t.total_ms / double(t.launches) : 0.0;
    if (launches) *launches = t.launches;
    return MPM_OK;
}

int mpm_reset_kernel_times(mpm_handle h) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    RC_TRY(h->resolve_events());
    for (auto& t : h->timers) t = KernelTimer{};
    return MPM_OK;
}

int mpm_set_timing(mpm_handle h, int enable) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    RC_TRY(h->resolve_events());
    h->opts.timing = enable ? 1 : 0;
    return MPM_OK;
}

int mpm_set_tuning(mpm_handle h, int p2g_chunk, int g2p_threads) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    return h->set_tuning(p2g_chunk, g2p_threads);
}

int mpm_get_tuning(mpm_handle h, int* p2g_chunk, int* g2p_threads) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (p2g_chunk) *p2g_chunk = h->p.chunk;
    if (g2p_threads) *g2p_threads = h->g2p_threads;
    return MPM_OK;
}

int mpm_set_compact(mpm_handle h, int particles_per_thread) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    HIP_TRY(hipSetDevice(h->opts.device));
    return h->set_compact(particles_per_thread);
}

int mpm_get_compact(mpm_handle h, int* particles_per_thread) {
    if (!h || !particles_per_thread) return fail(MPM_ERR_INVALID, "NULL argument");
    *particles_per_thread = h->compact_ppt;
    return MPM_OK;
}

int mpm_autotune(mpm_handle h, int steps_per_trial) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (h->p.n == 0) return MPM_OK;
    return h->autotune(steps_per_trial);
}

int mpm_active_nodes_scene(mpm_handle h, int scene, int64_t* count) {
    if (!h || !count) return fail(MPM_ERR_INVALID, "NULL argument");
    if (scene < -1 || scene >= int(h->scenes.size())) return fail(MPM_ERR_INVALID, "scene index out of range");
    RC_TRY(h->sync());
    if (h->steps == 0) {
        *count = 0;
        return MPM_OK;
    }
    uint32_t first = 0, tiles = h->p.node_tiles;
    if (scene >= 0) {
        first = h->scenes[size_t(scene)].node_tile_base;
        tiles = (size_t(scene) + 1 < h->scenes.size() ? h->scenes[size_t(scene) + 1].node_tile_base : h->p.node_tiles) - first;
    }
    unsigned long long c = 0;
    HIP_TRY(nclr_host::launch_active_sum(h->args(h->cur), first, tiles, h->active_sum));
    HIP_TRY(hipMemcpyAsync(&c, h->active_sum, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *count = int64_t(c);
    return MPM_OK;
}

int mpm_active_nodes(mpm_handle h, int64_t* count) { return mpm_active_nodes_scene(h, -1, count); }

int mpm_sort_keys(const uint32_t* keys, int64_t n, int bits, uint32_t* perm, int device) {
    if (n < 0 || (n > 0 && (!keys || !perm))) return fail(MPM_ERR_INVALID, "bad arguments");
    if (bits < 0 || bits > 32) return fail(MPM_ERR_INVALID, "bits must be in [0, 32]");
    if (n == 0) return MPM_OK;
    if (n > (int64_t(1) << 31)) return fail(MPM_ERR_INVALID, "too many keys");
    HIP_TRY(hipSetDevice(device));
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    uint32_t *dk = nullptr, *dv = nullptr, *tk = nullptr, *tv = nullptr, *hist = nullptr;
    hipError_t e = dmalloc(&dk, size_t(n));
    if (e == hipSuccess) e = dmalloc(&dv, size_t(n));
    if (e == hipSuccess) e = dmalloc(&tk, size_t(n));
    if (e == hipSuccess) e = dmalloc(&tv, size_t(n));
    if (e == hipSuccess) e = dmalloc(&hist, size_t(256) * size_t(nclr_host::sort_tiles(n)) + 256);
    if (e == hipSuccess) e = hipMemcpyAsync(dk, keys, size_t(n) * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = nclr_host::radix_sort_pairs(dk, dv, tk, tv, hist, n, bits, true, s);
    if (e == hipSuccess) e = hipMemcpyAsync(perm, dv, size_t(n) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    void* bufs[] = {dk, dv, tk, tv, hist};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("mpm_sort_keys: ") + hipGetErrorString(e));
    return MPM_OK;
}

int mpm_svd(int dim, int64_t n, const float* A, float* U, float* sigma, float* V, int exact, int device) {
    if (dim != 2 && dim != 3) return fail(MPM_ERR_INVALID, "dim must be 2 or 3");
    if (n < 0 || (n > 0 && (!A || !U || !sigma || !V))) return fail(MPM_ERR_INVALID, "bad arguments");
    if (exact < 0 || exact > 2) return fail(MPM_ERR_INVALID, "exact must be 1 (exact), 0 (fast) or 2 (accurate)");
    if (n == 0) return MPM_OK;
    HIP_TRY(hipSetDevice(device));
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t mats = size_t(n) * dim * dim, vecs = size_t(n) * dim;
    float *dA = nullptr, *dU = nullptr, *dS = nullptr, *dV = nullptr;
    hipError_t e = dmalloc(&dA, mats);
    if (e == hipSuccess) e = dmalloc(&dU, mats);
    if (e == hipSuccess) e = dmalloc(&dS, vecs);
    if (e == hipSuccess) e = dmalloc(&dV, mats);
    if (e == hipSuccess) e = hipMemcpyAsync(dA, A, mats * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = nclr_host::launch_svd(dim, exact == 1 ? 0 : exact == 0 ? 1 : 2, dA, dU, dS, dV, n, s);
    if (e == hipSuccess) e = hipMemcpyAsync(U, dU, mats * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(sigma, dS, vecs * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(V, dV, mats * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    void* bufs[] = {dA, dU, dS, dV};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string("mpm_svd: ") + hipGetErrorString(e));
    return MPM_OK;
}

int mpm_grad_channels(int dim) { return (dim == 2 || dim == 3) ? 2 * dim + 2 * dim * dim + 1 : -1; }

int64_t mpm_state_bytes(mpm_handle h) { return h ? int64_t(h->state_bytes()) : -1; }

int mpm_save_state(mpm_handle h, void* dst) {
    if (!h || !dst) return fail(MPM_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->opts.device));
    return h->save_state(dst);
}

int mpm_load_state(mpm_handle h, const void* src) {
    if (!h || !src) return fail(MPM_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->opts.device));
    return h->load_state(src);
}

int mpm_get_state_device(mpm_handle h, float* dst) {
    if (!h || (!dst && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->opts.device));
    HIP_TRY(nclr_host::launch_permute_channels(h->pid[h->cur], h->state[h->cur], h->p.stride, dst, h->p.n,
                                               h->grad_channels(), h->p.n, true, h->stream));
    return MPM_OK;
}

int mpm_set_state_device(mpm_handle h, const float* src) {
    if (!h || (!src && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->opts.device));
    const int64_t n = h->p.n, stride = h->p.stride;
    if (n == 0) return MPM_OK;
    const int o = h->cur ^ 1, ng = h->grad_channels();
    // Jp, mass and volume into particle order, then x, v, F, C from the caller; slots = particles
    HIP_TRY(nclr_host::launch_permute_channels(h->pid[h->cur], h->state[h->cur] + int64_t(ng) * stride, stride,
                                               h->state[o] + int64_t(ng) * stride, stride, h->nf() - ng, n, true,
                                               h->stream));
    HIP_TRY(hipMemcpy2DAsync(h->state[o], size_t(stride) * sizeof(float), src, size_t(n) * sizeof(float),
                             size_t(n) * sizeof(float), size_t(ng), hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(nclr_host::launch_iota_u32(h->pid[o], stride, h->stream));
    h->cur = o;
    h->sorted_valid = false;
    return h->clear_scene_state();
}

int mpm_grad_from_particle_order(mpm_handle h, const float* src, float* dst) {
    if (!h || ((!src || !dst) && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    if (src == dst && h->p.n > 0) return fail(MPM_ERR_INVALID, "src and dst must not alias");
    HIP_TRY(hipSetDevice(h->opts.device));
    HIP_TRY(nclr_host::launch_permute_channels(h->pid[h->cur], src, h->p.n, dst, h->p.n, h->grad_channels(), h->p.n,
                                               false, h->stream));
    return MPM_OK;
}

int mpm_grad_to_particle_order(mpm_handle h, const float* src, float* dst) {
    if (!h || ((!src || !dst) && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    if (src == dst && h->p.n > 0) return fail(MPM_ERR_INVALID, "src and dst must not alias");
    HIP_TRY(hipSetDevice(h->opts.device));
    HIP_TRY(nclr_host::launch_permute_channels(h->pid[h->cur], src, h->p.n, dst, h->p.n, h->grad_channels(), h->p.n,
                                               true, h->stream));
    return MPM_OK;
}

int mpm_backward_step(mpm_handle h, const float* grad_out, float* grad_in, double* grad_E) {
    if (!h || ((!grad_out || !grad_in) && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    if (grad_out == grad_in && h->p.n > 0) return fail(MPM_ERR_INVALID, "grad_out and grad_in must not alias");
    HIP_TRY(hipSetDevice(h->opts.device));
    RC_TRY(h->require_live());
    return h->backward_step(grad_out, grad_in, grad_E);
}

int mpm_backward_segment(mpm_handle h, int nsteps, const float* grad_out, float* grad_in, double* grad_E) {
    if (!h || ((!grad_out || !grad_in) && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->opts.device));
    RC_TRY(h->require_live());
    return h->backward_segment(nsteps, grad_out, grad_in, grad_E);
}

int64_t mpm_tape_bytes(mpm_handle h, int nsteps) {
    if (!h || nsteps < 1) return -1;
    const int64_t grid = 2 * int64_t(h->p.node_tiles) * int64_t(nclr_host::tile_nodes(h->dim)) * int64_t(sizeof(float4));
    return int64_t(nsteps - 1) * int64_t(h->state_bytes() + h->sort_block_bytes) + int64_t(nsteps) * grid;
}

int64_t mpm_record_bytes(mpm_handle h, int nsteps, int with_grid) {
    if (!h || nsteps < 1) return -1;
    return int64_t(h->record_layout(nsteps, with_grid != 0, nullptr, nullptr));
}

int mpm_advance_recorded(mpm_handle h, int nsteps, void* tape, int with_grid) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    HIP_TRY(hipSetDevice(h->opts.device));
    return h->advance_recorded(nsteps, static_cast<char*>(tape), with_grid != 0);
}

int mpm_backward_recorded(mpm_handle h, int nsteps, void* tape, int with_grid, const float* grad_out,
                          float* grad_in, double* grad_E) {
    if (!h || !tape || ((!grad_out || !grad_in) && h->p.n > 0)) return fail(MPM_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->opts.device));
    RC_TRY(h->require_live());
    return h->backward_recorded(nsteps, static_cast<char*>(tape), with_grid != 0, grad_out, grad_in, grad_E);
}

int mpm_set_scene_elasticity(mpm_handle h, int scene, float E, float nu) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (scene < 0 || scene >= int(h->scenes.size())) return fail(MPM_ERR_INVALID, "scene index out of range");
    HIP_TRY(hipSetDevice(h->opts.device));
    return h->set_elasticity(scene, E, nu);
}

}  // extern "C"

// ---- scene generators and viewer hooks: null stream, synchronous --------------------------------

namespace {

// Uploads `data` (count elements) to a temporary device buffer.
template <typename T>
hipError_t upload_tmp(const T* data, size_t count, T** dev) {
    hipError_t e = dmalloc(dev, count);
    if (e == hipSuccess && count) e = hipMemcpy(*dev, data, count * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

int finish(hipError_t e, std::initializer_list<void*> bufs, const char* what) {
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(MPM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return MPM_OK;
}

bool valid_cube(const mpm_cube& c, int dim) {
    for (int a = 0; a < dim; ++a)
        if (c.res[a] < 1 || !std::isfinite(c.lo[a]) || !std::isfinite(c.hi[a]) || !std::isfinite(c.v[a])) return false;
    return true;
}

}  // namespace

extern "C" {

int mpm_gen_cubes(int dim, int ncubes, const mpm_cube* cubes, int identity_mode, int64_t n, float* image,
                  int device) {
    if (dim != 2 && dim != 3) return fail(MPM_ERR_INVALID, "dim must be 2 or 3");
    if (ncubes < 0 || (ncubes > 0 && !cubes)) return fail(MPM_ERR_INVALID, "bad cube list");
    if (identity_mode != 0 && identity_mode != 1) return fail(MPM_ERR_INVALID, "identity_mode must be 0 or 1");
    std::vector<int64_t> offsets(size_t(ncubes) + 1, 0);
    for (int c = 0; c < ncubes; ++c) {
        if (!valid_cube(cubes[c], dim)) return fail(MPM_ERR_INVALID, "cube " + std::to_string(c) + " is invalid");
        int64_t pts = 1;
        for (int a = 0; a < dim; ++a) pts *= cubes[c].res[a];
        offsets[size_t(c) + 1] = offsets[size_t(c)] + pts;
    }
    if (n != offsets.back()) return fail(MPM_ERR_INVALID, "n must equal the total lattice point count");
    if (n == 0) return MPM_OK;
    if (!image) return fail(MPM_ERR_INVALID, "NULL image");
    HIP_TRY(hipSetDevice(device));
    mpm_cube* dc = nullptr;
    int64_t* doff = nullptr;
    hipError_t e = upload_tmp(cubes, size_t(ncubes), &dc);
    if (e == hipSuccess) e = upload_tmp(offsets.data(), offsets.size(), &doff);
    if (e == hipSuccess) e = nclr_host::launch_gen_cubes(dim, dc, doff, ncubes, identity_mode, n, image, nullptr);
    return finish(e, {dc, doff}, "mpm_gen_cubes");
}

int mpm_rand_advance(uint32_t state[4], uint64_t k) {
    if (!state) return fail(MPM_ERR_INVALID, "NULL state");
    nclr_host::rand_jump(state, k);
    return MPM_OK;
}

int mpm_gen_rand(int64_t count, uint32_t state[4], float* out, int device) {
    if (!state || count < 0 || (count > 0 && !out)) return fail(MPM_ERR_INVALID, "bad arguments");
    if (count == 0) return MPM_OK;
    HIP_TRY(hipSetDevice(device));
    const std::vector<uint4> starts = nclr_host::rand_starts(state, count);
    uint4* ds = nullptr;
    hipError_t e = upload_tmp(starts.data(), starts.size(), &ds);
    if (e == hipSuccess) e = nclr_host::launch_gen_rand(ds, count, out, nullptr);
    const int rc = finish(e, {ds}, "mpm_gen_rand");
    if (rc == MPM_OK) nclr_host::rand_jump(state, uint64_t(count));
    return rc;
}

int mpm_gen_sample(int dim, int ncenters, const float* centers, int per_center, float half_width, uint32_t state[4],
                   int identity_mode, float* image, int device) {
    if (dim != 2 && dim != 3) return fail(MPM_ERR_INVALID, "dim must be 2 or 3");
    if (ncenters < 0 || per_center < 0 || !state || (ncenters > 0 && !centers))
        return fail(MPM_ERR_INVALID, "bad arguments");
    if (identity_mode != 0 && identity_mode != 1) return fail(MPM_ERR_INVALID, "identity_mode must be 0 or 1");
    const int64_t n = int64_t(ncenters) * per_center;
    if (n == 0) return MPM_OK;
    if (!image) return fail(MPM_ERR_INVALID, "NULL image");
    HIP_TRY(hipSetDevice(device));
    const std::vector<uint4> starts = nclr_host::rand_starts(state, n * dim);
    uint4* ds = nullptr;
    float* dc = nullptr;
    hipError_t e = upload_tmp(starts.data(), starts.size(), &ds);
    if (e == hipSuccess) e = upload_tmp(centers, size_t(ncenters) * dim, &dc);
    if (e == hipSuccess)
        e = nclr_host::launch_gen_sample(dim, ds, dc, per_center, half_width, identity_mode, n, image, nullptr);
    const int rc = finish(e, {ds, dc}, "mpm_gen_sample");
    if (rc == MPM_OK) nclr_host::rand_jump(state, uint64_t(n * dim));
    return rc;
}

int mpm_gyroid(int64_t n, const float* pos, float k, float t, float* out, int device) {
    if (n < 0 || (n > 0 && (!pos || !out))) return fail(MPM_ERR_INVALID, "bad arguments");
    if (n == 0) return MPM_OK;
    HIP_TRY(hipSetDevice(device));
    return finish(nclr_host::launch_eval_gyroid(n, pos, k, t, out, nullptr), {}, "mpm_gyroid");
}

int mpm_gen_gyroid(const mpm_cube* box, float k, float t, float thickness, int identity_mode, float* image,
                   int64_t capacity, int64_t* n_out, int device) {
    if (!box || !n_out || !valid_cube(*box, 3)) return fail(MPM_ERR_INVALID, "bad box");
    if (identity_mode != 0 && identity_mode != 1) return fail(MPM_ERR_INVALID, "identity_mode must be 0 or 1");
    if (!(thickness >= 0.0f) || !std::isfinite(k) || !std::isfinite(t)) return fail(MPM_ERR_INVALID, "bad k, t or thickness");
    if (int64_t(box->res[0]) * box->res[1] * box->res[2] > (int64_t(1) << 40))
        return fail(MPM_ERR_INVALID, "box has too many lattice points");
    HIP_TRY(hipSetDevice(device));
    const int64_t nb = nclr_host::gyroid_blocks(*box);
    uint32_t* counts = nullptr;
    int64_t* offsets = nullptr;
    hipError_t e = dmalloc(&counts, size_t(nb));
    if (e == hipSuccess) e = dmalloc(&offsets, size_t(nb) + 1);
    if (e == hipSuccess) e = nclr_host::launch_gyroid_count(*box, k, t, thickness, counts, offsets, nullptr);
    int64_t total = 0;
    if (e == hipSuccess) e = hipMemcpy(&total, offsets + nb, sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        *n_out = total;
        if (image && capacity != total) {
            (void)finish(e, {counts, offsets}, "mpm_gen_gyroid");
            return fail(MPM_ERR_INVALID, "capacity must equal the point count (" + std::to_string(total) + ")");
        }
        if (image)
            e = nclr_host::launch_gyroid_scatter(*box, k, t, thickness, offsets, identity_mode, total, image, nullptr);
    }
    return finish(e, {counts, offsets}, "mpm_gen_gyroid");
}

int mpm_project_points(int64_t n, const float* pos, int centered, float* out, int device) {
    if (n < 0 || (n > 0 && (!pos || !out))) return fail(MPM_ERR_INVALID, "bad arguments");
    if (n == 0) return MPM_OK;
    HIP_TRY(hipSetDevice(device));
    return finish(nclr_host::launch_project(n, pos, centered ? 1 : 0, out, nullptr), {}, "mpm_project_points");
}

int mpm_destroy(mpm_handle h) {
    if (!h) return MPM_OK;
    (void)hipSetDevice(h->opts.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->release();
    delete h;
    return MPM_OK;
}

const char* mpm_last_error(void) { return g_err.c_str(); }

const char* mpm_version(void) {
    return "nclr_mpm 0.3 gfx950 (fp32, no fp contraction, IEEE div/sqrt, tile-sorted SoA particles, "
           "fixed-point LDS tile P2G, batched scenes, hipGraph steps)";
}

}  // extern "C"
