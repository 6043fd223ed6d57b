// C ABI of the MLS-MPM engine (declarations and contracts: include/nclr_mpm.h).
//
// One handle = one simulation = one HIP stream.  The handle owns two ping-pong buffers of
// structure-of-arrays particle state (G2P writes the state in tile order into the other buffer),
// the original particle index of every slot, the sort buffers, the per-tile P2G slots, the
// tile-blocked updated grid and a device status word that kernels set when a particle leaves
// the grid.  The status is read back (and the error reported) only by synchronising calls, so
// mpm_advance() never waits on the device.
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

enum KernelId { kSort = 0, kP2G = 1, kGrid = 2, kG2P = 3, kNumKernels = 4 };
const char* const kKernelNames[kNumKernels] = {"sort", "p2g", "grid", "g2p"};

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

}  // namespace

struct nclr_mpm_sim {
    int dim = 0, material = 0, res = 0;
    mpm_options opts{};
    nclr_dev::StepParams p{};
    hipStream_t stream = nullptr;
    float* state[2] = {nullptr, nullptr};
    uint32_t* pid[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t *keys = nullptr, *perm = nullptr, *keys_tmp = nullptr, *perm_tmp = nullptr, *hist = nullptr;
    uint32_t *tile_start = nullptr, *tile_end = nullptr, *item_offset = nullptr;
    float4* slots = nullptr;
    float4* gv = nullptr;
    uint32_t* stamps = nullptr;
    uint32_t* status = nullptr;
    uint32_t* active = nullptr;
    unsigned long long* active_sum = nullptr;
    int64_t node_tiles = 0;
    int64_t steps = 0;
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> pool;
    KernelTimer timers[kNumKernels];

    nclr_host::LaunchArgs args() const {
        nclr_host::LaunchArgs a{};
        a.dim = dim;
        a.material = material;
        a.p = p;
        a.state_in = state[cur];
        a.state_out = state[cur ^ 1];
        a.pid_in = pid[cur];
        a.pid_out = pid[cur ^ 1];
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
        a.stamps = stamps;
        a.status = status;
        a.active = active;
        a.stream = stream;
        return a;
    }

    int nf() const { return nclr_host::num_fields(dim); }
    int64_t num_nodes() const {
        int64_t n = 1;
        for (int a = 0; a < dim; ++a) n *= (res + 1);
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

    template <typename Fn>
    int timed(int kernel, Fn&& launch) {
        if (!opts.timing) {
            HIP_TRY(launch());
            return MPM_OK;
        }
        if (pending.size() > 8192) {
            const int rc = resolve_events();
            if (rc != MPM_OK) return rc;
        }
        PendingEvent pe{kernel, take_event(), take_event()};
        if (!pe.start || !pe.stop) return fail(MPM_ERR_HIP, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(pe.start, stream));
        HIP_TRY(launch());
        HIP_TRY(hipEventRecord(pe.stop, stream));
        pending.push_back(pe);
        return MPM_OK;
    }

    int step() {
        p.stamp = p.stamp + 1;
        if (p.stamp == 0) p.stamp = 1;
        const nclr_host::LaunchArgs a = args();
        int rc = timed(kSort, [&] { return nclr_host::launch_sort(a); });
        if (rc != MPM_OK) return rc;
        rc = timed(kP2G, [&] { return nclr_host::launch_p2g(a); });
        if (rc != MPM_OK) return rc;
        rc = timed(kGrid, [&] { return nclr_host::launch_grid(a); });
        if (rc != MPM_OK) return rc;
        rc = timed(kG2P, [&] { return nclr_host::launch_g2p(a); });
        if (rc != MPM_OK) return rc;
        if (p.n > 0) cur ^= 1;
        steps += 1;
        return MPM_OK;
    }

    int sync() {
        HIP_TRY(hipSetDevice(opts.device));
        HIP_TRY(hipStreamSynchronize(stream));
        uint32_t st = 0;
        HIP_TRY(hipMemcpy(&st, status, sizeof(st), hipMemcpyDeviceToHost));
        if (opts.timing) {
            const int rc = resolve_events();
            if (rc != MPM_OK) return rc;
        }
        if (st != 0) {
            HIP_TRY(nclr_host::launch_fill_u32(status, 1, 0u, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            if (st & nclr_dev::kStatusOutOfRange)
                return fail(MPM_ERR_OUT_OF_RANGE,
                            "out_of_range: a particle's interpolation stencil left the grid "
                            "(position not finite or outside [0, 1 - 1.5 dx) on some axis)");
            return fail(MPM_ERR_HIP, "device status " + std::to_string(st));
        }
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

    void release() {
        if (opts.timing) (void)resolve_events();
        for (auto e : pool) (void)hipEventDestroy(e);
        pool.clear();
        void* bufs[] = {state[0], state[1], pid[0], pid[1], keys, perm, keys_tmp, perm_tmp, hist, tile_start,
                        tile_end, item_offset, slots, gv, stamps, status, active, active_sum};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        state[0] = state[1] = nullptr;
        pid[0] = pid[1] = nullptr;
        keys = perm = keys_tmp = perm_tmp = hist = tile_start = tile_end = item_offset = nullptr;
        slots = gv = nullptr;
        stamps = status = active = nullptr;
        active_sum = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

extern "C" {

void mpm_default_options(mpm_options* opts) {
    if (!opts) return;
    std::memset(opts, 0, sizeof(*opts));
}

int mpm_create(int dim, int material, int res, float dt, float E, float nu, float gravity, int64_t n, const float* x,
               const float* v, const float* F, const float* C, const float* Jp, const float* mass,
               const float* volume, const mpm_options* opts, mpm_handle* out) {
    if (!out) return fail(MPM_ERR_INVALID, "out handle pointer is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(MPM_ERR_INVALID, "dim must be 2 or 3");
    if (material < MPM_SNOW || material > MPM_LIQUID) return fail(MPM_ERR_INVALID, "unknown material model");
    if (res < 3 || res > (dim == 3 ? 512 : 8192)) return fail(MPM_ERR_INVALID, "res out of range");
    if (n < 0 || n > (int64_t(1) << 31) - 64) return fail(MPM_ERR_INVALID, "particle count out of range");
    if (n > 0 && !x) return fail(MPM_ERR_INVALID, "x must not be NULL");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return fail(MPM_ERR_INVALID, "dt must be positive and finite");

    auto* h = new nclr_mpm_sim();
    if (opts)
        h->opts = *opts;
    else
        mpm_default_options(&h->opts);
    h->dim = dim;
    h->material = material;
    h->res = res;
    auto& p = h->p;
    p.res = res;
    const int lg = nclr_host::tile_log(dim);
    p.nt = ((res - 2) >> lg) + 1;  // stencil bases lie in [0, res-2]
    p.ntn = (res >> lg) + 1;       // nodes lie in [0, res]
    uint64_t ntiles = 1;
    h->node_tiles = 1;
    for (int a = 0; a < dim; ++a) {
        ntiles *= uint64_t(p.nt);
        h->node_tiles *= p.ntn;
    }
    p.ntiles = uint32_t(ntiles);
    p.identity_mode = h->opts.identity_mode;
    p.stamp = 0;
    p.dt = dt;
    p.dx = float(1.0 / res);
    p.inv_dx = 1.0f / p.dx;
    p.gravity = gravity;
    // src/nclr.h:77-78, float arithmetic
    p.mu0 = E / (2.0f * (1.0f + nu));
    p.lambda0 = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    p.allowed = float(double(p.dx) * 0.9 / double(dt));
    p.n = n;
    p.stride = std::max<int64_t>(64, (n + 63) / 64 * 64);

    auto cleanup = [&](int rc) {
        h->release();
        delete h;
        return rc;
    };
    if (hipSetDevice(h->opts.device) != hipSuccess) return cleanup(fail(MPM_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "hipStreamCreate failed"));
    const size_t state_floats = size_t(h->nf()) * size_t(p.stride);
    const size_t nbins = size_t(ntiles) + 1;
    const size_t gv_nodes = size_t(h->node_tiles) * nclr_host::tile_nodes(dim);
    const size_t slot_nodes = size_t(nclr_host::max_items(dim, n, uint32_t(ntiles))) * nclr_host::region_nodes(dim);
    const size_t nn = size_t(std::max<int64_t>(n, 1));
    if (dmalloc(&h->state[0], state_floats) != hipSuccess || dmalloc(&h->state[1], state_floats) != hipSuccess ||
        dmalloc(&h->pid[0], size_t(p.stride)) != hipSuccess || dmalloc(&h->pid[1], size_t(p.stride)) != hipSuccess ||
        dmalloc(&h->keys, nn) != hipSuccess || dmalloc(&h->perm, nn) != hipSuccess ||
        dmalloc(&h->keys_tmp, nn) != hipSuccess || dmalloc(&h->perm_tmp, nn) != hipSuccess ||
        dmalloc(&h->hist, size_t(256) * size_t(nclr_host::sort_tiles(int64_t(nn))) + 256) != hipSuccess ||
        dmalloc(&h->tile_start, nbins) != hipSuccess || dmalloc(&h->tile_end, nbins) != hipSuccess ||
        dmalloc(&h->item_offset, nbins) != hipSuccess || dmalloc(&h->slots, slot_nodes) != hipSuccess ||
        dmalloc(&h->gv, gv_nodes) != hipSuccess || dmalloc(&h->stamps, size_t(h->node_tiles)) != hipSuccess ||
        dmalloc(&h->status, 1) != hipSuccess || dmalloc(&h->active, size_t(h->node_tiles)) != hipSuccess ||
        dmalloc(&h->active_sum, 1) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "hipMalloc failed (out of device memory?)"));
    if (nclr_host::launch_zero_f4(h->gv, int64_t(gv_nodes), h->stream) != hipSuccess ||
        nclr_host::launch_zero_f4(h->slots, int64_t(slot_nodes), h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->stamps, h->node_tiles, 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->status, 1, 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->active, h->node_tiles, 0u, h->stream) != hipSuccess ||
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
    const int rc = h->upload(x, v, F, C, Jp, mass, volume, host);
    if (rc != MPM_OK) return cleanup(rc);
    if (hipMemcpyAsync(h->state[1], h->state[0], state_floats * sizeof(float), hipMemcpyDeviceToDevice, h->stream) !=
            hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "device initialisation failed"));
    *out = h;
    return MPM_OK;
}

int mpm_advance(mpm_handle h, int nsteps) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (nsteps < 0) return fail(MPM_ERR_INVALID, "nsteps must be >= 0");
    HIP_TRY(hipSetDevice(h->opts.device));
    for (int s = 0; s < nsteps; ++s) {
        const int rc = h->step();
        if (rc != MPM_OK) return rc;
    }
    return MPM_OK;
}

int mpm_synchronize(mpm_handle h) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    return h->sync();
}

int mpm_get_particles(mpm_handle h, float* x, float* v, float* F, float* C, float* Jp, float* mass,
                      float* volume) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    int rc = h->sync();
    if (rc != MPM_OK) return rc;
    std::vector<float> host;
    std::vector<uint32_t> ids;
    rc = h->download(host, ids);
    if (rc != MPM_OK) return rc;
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
    int rc = h->sync();
    if (rc != MPM_OK) return rc;
    std::vector<float> host;
    std::vector<uint32_t> ids;
    rc = h->download(host, ids);
    if (rc != MPM_OK) return rc;
    h->unpermute(host, ids);
    return h->upload(x, v, F, C, Jp, mass, volume, host);
}

int mpm_get_grid(mpm_handle h, float* velocity, float* mass) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    int rc = h->sync();
    if (rc != MPM_OK) return rc;
    const int64_t total = h->num_nodes();
    float* dvel = nullptr;
    float* dmass = nullptr;
    HIP_TRY(dmalloc(&dvel, size_t(total) * h->dim));
    if (dmalloc(&dmass, size_t(total)) != hipSuccess) {
        (void)hipFree(dvel);
        return fail(MPM_ERR_HIP, "hipMalloc failed");
    }
    hipError_t e = nclr_host::launch_export_grid(h->args(), dvel, dmass, total);
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

int mpm_get_params(mpm_handle h, float* mu_0, float* lambda_0, float* dx, float* inv_dx) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    if (mu_0) *mu_0 = h->p.mu0;
    if (lambda_0) *lambda_0 = h->p.lambda0;
    if (dx) *dx = h->p.dx;
    if (inv_dx) *inv_dx = h->p.inv_dx;
    return MPM_OK;
}

int64_t mpm_num_particles(mpm_handle h) { return h ? h->p.n : -1; }
int64_t mpm_num_nodes(mpm_handle h) { return h ? h->num_nodes() : -1; }
int64_t mpm_steps_taken(mpm_handle h) { return h ? h->steps : -1; }

int mpm_kernel_time(mpm_handle h, const char* name, double* avg_ms, int64_t* launches) {
    if (!h || !name) return fail(MPM_ERR_INVALID, "NULL argument");
    int k = -1;
    for (int i = 0; i < kNumKernels; ++i)
        if (std::strcmp(name, kKernelNames[i]) == 0) k = i;
    if (k < 0) return fail(MPM_ERR_INVALID, std::string("unknown kernel name: ") + name);
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int rc = h->resolve_events();
    if (rc != MPM_OK) return rc;
    const auto& t = h->timers[k];
    if (avg_ms) *avg_ms = t.launches ? t.total_ms / double(t.launches) : 0.0;
    if (launches) *launches = t.launches;
    return MPM_OK;
}

int mpm_reset_kernel_times(mpm_handle h) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int rc = h->resolve_events();
    if (rc != MPM_OK) return rc;
    for (auto& t : h->timers) t = KernelTimer{};
    return MPM_OK;
}

int mpm_set_timing(mpm_handle h, int enable) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int rc = h->resolve_events();
    if (rc != MPM_OK) return rc;
    h->opts.timing = enable ? 1 : 0;
    return MPM_OK;
}

int mpm_active_nodes(mpm_handle h, int64_t* count) {
    if (!h || !count) return fail(MPM_ERR_INVALID, "NULL argument");
    int rc = h->sync();
    if (rc != MPM_OK) return rc;
    if (h->p.stamp == 0) {
        *count = 0;
        return MPM_OK;
    }
    unsigned long long c = 0;
    HIP_TRY(nclr_host::launch_active_sum(h->args(), h->active_sum));
    HIP_TRY(hipMemcpyAsync(&c, h->active_sum, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *count = int64_t(c);
    return MPM_OK;
}

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
    return "nclr_mpm 0.2 gfx950 (fp32, no fp contraction, IEEE div/sqrt, tile-sorted SoA particles, "
           "LDS tile P2G)";
}

}  // extern "C"
