// C ABI of the MLS-MPM engine (declarations and contracts: include/nclr_mpm.h).
//
// One handle = one simulation = one HIP stream.  The handle owns the
// structure-of-arrays particle state, the blocked grid buffers and a device
// status word that kernels set when a particle leaves the grid; the status is
// read back (and the error reported) only by synchronising calls, so that
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

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
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

constexpr int kBlockNodes = 64;

int block_log(int dim) { return dim == 2 ? 3 : 2; }

}  // namespace

struct nclr_mpm_sim {
    int dim = 0, material = 0, res = 0;
    float E = 0, nu = 0;
    mpm_options opts{};
    nclr_dev::StepParams p{};
    hipStream_t stream = nullptr;
    float* state = nullptr;
    float4* acc = nullptr;
    float4* gv = nullptr;
    uint32_t* stamps = nullptr;
    uint32_t* status = nullptr;
    uint32_t* active_count = nullptr;
    int64_t nblocks = 0;
    int64_t steps = 0;
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> pool;
    KernelTimer timers[kNumKernels];

    nclr_host::LaunchArgs args() const {
        nclr_host::LaunchArgs a{};
        a.dim = dim;
        a.material = material;
        a.p = p;
        a.state = state;
        a.acc = acc;
        a.gv = gv;
        a.stamps = stamps;
        a.nblocks = nblocks;
        a.status = status;
        a.active_count = active_count;
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
        int rc = timed(kP2G, [&] { return nclr_host::launch_p2g(a); });
        if (rc != MPM_OK) return rc;
        rc = timed(kGrid, [&] { return nclr_host::launch_grid(a); });
        if (rc != MPM_OK) return rc;
        rc = timed(kG2P, [&] { return nclr_host::launch_g2p(a); });
        if (rc != MPM_OK) return rc;
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

    // Host particle-major arrays -> device SoA for the fields given (NULL = keep `fallback`).
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
            if (Jp) host[(fJ)*stride + i] = Jp[i];
            if (mass) host[(fJ + 1) * stride + i] = mass[i];
            if (volume) host[(fJ + 2) * stride + i] = volume[i];
        }
        HIP_TRY(hipMemcpyAsync(state, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    int download(std::vector<float>& host) {
        host.assign(size_t(nf()) * size_t(p.stride), 0.0f);
        HIP_TRY(hipMemcpyAsync(host.data(), state, host.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return MPM_OK;
    }

    void release() {
        if (opts.timing) (void)resolve_events();
        for (auto e : pool) (void)hipEventDestroy(e);
        pool.clear();
        if (state) (void)hipFree(state);
        if (acc) (void)hipFree(acc);
        if (gv) (void)hipFree(gv);
        if (stamps) (void)hipFree(stamps);
        if (status) (void)hipFree(status);
        if (active_count) (void)hipFree(active_count);
        if (stream) (void)hipStreamDestroy(stream);
        state = nullptr;
        acc = gv = nullptr;
        stamps = status = active_count = nullptr;
        stream = nullptr;
    }
};

extern "C" {

void mpm_default_options(mpm_options* opts) {
    if (!opts) return;
    std::memset(opts, 0, sizeof(*opts));
    opts->identity_mode = 0;
    opts->sort_every = 0;
    opts->p2g_variant = 0;
    opts->block_threads = 0;
    opts->device = 0;
    opts->timing = 0;
}

int mpm_create(int dim, int material, int res, float dt, float E, float nu, float gravity, int64_t n, const float* x,
               const float* v, const float* F, const float* C, const float* Jp, const float* mass,
               const float* volume, const mpm_options* opts, mpm_handle* out) {
    if (!out) return fail(MPM_ERR_INVALID, "out handle pointer is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(MPM_ERR_INVALID, "dim must be 2 or 3");
    if (material < MPM_SNOW || material > MPM_LIQUID) return fail(MPM_ERR_INVALID, "unknown material model");
    if (res < 3 || res > 4096) return fail(MPM_ERR_INVALID, "res must be in [3, 4096]");
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
    h->E = E;
    h->nu = nu;
    auto& p = h->p;
    p.res = res;
    const int B = 1 << block_log(dim);
    p.nb = (res + 1 + B - 1) / B;
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
    h->nblocks = 1;
    for (int a = 0; a < dim; ++a) h->nblocks *= p.nb;

    auto cleanup = [&](int rc) {
        h->release();
        delete h;
        return rc;
    };
    if (hipSetDevice(h->opts.device) != hipSuccess) return cleanup(fail(MPM_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "hipStreamCreate failed"));
    const size_t nodes = size_t(h->nblocks) * kBlockNodes;
    const size_t state_floats = size_t(h->nf()) * size_t(p.stride);
    if (hipMalloc(&h->state, state_floats * sizeof(float)) != hipSuccess ||
        hipMalloc(&h->acc, nodes * sizeof(float4)) != hipSuccess ||
        hipMalloc(&h->gv, nodes * sizeof(float4)) != hipSuccess ||
        hipMalloc(&h->stamps, size_t(h->nblocks) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&h->status, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&h->active_count, 2 * sizeof(uint32_t)) != hipSuccess)
        return cleanup(fail(MPM_ERR_HIP, "hipMalloc failed (out of device memory?)"));
    if (nclr_host::launch_zero_f4(h->acc, int64_t(nodes), h->stream) != hipSuccess ||
        nclr_host::launch_zero_f4(h->gv, int64_t(nodes), h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->stamps, h->nblocks, 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->status, 1, 0u, h->stream) != hipSuccess ||
        nclr_host::launch_fill_u32(h->active_count, 2, 0u, h->stream) != hipSuccess)
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
    rc = h->download(host);
    if (rc != MPM_OK) return rc;
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
    rc = h->download(host);
    if (rc != MPM_OK) return rc;
    return h->upload(x, v, F, C, Jp, mass, volume, host);
}

int mpm_get_grid(mpm_handle h, float* velocity, float* mass) {
    if (!h) return fail(MPM_ERR_INVALID, "NULL handle");
    int rc = h->sync();
    if (rc != MPM_OK) return rc;
    const int64_t total = h->num_nodes();
    float* dvel = nullptr;
    float* dmass = nullptr;
    HIP_TRY(hipMalloc(&dvel, size_t(total) * h->dim * sizeof(float)));
    if (hipMalloc(&dmass, size_t(total) * sizeof(float)) != hipSuccess) {
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
    uint32_t c[2] = {0, 0};
    HIP_TRY(hipMemcpy(c, h->active_count, sizeof(c), hipMemcpyDeviceToHost));
    *count = h->p.stamp == 0 ? 0 : int64_t(c[h->p.stamp & 1]);
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
    return "nclr_mpm 0.1 gfx950 (fp32, no fp contraction, IEEE div/sqrt, SoA particles, blocked grid)";
}

}  // extern "C"
