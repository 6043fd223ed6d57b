// mpm_sort_keys_noalloc: the engine's stable sort of device keys with no allocation.  Its
// temporary buffers (a copy of the keys, the ping-pong key and value arrays of the radix passes
// and their digit histograms) live in `scratch`, a fixed device array of this library; the caller
// passes only its input and output buffers.  For callers that cannot allocate device memory
// (code under stream capture, serving loops that must not call hipMalloc).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mpm_internal.hpp"

namespace {

constexpr int64_t kNoAllocMaxKeys = int64_t(1) << 20;
// keys copy, keys_tmp, vals_tmp (kNoAllocMaxKeys each) + the histogram of the largest sort
constexpr int64_t kHistWords = 256 * ((kNoAllocMaxKeys + 4095) / 4096) + 256;
constexpr int64_t kScratchWords = 3 * kNoAllocMaxKeys + kHistWords;

// One copy per GPU the library runs on, held until the process ends.  In host code this name is
// only a handle for hipGetSymbolAddress: every use takes the device address from device_scratch.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
uint32_t* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (made current by the caller), looked up once per device.
hipError_t device_scratch(int device, uint32_t** out) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace

extern "C" int mpm_sort_keys_noalloc(const uint32_t* keys, int64_t n, int bits, uint32_t* perm, void* stream,
                                     int device) {
    if (n < 0 || (n > 0 && (!keys || !perm))) return nclr_host::set_last_error(MPM_ERR_INVALID, "bad arguments");
    if (bits < 0 || bits > 32) return nclr_host::set_last_error(MPM_ERR_INVALID, "bits must be in [0, 32]");
    if (n > kNoAllocMaxKeys)
        return nclr_host::set_last_error(
            MPM_ERR_INVALID, ("mpm_sort_keys_noalloc sorts at most " + std::to_string(kNoAllocMaxKeys) +
                              " keys (its fixed scratch); use mpm_sort_keys for more").c_str());
    if (device < 0 || device >= kMaxDevices) return nclr_host::set_last_error(MPM_ERR_INVALID, "bad device");
    if (n == 0) return MPM_OK;
    hipError_t e = hipSetDevice(device);
    uint32_t* base = nullptr;
    if (e == hipSuccess) e = device_scratch(device, &base);
    if (e != hipSuccess)
        return nclr_host::set_last_error(MPM_ERR_HIP, (std::string("mpm_sort_keys_noalloc: ") + hipGetErrorString(e)).c_str());
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t* kcopy = base;
    uint32_t* ktmp = base + kNoAllocMaxKeys;
    uint32_t* vtmp = base + 2 * kNoAllocMaxKeys;
    uint32_t* hist = base + 3 * kNoAllocMaxKeys;
    // every histogram entry and digit total is written by its pass before it is read: nothing
    // carried over from an earlier call is used
    e = hipMemcpyAsync(kcopy, keys, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = nclr_host::radix_sort_pairs(kcopy, perm, ktmp, vtmp, hist, n, bits, true, s);
    if (e != hipSuccess)
        return nclr_host::set_last_error(MPM_ERR_HIP, (std::string("mpm_sort_keys_noalloc: ") + hipGetErrorString(e)).c_str());
    return MPM_OK;
}
