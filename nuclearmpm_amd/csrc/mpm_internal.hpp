// Internal types shared by the kernels (mpm_kernels.hip) and the C ABI (mpm_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/nclr_mpm.h"

namespace nclr_dev {

constexpr uint32_t kStatusOutOfRange = 1u;

// Scalars of one simulation, passed by value to every kernel.
struct StepParams {
    int res;            // grid resolution; nodes per axis = res + 1
    int nb;             // node blocks per axis
    int identity_mode;  // 0 reference diag<dim>(1), 1 physical identity
    uint32_t stamp;     // step stamp written by P2G into the touched node blocks
    float dt, dx, inv_dx, gravity;
    float mu0, lambda0;
    float allowed;      // grid velocity clamp, float(double(dx) * 0.9 / double(dt))
    int64_t n;          // particles
    int64_t stride;     // floats between consecutive fields of the SoA state
};

}  // namespace nclr_dev

namespace nclr_host {

struct LaunchArgs {
    int dim;
    int material;
    nclr_dev::StepParams p;
    float* state;
    float4* acc;
    float4* gv;
    uint32_t* stamps;
    int64_t nblocks;
    uint32_t* status;
    uint32_t* active_count;
    hipStream_t stream;
};

hipError_t launch_p2g(const LaunchArgs& a);
hipError_t launch_grid(const LaunchArgs& a);
hipError_t launch_g2p(const LaunchArgs& a);
hipError_t launch_export_grid(const LaunchArgs& a, float* vel, float* mass, int64_t total);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s);
hipError_t launch_zero_f4(float4* p, int64_t n, hipStream_t s);
int num_fields(int dim);

}  // namespace nclr_host
