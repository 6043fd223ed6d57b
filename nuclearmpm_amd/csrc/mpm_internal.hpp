// Internal types shared by the kernels (mpm_step.hip, mpm_sort.hip) and the C ABI (mpm_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/nclr_mpm.h"

namespace nclr_dev {

constexpr uint32_t kStatusOutOfRange = 1u;

// Particle tiles: S^D cells (S = 16 in 2D, 8 in 3D).  A tile's P2G region covers the nodes its
// particles can touch, (S+2)^D nodes from the tile origin.  Node tiles hold S^D nodes each.
template <int D>
struct Tile;
template <>
struct Tile<2> {
    static constexpr int LOG = 4;
    static constexpr int S = 16;
    static constexpr int R = 18;           // region side
    static constexpr int REGION = 18 * 18;  // region nodes
    static constexpr int NODES = 16 * 16;   // nodes per node tile
    static constexpr int CHUNK = 1024;      // particles per P2G work item
};
template <>
struct Tile<3> {
    static constexpr int LOG = 3;
    static constexpr int S = 8;
    static constexpr int R = 10;
    static constexpr int REGION = 10 * 10 * 10;
    static constexpr int NODES = 8 * 8 * 8;
    static constexpr int CHUNK = 512;
};

// Scalars of one simulation, passed by value to every kernel.
struct StepParams {
    int res;            // grid resolution; nodes per axis = res + 1
    int nt;             // particle tiles per axis
    int ntn;            // node tiles per axis
    int identity_mode;  // 0 reference diag<dim>(1), 1 physical identity
    uint32_t stamp;     // step stamp written into the node tiles the step touched
    uint32_t ntiles;    // nt^D; also the key of particles outside the grid
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
    const float* state_in;     // SoA state, current physical order
    float* state_out;          // SoA state written by G2P in sorted order
    const uint32_t* pid_in;    // original particle index of each slot
    uint32_t* pid_out;
    uint32_t* keys;            // n sort keys (tile of each particle) -> sorted keys
    uint32_t* perm;            // n: sorted position -> current slot
    uint32_t* keys_tmp;
    uint32_t* perm_tmp;
    uint32_t* hist;            // 256 * sort tiles + 256
    uint32_t* tile_start;      // ntiles + 1
    uint32_t* tile_end;        // ntiles + 1
    uint32_t* item_offset;     // ntiles + 1: exclusive scan of P2G work items per tile
    float4* slots;             // max_items * REGION: the P2G region of every work item
    float4* gv;                // node tiles * NODES, updated grid
    uint32_t* stamps;          // node tiles
    uint32_t* status;
    uint32_t* active;          // node tiles: nodes with post-P2G mass > 0
    hipStream_t stream;
};

hipError_t launch_sort(const LaunchArgs& a);
hipError_t launch_p2g(const LaunchArgs& a);
hipError_t launch_grid(const LaunchArgs& a);
hipError_t launch_g2p(const LaunchArgs& a);
hipError_t launch_export_grid(const LaunchArgs& a, float* vel, float* mass, int64_t total);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s);
hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t s);
hipError_t launch_zero_f4(float4* p, int64_t n, hipStream_t s);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* hist,
                            int64_t n, int bits, bool identity_values, hipStream_t s);
hipError_t launch_active_sum(const LaunchArgs& a, unsigned long long* out);
int64_t sort_tiles(int64_t n);
int64_t max_items(int dim, int64_t n, uint32_t ntiles);
int num_fields(int dim);
int region_nodes(int dim);
int tile_nodes(int dim);
int tile_log(int dim);

}  // namespace nclr_host
