// Internal types shared by the kernels (mpm_step.hip, mpm_sort.hip) and the C ABI (mpm_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/nclr_mpm.h"

namespace nclr_dev {

constexpr uint32_t kStatusOutOfRange = 1u;
// Tiles per round of the tile scans (one workgroup of 1024 threads, 4 tiles per thread).
constexpr uint32_t kScanRound = 4096;
constexpr uint32_t kStatusMovedTooFar = 2u;
constexpr uint32_t kStatusBarrierTimeout = 4u;  // a grid-wide barrier of the compact step waited past its limit  // a particle crossed more than one tile in a step
constexpr int kMixed = 3;  // material template value: per-scene material chosen at run time

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
};
template <>
struct Tile<3> {
    static constexpr int LOG = 3;
    static constexpr int S = 8;
    static constexpr int R = 10;
    static constexpr int REGION = 10 * 10 * 10;
    static constexpr int NODES = 8 * 8 * 8;
};

// Smallest P2G work item (particles per workgroup); the slot buffer is sized for it.
constexpr int kMinChunk = 256;

// One scene (an independent simulation) of a batch.
struct SceneParams {
    int res;             // grid resolution; nodes per axis = res + 1
    int nt;              // particle tiles per axis
    int ntn;             // node tiles per axis
    int material;        // MPM_SNOW / MPM_JELLY / MPM_LIQUID
    int identity_mode;   // 0 reference diag<dim>(1), 1 physical identity
    uint32_t tile_base;  // first global particle tile
    uint32_t node_tile_base;
    float dt, dx, inv_dx, gravity;
    float mu0, lambda0;
    float allowed;       // grid velocity clamp, float(double(dx) * 0.9 / double(dt))
    float dmu_dE, dlambda_dE;  // d mu0 / dE = 1 / (2 (1 + nu)), d lambda0 / dE = nu / ((1 + nu)(1 - 2 nu))
    int64_t particle_base;  // first particle (original index)
    int64_t n;
};

// Launch-wide values, passed by value to every kernel.
struct StepParams {
    int nscenes;
    uint32_t ntiles;      // particle tiles of all scenes; also the key of particles off their grid
    uint32_t node_tiles;  // node tiles of all scenes
    int chunk;            // P2G particles per work item (256, 512 or 1024)
    int64_t n;            // particles of all scenes
    int64_t stride;       // floats between consecutive fields of the SoA state
    const SceneParams* scenes;
    const uint32_t* stamp;  // device word: stamp of the current step (incremented by the key kernel)
    uint4* items;           // max_items: per P2G work item {tile, first, end, 0}, its tile and sorted range
                            // (written with item_offset)
    const uint32_t* tile_scene;       // ntiles: scene of each tile (batches of more than one scene)
    const uint32_t* node_tile_scene;  // node_tiles: scene of each node tile (likewise)
    // Per scene: status bits the sort raised since the host last read them (a particle off its
    // grid), and the device step stamp of the step that found the scene's first failure (0: live).
    // A scene with either word set is stopped: its particles are carried through unchanged and
    // its grid is not touched, while the other scenes keep stepping.
    uint32_t* scene_flags;
    uint32_t* scene_frozen;
};

// The compact step (mpm_compact.hip): dynamic LDS per workgroup for its box of nodes.
constexpr int kCompactLdsBytes = 150 * 1024;
constexpr int kCompactMaxWorkgroups = 256;
constexpr int kCompactTableWords = 3 * kCompactMaxWorkgroups * 4;  // 64-bit words of the sub-box tables
constexpr int kCompactMaxRes = 65000;  // node coordinates travel in 16 bits in the sub-box table

// Words of CompactArgs::info, written by the compact step at the end of its launch.
enum CompactInfo : int {
    kInfoDone = 0,     // steps completed by the launch
    kInfoLo = 1,       // [3] origin of the box of the last completed step's grid
    kInfoExt = 4,      // [3] its extent
    kInfoNodes = 7,    // its nodes
    kInfoReason = 8,   // why the launch ended: kCompactDone, kCompactOverflow or kCompactStopped
    kInfoOutside = 9,  // raised by a particle outside the box (the step is retried over the whole grid)
    kInfoStamp = 10,   // device stamp of the last completed step
    kInfoWords = 16
};
enum CompactReason : int {
    kCompactDone = 0,      // every step completed
    kCompactOverflow = 1,  // the box outgrew the LDS: the remaining steps are left to the tiled step
    kCompactStopped = 2    // a particle left the grid: the scene stopped (the stamp counts every step)
};

// Arguments of the compact step kernel (one scene, many steps per launch).
struct CompactArgs {
    float* st;             // SoA state, updated in place
    int64_t stride;
    int64_t n;
    int nsteps;
    int slot_sets;         // workgroup slots per step parity (>= the launch's workgroups)
    SceneParams sp;
    StepParams step;       // the scene's status and stop words (scene_flags, scene_frozen)
    uint32_t* stamp;       // device step stamp: + 1 per completed step
    uint32_t* status;      // engine status word (kStatusBarrierTimeout)
    float4* slots;         // 2 x slot_sets x box capacity: each workgroup's deposits, by step parity
    // 3 x kCompactMaxWorkgroups x 4 words, by step barrier count mod 3: each workgroup's sub-box
    // [lo[3], hi[3]] (the nodes it deposited to) and its flags, one 32-bit half per value, each half
    // (stamp << 16) | value
    unsigned long long* subbox;
    float4* grid;          // box capacity: the updated grid of the box of the last completed step
    uint8_t* massive;      // box capacity: 1 where that step's node mass after P2G is > 0
    int* box;              // [min[D], max[D]]: base range of the positions before the first step
    uint32_t* bar;         // [2]: grid barrier arrivals and generation
    int* info;             // kInfoWords words (CompactInfo)
};

}  // namespace nclr_dev

namespace nclr_host {

// Fewer full work items than this (the chip's CUs) make a 2D launch latency-bound: P2G and the G2P
// adjoint then run one particle per thread.
constexpr int kFewItems = 256;

struct LaunchArgs {
    int dim;
    int material;  // common material of all scenes, or nclr_dev::kMixed
    int g2p_threads;
    int svd_mode;   // 3D SVDs and polar factors: 0 nclr_svd's operation order (bitwise the oracle), 1 fast
                    // FP32 path, 2 accurate FP64 path (the adjoint kernels take nclr_svd's order for 0 and 2)
    nclr_dev::StepParams p;
    uint32_t* stamp;           // device stamp word (same as p.stamp)
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
    float4* gpre;              // optional (nullptr): node tiles * NODES, the grid's {momentum, mass} before
                               // the update (written by the grid kernel, read by the grid adjoint)
    uint32_t* stamps;          // node tiles
    uint32_t* status;
    uint32_t* active;          // node tiles: nodes with post-P2G mass > 0
    // incremental sort (mpm_bucket.hip)
    uint32_t* uidx;            // n: old work item of each slot
    uint32_t* ucnt;            // max_items * offsets: per old item, particles per neighbour offset
    uint32_t* upref;           // max_items * offsets: per (old item, offset), earlier particles bound alike
    uint32_t* ttotal;          // ntiles: particles per new tile
    uint32_t* meta;            // 4 counters
    uint32_t* rtot;            // 2 per round of the tile scans: particle and work-item totals
    // the incremental sort's inputs from the step before: its item offsets, item table and meta.
    // The same buffers as item_offset, p.items and meta, unless the step sorts into another block.
    const uint32_t* prev_item_offset;
    const uint4* prev_items;
    const uint32_t* prev_meta;
    hipStream_t stream;
};

// Buffers of one adjoint step (mpm_adjoint.hip).  Cotangents are structure-of-arrays with
// Layout<D>::NG channels (x, v, F, C) of `gstride` floats each.
struct AdjointArgs {
    const float* grad_out;  // cotangent of the state after the step, in sorted order (its slot order)
    float* grad_in;         // cotangent of the state before the step, in the current slot order
    int64_t gstride;
    float4* slots_b;        // max_items * REGION: per work item, the G2P adjoint's grid cotangent
    float4* gbar;           // node tiles * NODES: cotangent of the pre-update grid {momentum, mass}
    float* epart;           // n: per sorted particle, its contribution to dL/dE of its scene
    double* grad_E;         // nscenes: accumulated dL/dE
    int64_t* erange;        // 2 nscenes: each scene's sorted range [lo, hi)
    double* epartial;       // nscenes * echunks: partial sums of dL/dE
    int echunks;            // e_chunks(largest scene)
};

// Chunks of the fixed-order dL/dE sum for a scene of n particles.
int e_chunks(int64_t n);

hipError_t launch_g2p_adjoint(const LaunchArgs& a, const AdjointArgs& b);
hipError_t launch_grid_adjoint(const LaunchArgs& a, const AdjointArgs& b);
hipError_t launch_p2g_adjoint(const LaunchArgs& a, const AdjointArgs& b);
hipError_t launch_permute_channels(const uint32_t* pid, const float* src, int64_t src_stride, float* dst,
                                   int64_t dst_stride, int channels, int64_t n, bool to_particle_order,
                                   hipStream_t s);
hipError_t launch_sort(const LaunchArgs& a);
// The stable sort by tile for a state that is stored in the order of the previous sort, whose
// tile bounds and work items are still in tile_start / tile_end / item_offset (same chunk size).
hipError_t launch_sort_incremental(const LaunchArgs& a);
int bucket_offsets(int dim);
hipError_t launch_p2g(const LaunchArgs& a);
hipError_t launch_grid(const LaunchArgs& a);
hipError_t launch_g2p(const LaunchArgs& a);
hipError_t launch_export_grid(const LaunchArgs& a, int scene, int res, uint32_t node_tile_base, int ntn,
                              float* vel, float* mass, int64_t total);
hipError_t launch_svd(int dim, int mode, const float* A, float* U, float* S, float* V, int64_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s);
hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t s);
hipError_t launch_bump_u32(uint32_t* p, hipStream_t s);  // *p += 1 on the device
hipError_t launch_zero_f4(float4* p, int64_t n, hipStream_t s);
hipError_t launch_active_sum(const LaunchArgs& a, uint32_t first_tile, uint32_t tiles, unsigned long long* out);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* hist,
                            int64_t n, int bits, bool identity_values, hipStream_t s);
int64_t sort_tiles(int64_t n);
// scene generators and the viewer projection (mpm_scene.hip)
void rand_jump(uint32_t state[4], uint64_t k);
std::vector<uint4> rand_starts(const uint32_t state[4], int64_t count);
hipError_t launch_gen_cubes(int dim, const mpm_cube* cubes, const int64_t* offsets, int ncubes, int identity_mode,
                            int64_t n, float* img, hipStream_t s);
hipError_t launch_gen_rand(const uint4* starts, int64_t count, float* out, hipStream_t s);
hipError_t launch_gen_sample(int dim, const uint4* starts, const float* centers, int per_center, float half,
                             int identity_mode, int64_t n, float* img, hipStream_t s);
hipError_t launch_eval_gyroid(int64_t n, const float* pos, float k, float t, float* out, hipStream_t s);
int64_t gyroid_blocks(const mpm_cube& box);
hipError_t launch_gyroid_count(const mpm_cube& box, float k, float t, float thickness, uint32_t* counts,
                               int64_t* offsets, hipStream_t s);
hipError_t launch_gyroid_scatter(const mpm_cube& box, float k, float t, float thickness, const int64_t* offsets,
                                 int identity_mode, int64_t n, float* img, hipStream_t s);
hipError_t launch_project(int64_t n, const float* pos, int centered, float* out, hipStream_t s);
int64_t max_items(int64_t n, uint32_t ntiles);
// The compact step: nodes of the largest box; one cooperative launch of nwg workgroups of 256
// threads with ppt particles each running a.nsteps steps of a single-scene engine.
int compact_capacity(int dim);
hipError_t launch_compact_steps(int dim, int material, int mode, int ppt, int nwg, const nclr_dev::CompactArgs& a,
                                hipStream_t s);
// The slot order of a compact launch: the state (every field) and the slot -> particle map
// gathered into st_out / pid_out in Hilbert order of the particles' cells (keys, perm and the
// temporaries are the tile sort's buffers).
hipError_t launch_compact_order(int dim, const float* st, const uint32_t* pid, float* st_out, uint32_t* pid_out,
                                int64_t n, int64_t stride, int fields, const nclr_dev::SceneParams& sp, uint32_t* keys,
                                uint32_t* perm, uint32_t* keys_tmp, uint32_t* perm_tmp, uint32_t* hist,
                                hipStream_t s);
// Writes the compact step's last grid into the tiled grid (gv, stamps, active) of a single-scene
// engine, so that the grid export and active-node count read it as the tiled step's.
hipError_t launch_compact_to_tiles(int dim, const nclr_dev::CompactArgs& a, float4* gv, uint32_t* stamps,
                                   uint32_t* active, hipStream_t s);
// Records msg as the calling thread's mpm_last_error() and returns code.
int set_last_error(int code, const char* msg);
int num_fields(int dim);
int region_nodes(int dim);
int tile_nodes(int dim);
int tile_log(int dim);

}  // namespace nclr_host
