/*
 * nclr_mpm.h - C ABI of the MI355X (gfx950) MLS-MPM engine.
 *
 * The reference engine (jparr721/NuclearMPM) is a header-only C++ class with no FFI; these entry
 * points are what a binding of its step loop needs.  Each declaration names the reference
 * interface it replaces (or says that it has none).  The C++ drop-in class in include/nclr/nclr.h
 * and the Python package nuclearmpm_amd are both written on top of this ABI.
 *
 * Conventions
 *  - Host arrays are particle-major: x[n][dim], v[n][dim], F[n][dim][dim] (row-major),
 *    C[n][dim][dim], Jp[n], mass[n], volume[n].  Device state is structure-of-arrays inside the
 *    library; host buffers are copied in/out.
 *  - Grid arrays use the reference's flat node order, x-major with the last axis fastest,
 *    (res+1)^dim nodes: velocity[node][dim], mass[node].
 *  - A handle advances one or more independent scenes (a batch); the particles of scene s are
 *    rows [sum of earlier scene sizes, + n_s) of every particle array.
 *  - Every function returns MPM_OK (0) or an MPM_ERR_* code; mpm_last_error() gives the message
 *    of the last failure on the calling thread.
 *  - One handle owns one HIP stream and must be driven from one host thread at a time;
 *    different handles are independent.
 */
#ifndef NCLR_MPM_H
#define NCLR_MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nclr_mpm_sim* mpm_handle;

/* MaterialModel (src/nclr.h:57-61): same values. */
enum mpm_material { MPM_SNOW = 0, MPM_JELLY = 1, MPM_LIQUID = 2 };

enum mpm_status {
    MPM_OK = 0,
    MPM_ERR_INVALID = 1,      /* bad argument */
    MPM_ERR_OUT_OF_RANGE = 2, /* a particle's stencil left the grid (std::out_of_range in src/nclr.h:163-164) */
    MPM_ERR_HIP = 3,          /* HIP runtime failure */
    MPM_ERR_UNSUPPORTED = 4   /* feature not available for this configuration */
};

typedef struct mpm_options {
    /* 0: reference mode, diag<dim>(1) sets (0,0) and (1,1) only (src/nclr_math.h:13-19), so the
     *    3D "identity" of the particle ctor, the F update and the liquid reset is diag(1,1,0)
     *    (src/nclr.h:47,230,256).  1: physical identity. */
    int32_t identity_mode;
    /* HIP device ordinal. */
    int32_t device;
    /* Record per-kernel HIP events (mpm_kernel_time). */
    int32_t timing;
    /* Particles per P2G work item: 256, 512 or 1024 (0 = 512). */
    int32_t p2g_chunk;
    /* G2P workgroup size: 64, 128 or 256 (0 = 256). */
    int32_t g2p_threads;
    /* 1: launch each step as one replay of a captured hipGraph (ignored while timing is on). */
    int32_t graph;
    /* 1: choose p2g_chunk, g2p_threads and (single-scene engines) the compact step at creation by
     * timing 20 trial steps of each (state restored); see mpm_autotune. */
    int32_t autotune;
    /* 3D singular value decompositions (the polar factor in P2G, the snow / liquid SVD in G2P):
     * 0 = nclr_svd's two-sided Jacobi in Eigen's operation order (src/nclr_math.h:50-74), bitwise
     * the CPU restatement's FP32 arithmetic (the default; the drop-in's mode);
     * 1 = a fixed-sweep FP32 symmetric Jacobi SVD of A^T A with the same convention (U, V proper
     * rotations, descending order, the sign of det on the last singular value): cheaper, but its
     * rotation factor carries ~1e-7 absolute error, which the elastic stress (F - R) amplifies at
     * small strains;
     * 2 = accurate: F - R by Newton's polar iteration in FP64 (two-sided Jacobi where det F <= 0 or
     * F is ill-conditioned) and the snow / liquid SVD by an FP64 Jacobi eigen-decomposition of
     * A^T A; closer to the FP64 step than the reference's FP32 arithmetic is, and faster than 0.
     * The P2G adjoint takes mode 2's Newton polar factor in mode 2; the G2P adjoint takes mode 0's
     * factors for modes 0 and 2. */
    int32_t fast_svd;
    /* Single-scene engines: the compact step with this many particles per thread (1, 2 or 4; 3D: 1
     * or 2), 0 = off (the default).  See mpm_set_compact. */
    int32_t compact;
    int32_t reserved[7];
} mpm_options;

/* One independent scene of a batch: the reference ctor's arguments (src/nclr.h:74-75). */
typedef struct mpm_scene {
    int32_t material;
    int32_t res;
    float dt, E, nu, gravity;
    int64_t n; /* particles of this scene */
} mpm_scene;

/* Fills *opts with the defaults (reference identity mode, library-chosen launch shapes). */
void mpm_default_options(mpm_options* opts);

/* Replaces MPMSimulation<dim>::MPMSimulation(particles, model, res, dt, E, nu, gravity)
 * (src/nclr.h:74-78).  Copies the particle arrays to the device.  Any of v, F, C, Jp, mass,
 * volume may be NULL for the Particle ctor defaults (src/nclr.h:46-47).  opts may be NULL. */
int mpm_create(int dim, int material, int res, float dt, float E, float nu, float gravity, int64_t n,
               const float* x, const float* v, const float* F, const float* C, const float* Jp,
               const float* mass, const float* volume, const mpm_options* opts, mpm_handle* out);

/* Many independent simulations advanced by one launch sequence per step.  No reference counterpart
 * (the reference needs one MPMSimulation object per scene); scene s behaves exactly like
 * mpm_create with scenes[s]'s arguments on its particle rows. */
int mpm_create_batch(int dim, int nscenes, const mpm_scene* scenes, const float* x, const float* v,
                     const float* F, const float* C, const float* Jp, const float* mass, const float* volume,
                     const mpm_options* opts, mpm_handle* out);

/* Replaces nsteps calls of MPMSimulation<dim>::advance() (src/nclr.h:80-84) for every scene.  The
 * steps are queued on the handle's stream; the call returns without waiting.  Errors raised by
 * queued steps (a particle leaving the grid) are reported by the next synchronising call. */
int mpm_advance(mpm_handle h, int nsteps);

/* Waits for queued work; returns MPM_ERR_OUT_OF_RANGE if any step since the last check found a
 * particle outside its scene's grid (the reference throws std::out_of_range from std::vector::at,
 * src/nclr.h:163-164); the message names the scenes and the step of each failure.
 * Each scene fails on its own, as each MPMSimulation object of the reference does: the step that
 * finds one of a scene's particles off its grid leaves that scene's particles as they were before
 * the step (the state the reference's throwing advance() leaves), and the scene stays stopped;
 * later steps advance the other scenes and find the stopped scene again (every later advance() of
 * the reference throws again).  Setting or loading a state (mpm_set_particles, mpm_load_state,
 * mpm_set_state_device) restarts every scene.  A stopped scene's grid export shows its last
 * completed step. */
int mpm_synchronize(mpm_handle h);

/* Per-scene status without raising it as an error (no reference counterpart: a batch of the
 * reference is one MPMSimulation per scene, each throwing on its own): waits for queued work and
 * fills, for each scene s, flags[s] = the status bits found since the state was last set
 * (1: a particle's stencil left the grid, 2: a particle crossed more than one tile) and
 * failed_step[s] = the step that stopped the scene (1-based count of steps taken) or -1.  Either
 * pointer may be NULL.  The failures it returns are not reported again by mpm_synchronize until
 * a later step finds them again. */
int mpm_scene_status(mpm_handle h, uint32_t* flags, int64_t* failed_step);

/* Replaces MPMSimulation<dim>::particles() (src/nclr.h:86): copies the current particle state of
 * all scenes to host arrays in the original particle order.  Any pointer may be NULL. */
int mpm_get_particles(mpm_handle h, float* x, float* v, float* F, float* C, float* Jp, float* mass,
                      float* volume);

/* Replaces the particle vector the constructor takes by value (src/nclr.h:74,76): overwrites the
 * state of all particles (NULL pointers leave a field unchanged). */
int mpm_set_particles(mpm_handle h, const float* x, const float* v, const float* F, const float* C,
                      const float* Jp, const float* mass, const float* volume);

/* Replaces MPMSimulation<dim>::grid() (src/nclr.h:87) for scene 0: the grid after the last step's
 * grid update, velocity (or momentum where mass == 0) and mass per node, dense flat layout.
 * Before the first step the grid is empty and both arrays are zero-filled. */
int mpm_get_grid(mpm_handle h, float* velocity, float* mass);
/* The same for scene `scene` of a batch. */
int mpm_get_grid_scene(mpm_handle h, int scene, float* velocity, float* mass);

/* Scalars of scene 0: mu_0, lambda_0 (src/nclr.h:71-72), dx, inv_dx. */
int mpm_get_params(mpm_handle h, float* mu_0, float* lambda_0, float* dx, float* inv_dx);
int mpm_get_scene_params(mpm_handle h, int scene, float* mu_0, float* lambda_0, float* dx, float* inv_dx);
int64_t mpm_num_particles(mpm_handle h);
int64_t mpm_num_nodes(mpm_handle h);
int64_t mpm_steps_taken(mpm_handle h);
int mpm_num_scenes(mpm_handle h);

/* Kernel timing (timing on): average device milliseconds per launch of each kernel group since
 * the last reset, measured with HIP events on the handle's stream, and the launch counts.
 * names: "sort", "p2g", "grid", "g2p", "step" (whole steps, also recorded in graph mode).
 * Returns MPM_ERR_INVALID for an unknown name. */
int mpm_kernel_time(mpm_handle h, const char* name, double* avg_ms, int64_t* launches);
int mpm_reset_kernel_times(mpm_handle h);
/* Turns the per-kernel HIP event recording of mpm_kernel_time on (1) or off (0). */
int mpm_set_timing(mpm_handle h, int enable);

/* Launch shape of the step kernels (engine extras, no reference counterpart). */
int mpm_set_tuning(mpm_handle h, int p2g_chunk, int g2p_threads);
int mpm_get_tuning(mpm_handle h, int* p2g_chunk, int* g2p_threads);
/* The compact step (engine extra, no reference counterpart; for small single-scene engines, where
 * the tiled step's per-step launches and sort dominate): mpm_advance runs all its steps in one
 * cooperative launch of ceil(n / (256 particles_per_thread)) workgroups, at most one per CU, with no
 * sort: each workgroup deposits its particles into a fixed-point copy of the box of grid nodes the
 * particles can touch (at most 4266 nodes, in LDS), the grid update sums the workgroups' copies in
 * a fixed order, and G2P updates the particles in place.  Same step, same parity bound against the
 * reference and deterministic, but not bitwise the tiled step (its fixed-point sums are grouped by
 * workgroup, not by tile).  When the box outgrows the LDS the remaining steps run tiled (and the
 * step stays tiled until the state is set again); mpm_advance then waits for the launch, which it
 * does not when the whole grid fits (2D res <= 64, 3D res <= 15).  particles_per_thread: 1, 2 or 4
 * (3D: 1 or 2), 0 = off.  MPM_ERR_UNSUPPORTED for batches and for more particles than one
 * co-resident launch holds.  Off while timing is on. */
int mpm_set_compact(mpm_handle h, int particles_per_thread);
int mpm_get_compact(mpm_handle h, int* particles_per_thread);
/* Times `steps_per_trial` steps of every launch shape (and, on single-scene engines, of the compact
 * step) on the current state, keeps the fastest and restores the state; returns the choice through
 * mpm_get_tuning and mpm_get_compact. */
int mpm_autotune(mpm_handle h, int steps_per_trial);

/* Number of grid nodes with mass > 0 after the last step's P2G in scene `scene` (-1: all scenes);
 * the G_active count of the roofline bytes. */
int mpm_active_nodes(mpm_handle h, int64_t* count);
int mpm_active_nodes_scene(mpm_handle h, int scene, int64_t* count);

/* The engine's stable counting sort (the per-step particle-by-tile sort) on n host keys:
 * perm[j] = index of the j-th key in stable order by the low `bits` bits (no reference
 * counterpart; the reference steps particles in index order, src/nclr.h:112,169). */
int mpm_sort_keys(const uint32_t* keys, int64_t n, int bits, uint32_t* perm, int device);

/* mpm_sort_keys on device buffers without allocating: no counterpart in the reference, for callers
 * that cannot allocate device memory (code under stream capture, serving loops that must not call
 * hipMalloc).  keys (n, input, unchanged) and perm (n, output) are device memory on `device`; the
 * sort is queued on `stream` (a hipStream_t; NULL: the null stream) and the call returns without
 * waiting.  The temporary buffers live in a fixed device array of the library (about 12.9 MB per
 * GPU, kept until the process ends): n is at most 2^20 (larger n: MPM_ERR_INVALID).  Calls that
 * overlap on one GPU share that array and race: serialise them (one stream, or wait between). */
int mpm_sort_keys_noalloc(const uint32_t* keys, int64_t n, int bits, uint32_t* perm, void* stream, int device);

/* The step's device SVD on n host matrices (dim x dim, row-major): A = U diag(sigma) V^T with
 * nclr_svd's convention (src/nclr_math.h:50-74: descending order; in 3D U, V proper rotations and
 * the sign of det A on sigma[2]).  exact = 1: nclr_svd's operation order, 0: the fast FP32 path
 * (mpm_options.fast_svd = 1), 2: the accurate FP64 path (mpm_options.fast_svd = 2); 2D always uses
 * nclr_svd's order. */
int mpm_svd(int dim, int64_t n, const float* A, float* U, float* sigma, float* V, int exact, int device);

/* ---- Differentiation (no reference counterpart: the reference has no adjoint) ----------------
 *
 * Device-pointer calls.  Every pointer below is device memory on the handle's device; the calls
 * are stream-ordered on the handle's stream and return before the device finishes: the caller
 * makes its inputs ready before the call (e.g. synchronises the stream that produced them) and
 * calls mpm_synchronize before it reads the outputs from another stream.
 *
 * State cotangents and device state images are structure-of-arrays over the n particles of the
 * handle: channel c of particle i at g[c*n + i], with the mpm_grad_channels(dim) = 2 dim + 2 dim^2 + 1
 * channels x[dim], v[dim], F[dim*dim], C[dim*dim] (matrices row-major), Jp.  "Particle order" is the
 * order of the arrays given to mpm_create; "slot order" is the engine's current storage order
 * (particles are re-sorted by tile every step). */
int mpm_grad_channels(int dim);

/* Bytes of one state snapshot (all particle fields and the slot -> particle map). */
int64_t mpm_state_bytes(mpm_handle h);
/* Copies the current state into / out of a device snapshot of mpm_state_bytes(h) bytes
 * (checkpoints of a differentiable rollout). */
int mpm_save_state(mpm_handle h, void* dst);
int mpm_load_state(mpm_handle h, const void* src);

/* Device image of x, v, F, C, Jp in particle order (get), or sets them from one (set; mass and
 * volume are kept and the slot order becomes the particle order). */
int mpm_get_state_device(mpm_handle h, float* dst);
int mpm_set_state_device(mpm_handle h, const float* src);

/* Re-indexes a cotangent between particle order and the current slot order (src != dst). */
int mpm_grad_from_particle_order(mpm_handle h, const float* src, float* dst);
int mpm_grad_to_particle_order(mpm_handle h, const float* src, float* dst);

/* The adjoint entry points below return MPM_ERR_OUT_OF_RANGE while a scene is stopped.
 *
 * Vector-Jacobian product of one step taken from the current state (every material: jelly; snow
 * with the backward pass of its singular-value clamp, its Jp update and its hardening; liquid):
 * grad_out is the cotangent of the state after the step in ITS slot order,
 * grad_in (must not alias grad_out) receives the cotangent of the current state in the current
 * slot order, and grad_E[s] += dL/dE of scene s (nscenes doubles; NULL: not accumulated).  The
 * state and the step counter are unchanged; the grid export afterwards shows this step's grid. */
int mpm_backward_step(mpm_handle h, const float* grad_out, float* grad_in, double* grad_E);

/* Vector-Jacobian product of `nsteps` steps from the current state (a checkpoint): recomputes
 * the states inside the segment on the device into a tape of nsteps - 1 states and their sorts
 * (device memory, allocated on first use and grown to the longest segment), then runs nsteps
 * backward steps in reverse on the tape.  grad_out is in the slot order of the state nsteps steps ahead, grad_in in
 * the current slot order (they may alias).  The current state is left as it was. */
int mpm_backward_segment(mpm_handle h, int nsteps, const float* grad_out, float* grad_in, double* grad_E);

/* Device bytes of the tape mpm_backward_segment keeps for segments of `nsteps` steps (-1 for a
 * NULL handle or nsteps < 1): nsteps - 1 entries of one state snapshot (mpm_state_bytes) and one
 * sort (8 bytes per particle, 12 bytes per particle tile of every scene for its bounds and work
 * item offset, 16 bytes per possible work item, 256-byte aligned), and nsteps entries of the
 * forward grid and its pre-update {momentum, mass} (32 bytes per node of every scene's node tiles;
 * kept only while it leaves a quarter of the free device memory, else each adjoint step
 * recomputes its grid). */
int64_t mpm_tape_bytes(mpm_handle h, int nsteps);

/* Recorded segments: no counterpart in the reference (whose checkpointing recomputes every
 * segment); for callers with device memory to spare, which trade it for the recomputation.
 *
 * mpm_record_bytes: device bytes of the tape of one recorded segment of `nsteps` steps (-1 for a
 * NULL handle or nsteps < 1): nsteps - 1 state snapshots and sorts as in mpm_tape_bytes, and with
 * with_grid != 0 nsteps forward grids and pre-update sums.
 *
 * mpm_advance_recorded: advances `nsteps` steps from the current state like mpm_advance (on the
 * tiled step) and records into `tape` (caller-owned device memory of mpm_record_bytes bytes,
 * 256-byte aligned) what mpm_backward_segment would recompute from this state.
 *
 * mpm_backward_recorded: mpm_backward_segment for a segment recorded by mpm_advance_recorded, run
 * from the segment's first state (restored by the caller, e.g. with mpm_load_state) without
 * recomputing the recorded steps; the gradients are bitwise those of mpm_backward_segment.  The
 * tape must be unchanged since its recording, and MPM_ERR_INVALID is returned for a tape that was
 * not recorded on this engine with these nsteps and with_grid, or when the tuning or a scene's
 * elasticity changed since.  The tape may be passed again (a second backward pass) until it is
 * recorded over or freed. */
int64_t mpm_record_bytes(mpm_handle h, int nsteps, int with_grid);
int mpm_advance_recorded(mpm_handle h, int nsteps, void* tape, int with_grid);
int mpm_backward_recorded(mpm_handle h, int nsteps, void* tape, int with_grid, const float* grad_out,
                          float* grad_in, double* grad_E);

/* Sets Young's modulus and Poisson's ratio of one scene (mu_0, lambda_0 recomputed in float as
 * src/nclr.h:77-78 does). */
int mpm_set_scene_elasticity(mpm_handle h, int scene, float E, float nu);

/* ---- Scene generators and viewer hooks on the device ------------------------------------------
 *
 * The reference's scene helpers (src/nclr_math.h:22-39, 100-147) and pt_3d_to_2d (149-167) as
 * device kernels.  Generators write a state image (the layout mpm_set_state_device takes:
 * channel c of particle i at image[c*n + i], channels x[dim], v[dim], F[dim*dim], C[dim*dim], Jp) with
 * v from the shape, F = diag<dim>(1) (diag(1, 1, 0) in 3D for identity_mode 0, I for 1), C = 0 and
 * Jp = 1.
 * `image`, `pos` and `out` are device memory; the calls run on the device's null stream and return
 * when the device is done.  A state built this way enters an engine through mpm_create (any x,
 * e.g. zeros) followed by mpm_set_state_device. */

/* One cube of particles: a res[0] x res[1] (x res[2]) lattice, axis a spanning
 * LinSpaced(res[a], lo[a], hi[a]), last axis fastest: cube<dim>(res, min, max) of
 * src/nclr_math.h:100-129 is res[a] = res, lo[a] = min, hi[a] = max.  v: initial velocity. */
typedef struct mpm_cube {
    int32_t res[3];
    float lo[3], hi[3];
    float v[3];
} mpm_cube;

/* The lattices of `ncubes` cubes, one after another; n must be the total point count. */
int mpm_gen_cubes(int dim, int ncubes, const mpm_cube* cubes, int identity_mode, int64_t n, float* image,
                  int device);

/* The reference's random stream: nc_rand_int (src/nclr_math.h:22-29) is xorshift128 on the state
 * {x, y, z, w}, seeded {123456789, 362436069, 521288629, 88675123}; nc_rand = u32 * 2^-32 in float.
 * mpm_rand_advance moves `state` k draws ahead (a GF(2) jump, any k). */
int mpm_rand_advance(uint32_t state[4], uint64_t k);
/* out[k] = the k-th next nc_rand() for k < count; `state` advances by count. */
int mpm_gen_rand(int64_t count, uint32_t state[4], float* out, int device);
/* sample<dim>(center) (src/nclr_math.h:141-147, which has no return statement; here the points
 * are the result) for each of the ncenters host centres (ncenters x dim floats): per_center
 * particles (1000 in the reference) at (randvec<dim>() * 2 - 1) * half_width + center (0.08),
 * drawing from `state`, which advances by ncenters * per_center * dim.  n = ncenters * per_center. */
int mpm_gen_sample(int dim, int ncenters, const float* centers, int per_center, float half_width, uint32_t state[4],
                   int identity_mode, float* image, int device);

/* gyroid(k, t, pos) (src/nclr_math.h:131-139) at n positions (n x 3): out[i], float. */
int mpm_gyroid(int64_t n, const float* pos, float k, float t, float* out, int device);
/* The points of the 3D lattice `box` where gyroid(k, t, p) <= 0 (thickness == 0: the solid
 * network) or |gyroid(k, t, p)| <= thickness (a sheet), in lattice order.  *n_out receives the
 * count; image may be NULL to count only, else it must hold exactly *n_out particles: pass that
 * count in `capacity` (a different value is an error). */
int mpm_gen_gyroid(const mpm_cube* box, float k, float t, float thickness, int identity_mode, float* image,
                   int64_t capacity, int64_t* n_out, int device);

/* pt_3d_to_2d (src/nclr_math.h:154-167) of n positions (n x 3) into out (n x 2).  centered = 0
 * computes what the reference computes (its centred copy of the point is unused); 1 rotates
 * the point about (0.5, 0.5, 0.5). */
int mpm_project_points(int64_t n, const float* pos, int centered, float* out, int device);

/* Releases the device memory, graphs and stream of h. */
int mpm_destroy(mpm_handle h);

/* Message of the last failing call on this thread ("" if none). */
const char* mpm_last_error(void);

/* Library build string (architecture, flags). */
const char* mpm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NCLR_MPM_H */
