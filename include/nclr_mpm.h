/*
 * nclr_mpm.h - C ABI of the MI355X (gfx950) MLS-MPM engine.
 *
 * The reference engine (jparr721/NuclearMPM) is a header-only C++ class with
 * no FFI; these entry points are what a binding of its step loop needs.  Each
 * declaration names the reference interface it replaces.  The C++ drop-in
 * class in include/nclr/nclr.h and the Python package nuclearmpm_amd are both
 * written on top of this ABI.
 *
 * Conventions
 *  - Host arrays are particle-major: x[n][dim], v[n][dim], F[n][dim][dim]
 *    (row-major), C[n][dim][dim], Jp[n], mass[n], volume[n].  Device state is
 *    structure-of-arrays inside the library; host buffers are copied in/out.
 *  - Grid arrays use the reference's flat node order, x-major with the last
 *    axis fastest, (res+1)^dim nodes: velocity[node][dim], mass[node].
 *  - Every function returns MPM_OK (0) or an MPM_ERR_* code; mpm_last_error()
 *    gives the message of the last failure on the calling thread.
 *  - One handle owns one HIP stream and must be driven from one host thread
 *    at a time; different handles are independent.
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
     *    3D "identity" of the F update and the liquid reset is diag(1,1,0) (src/nclr.h:230,256).
     * 1: physical identity. */
    int32_t identity_mode;
    /* Stable counting sort of the particles by grid block every `sort_every` steps (0 = never). */
    int32_t sort_every;
    /* P2G variant: 0 auto, 1 per-particle global atomics, 2 LDS block tiles flushed with atomics. */
    int32_t p2g_variant;
    /* Threads per workgroup of the particle kernels (0 = library default). */
    int32_t block_threads;
    /* HIP device ordinal. */
    int32_t device;
    /* Record per-kernel HIP events (mpm_kernel_times). */
    int32_t timing;
    int32_t reserved[10];
} mpm_options;

/* Fills *opts with the defaults (reference identity mode, library-chosen launch shapes). */
void mpm_default_options(mpm_options* opts);

/* Replaces MPMSimulation<dim>::MPMSimulation(particles, model, res, dt, E, nu, gravity)
 * (src/nclr.h:74-78).  Copies the particle arrays to the device.  mass/volume may be NULL
 * (then 1, the Particle default of src/nclr.h:46).  opts may be NULL. */
int mpm_create(int dim, int material, int res, float dt, float E, float nu, float gravity, int64_t n,
               const float* x, const float* v, const float* F, const float* C, const float* Jp,
               const float* mass, const float* volume, const mpm_options* opts, mpm_handle* out);

/* Replaces nsteps calls of MPMSimulation<dim>::advance() (src/nclr.h:80-84).  The steps are
 * queued on the handle's stream; the call returns without waiting.  Errors raised by queued
 * steps (a particle leaving the grid) are reported by the next synchronising call. */
int mpm_advance(mpm_handle h, int nsteps);

/* Waits for queued work; returns MPM_ERR_OUT_OF_RANGE if any step since the last check found a
 * particle outside the grid (the reference throws std::out_of_range from std::vector::at). */
int mpm_synchronize(mpm_handle h);

/* Replaces MPMSimulation<dim>::particles() (src/nclr.h:86): copies the current particle state
 * to host arrays in the original particle order.  Any pointer may be NULL to skip that field. */
int mpm_get_particles(mpm_handle h, float* x, float* v, float* F, float* C, float* Jp, float* mass,
                      float* volume);

/* Replaces the particle vector the constructor takes by value (src/nclr.h:74,76): overwrites the
 * state of all n particles (NULL pointers leave a field unchanged). */
int mpm_set_particles(mpm_handle h, const float* x, const float* v, const float* F, const float* C,
                      const float* Jp, const float* mass, const float* volume);

/* Replaces MPMSimulation<dim>::grid() (src/nclr.h:87): the grid after the last step's grid
 * update, velocity (or momentum where mass == 0) and mass per node, dense flat layout.
 * Before the first step the grid is empty and both arrays are zero-filled. */
int mpm_get_grid(mpm_handle h, float* velocity, float* mass);

/* Scalars of the simulation: mu_0, lambda_0 (src/nclr.h:71-72), dx, inv_dx. */
int mpm_get_params(mpm_handle h, float* mu_0, float* lambda_0, float* dx, float* inv_dx);
int64_t mpm_num_particles(mpm_handle h);
int64_t mpm_num_nodes(mpm_handle h);
int64_t mpm_steps_taken(mpm_handle h);

/* Kernel timing (opts.timing = 1): average device milliseconds per launch of each kernel since the
 * last reset, measured with HIP events on the handle's stream, and the launch counts.
 * names: "sort", "p2g", "grid", "g2p".  Returns MPM_ERR_INVALID for an unknown name. */
int mpm_kernel_time(mpm_handle h, const char* name, double* avg_ms, int64_t* launches);
int mpm_reset_kernel_times(mpm_handle h);
/* Turns the per-kernel HIP event recording of mpm_kernel_time on (1) or off (0). */
int mpm_set_timing(mpm_handle h, int enable);

/* Number of grid nodes with mass > 0 after the last step's P2G (the G_active count of the
 * roofline bytes). */
int mpm_active_nodes(mpm_handle h, int64_t* count);

/* The engine's stable counting sort (the per-step particle-by-tile sort) on n host keys:
 * perm[j] = index of the j-th key in stable order by the low `bits` bits (no reference
 * counterpart; the reference steps particles in index order, src/nclr.h:112,169). */
int mpm_sort_keys(const uint32_t* keys, int64_t n, int bits, uint32_t* perm, int device);

/* Releases the device memory and stream of h. */
int mpm_destroy(mpm_handle h);

/* Message of the last failing call on this thread ("" if none). */
const char* mpm_last_error(void);

/* Library build string (architecture, flags). */
const char* mpm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NCLR_MPM_H */
