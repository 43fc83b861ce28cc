/*
 * kobayashi_gpu.h -- C ABI of the MI355X (gfx950) Kobayashi phase-field solver.
 *
 * Drop-in for the numerical core of jklae/CrystalGrowth's `Kobayashi` class
 * (src/Kobayashi.h:32-122, src/Kobayashi.cpp:7-249): plain C types, opaque
 * handles, int status codes.  Every function returns KG_OK (0) or a negative
 * KG_ERR_* code; kg_last_error() then describes the failure (per thread).
 *
 * Fields are nx*ny arrays, x fastest (index i + nx*j, src/Kobayashi.h:91), in
 * the handle's precision (KG_F32: float, KG_F64: double).  The persistent
 * state is (phase phi, temperature T, angle theta): theta is only rewritten
 * where the phase gradient is non-flat (src/Kobayashi.cpp:154-167), so it
 * carries state across steps.
 *
 * Threading: one HIP stream per handle; calls on one handle must not overlap;
 * distinct handles are independent.  Host buffers are caller-owned.
 */
#ifndef KOBAYASHI_GPU_H
#define KOBAYASHI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KG_ABI_VERSION 1

/* status codes */
#define KG_OK 0
#define KG_ERR_INVALID_ARGUMENT (-1)
#define KG_ERR_UNSTABLE (-2)      /* explicit-Euler stability bound violated */
#define KG_ERR_OUT_OF_MEMORY (-3)
#define KG_ERR_DEVICE (-4)        /* HIP runtime error or no usable GPU */
#define KG_ERR_SIZE (-5)          /* host buffer size does not match the field */

/* precision of a handle's fields and arithmetic */
#define KG_F32 0
#define KG_F64 1

/* boundary maps of every neighbour index (both passes) */
#define KG_PERIODIC 0             /* reference: i +- 1 wrapped (src/Kobayashi.cpp:133-136) */
#define KG_ZERO_FLUX 1            /* i +- 1 clamped to the grid */

/* step kernels */
#define KG_KERNEL_FUSED 0         /* one LDS-tiled launch per step (default) */
#define KG_KERNEL_TWO_PASS 1      /* pass 1 + pass 2 through global temporaries */
#define KG_KERNEL_FUSED_DENSE 2   /* the fused launch without tile status: every tile
                                     loads and stores all of its fields (same results) */

typedef struct kg_params {
    double dx, dy, dt;            /* ctor: _dx=_dy=0.03, _dt=timeStep (src/Kobayashi.cpp:61-63) */
    double tau, eps_bar, mu;      /* _parameterInit (src/Kobayashi.cpp:76-78); mu is unused, as in the reference */
    double K, delta, aniso;       /* latent heat, anisotropy strength, mode number j (:79-81) */
    double alpha, gamma, teq;     /* m(T) = alpha/pi atan(gamma (Teq - T)) (:82-84, :206) */
    double theta0;                /* anisotropy offset; 0 reproduces the reference */
    double noise_a;               /* noise amplitude a; 0 reproduces the reference */
    int32_t boundary;             /* KG_PERIODIC (reference) or KG_ZERO_FLUX */
    uint32_t instance;            /* ensemble member id, part of the noise counter */
    uint64_t seed;                /* Philox4x32-10 key of the noise */
} kg_params;

typedef struct kg_stats {
    double sum_phi;               /* FP64 sums over all cells */
    double sum_T;
    double sum_T_minus_K_phi;     /* invariant of the scheme up to rounding */
    double min_phi, max_phi;
    double solid_fraction;        /* fraction of cells with phi > 1/2 */
    int64_t nonfinite;            /* cells whose phi or T is NaN or infinite */
    int64_t cells;
} kg_stats;

typedef struct kg_solver_s *kg_handle;

/* ---- library ------------------------------------------------------------ */
int kg_abi_version(void);
const char *kg_last_error(void);
/* Number of visible GPUs (0 without a GPU; never an error). */
int kg_device_count(int *count);

/* ---- parameters (host only; no GPU needed) ------------------------------ */
/* Reference defaults: constructor + _parameterInit (src/Kobayashi.cpp:61-63,
 * :73-96) with dt = 1e-4 as src/main.cpp:16 passes it. */
int kg_default_params(kg_params *out);
/* Checks sizes (nx, ny >= 3), finite positive dx, dy, dt, tau, eps_bar, and
 * the explicit-Euler bound dt/dx^2 <= 3/8 of the 9-point Laplacian (also for
 * the phase equation's diffusion eps_bar^2 (1+|delta|)^2 / tau). */
int kg_validate(int nx, int ny, const kg_params *p);
/* Bit k of *mask set = parameter k outside the reference GUI's slider range
 * (src/Kobayashi.cpp:17-58; order tau, eps_bar, mu, K, delta, aniso, alpha,
 * gamma, teq).  Informational: out-of-range values are still accepted. */
int kg_reference_range_mask(const kg_params *p, uint32_t *mask);
/* Philox4x32-10 block (host evaluation of the device noise source). */
int kg_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- lifecycle ------------------------------------------------------------ */
/* Replaces Kobayashi::Kobayashi(int x, int y, float timeStep)
 * (src/Kobayashi.h:35, src/Kobayashi.cpp:7-67): allocates device fields and
 * applies the reference initial state (kg_reset). */
int kg_create(int nx, int ny, int precision, const kg_params *p, kg_handle *out);
/* Replaces Kobayashi::~Kobayashi (src/Kobayashi.cpp:69-71). */
int kg_destroy(kg_handle h);

/* ---- initial states --------------------------------------------------------- */
/* Replaces _vectorInit + _createNucleus(nx/2, ny/2) (src/Kobayashi.cpp:98-123),
 * as iResetSimulationState calls it (:241-249): all fields 0, phi = 1 on the
 * 5-cell plus at (nx/2, ny/2); the step counter returns to 0. */
int kg_reset(kg_handle h);
/* All fields 0, phi = 1 where (i-cx)^2 + (j-cy)^2 <= r^2; counter to 0. */
int kg_reset_disk(kg_handle h, int cx, int cy, double r);
/* All fields 0, `count` disks of radius r at Philox-drawn centres at least
 * `margin` cells from every edge (centres in cxy[2*count] if non-NULL). */
int kg_reset_seeds(kg_handle h, int count, double r, int margin, uint64_t seed, int *cxy);

/* ---- parameters of a handle ------------------------------------------------- */
/* Replaces the slider writes (src/Kobayashi.cpp:587-611) and RESET's
 * _parameterInit (:514).  Unlike the GUI it does not reset the fields. */
int kg_set_params(kg_handle h, const kg_params *p);
int kg_get_params(kg_handle h, kg_params *p);
int kg_get_size(kg_handle h, int *nx, int *ny, int *precision);

/* ---- stepping ------------------------------------------------------------------ */
/* nsteps explicit-Euler steps: _computeGradientLaplacian + _evolution
 * (src/Kobayashi.cpp:125-221) per step; Kobayashi::iUpdate is kg_step(h, 10)
 * (src/Kobayashi.cpp:227-239).  Asynchronous on the handle's stream; with
 * the default kernel the host stays at most about 14 steps ahead of the GPU
 * (every 7 steps it waits for the GPU's answer on which regions stay all
 * +0).  KG_ERR_INVALID_ARGUMENT while the handle's stream is being captured
 * into a graph (a replay would repeat the captured step; batches in
 * KG_BATCH_GRAPH mode replay correctly). */
int kg_step(kg_handle h, int nsteps);
/* kg_step bracketed by HIP events on the handle's stream; *ms = elapsed
 * milliseconds of the nsteps launches (waits for completion). */
int kg_step_timed(kg_handle h, int nsteps, double *ms);
int kg_synchronize(kg_handle h);
/* Steps taken since the last reset (the noise counter of the next step). */
int kg_get_step_count(kg_handle h, uint64_t *steps);
/* Sets that counter (the noise counter); the next steps cover every chunk
 * until a fresh answer on which regions stay all +0 arrives. */
int kg_set_step_count(kg_handle h, uint64_t steps);
/* Diagnostic, no counterpart in the reference: step launches since kg_create
 * that covered only the chunks listed as live (the rest all +0), and liveness
 * answers dropped because a flag was not yet visible to the host.  Chunk
 * lists are used on grids of at least 256 chunks of 128 x 120 cells. */
int kg_get_skip_stats(kg_handle h, uint64_t *listed_launches, uint64_t *stale_answers);
int kg_set_kernel(kg_handle h, int kernel);

/* ---- fields --------------------------------------------------------------------- */
/* Copies of _phi, _t, _angl (src/Kobayashi.h:107-115); bytes must equal
 * nx*ny*sizeof(element).  Synchronous. */
int kg_get_phase(kg_handle h, void *dst, size_t bytes);
int kg_get_temperature(kg_handle h, void *dst, size_t bytes);
int kg_get_angle(kg_handle h, void *dst, size_t bytes);
/* Replaces the state; theta may be NULL (then 0).  Does not touch the counter. */
int kg_set_fields(kg_handle h, const void *phi, const void *T, const void *theta, size_t bytes);
/* Device addresses of the current fields (valid until the next kg_step or
 * kg_reset*) and the handle's stream (hipStream_t). */
int kg_device_fields(kg_handle h, void **phi, void **T, void **theta);
int kg_get_stream(kg_handle h, void **stream);

/* ---- reductions and readback --------------------------------------------------------- */
/* FP64 reductions over the grid, deterministic (fixed two-level order). */
int kg_reduce(kg_handle h, kg_stats *out);
/* phi -> RGBA colour of iUpdateConstantBuffer's three-segment ramp
 * (src/Kobayashi.cpp:309-345), 4 floats per cell with alpha 1.  order 0: x
 * fastest; order 1: the reference's constant-buffer order, entry i showing
 * _phi[_INDEX(i / n, i % n)] (square grids only). */
int kg_colormap(kg_handle h, float *rgba, size_t bytes, int order);
/* The same colours as 8-bit unsigned normalised RGBA, 4 bytes per cell (R, G,
 * B, A = 255; each channel 255 c rounded to the nearest integer, ties to
 * even, c clamped to [0, 1]): the float-to-UNORM8 conversion a viewer applies
 * to the constant buffer's float colour (src/Kobayashi.cpp:344), done on the
 * device so a frame crosses PCIe at a quarter of the float size. */
int kg_colormap_rgba8(kg_handle h, uint8_t *rgba, size_t bytes, int order);
/* kg_colormap_rgba8 without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (render loops
 * that must not call hipMalloc per frame, code under stream capture,
 * harnesses that forbid allocation).  The colours are staged in a buffer of
 * the library's own, KG_COLORMAP_NOALLOC_MAX_CELLS * 4 bytes (16 MiB), that
 * every GPU it runs on holds until the process ends; larger grids return
 * KG_ERR_SIZE.  All calls on one GPU share that buffer: calls that overlap on
 * one GPU (from several threads, or on handles whose streams run
 * concurrently) race and must be serialised by the caller.  Same results as
 * kg_colormap_rgba8; synchronous like it. */
#define KG_COLORMAP_NOALLOC_MAX_CELLS (1 << 22)
int kg_colormap_rgba8_noalloc(kg_handle h, uint8_t *rgba, size_t bytes, int order);
/* Morphology of the phi = level front (no counterpart in the reference, which
 * only draws the field; the long-horizon comparisons use it): radii[k] = the
 * distance in cells from (cx, cy) to the first point where phi < level along
 * the ray at angle 2 pi k / n, marched in steps of 1/4 cell with bilinear
 * sampling (outside the grid reads 0), linearly interpolated; FP64.  The tip
 * radius is the largest entry; arm counts come from the profile's peaks. */
int kg_front_profile(kg_handle h, double cx, double cy, int n, double level, double *radii);

/* ---- adjoint (reverse mode) ------------------------------------------------------------
 * No counterpart in the reference (it has no gradients); needed to fit model
 * parameters to a target morphology.  From the handle's current state and
 * step counter, runs nsteps forward (storing (phi, T, theta) every
 * checkpoint_every steps), evaluates L = sum (phi_N - target)^2 in FP64, and
 * runs the hand-written reverse of each step (segments recomputed from their
 * checkpoint) to accumulate dL/d(parameter) in FP64.  d_T0 and d_phi0 (host,
 * nx*ny elements of the handle's precision, optional) receive dL/dT and dL/dphi
 * of the initial state.  Afterwards the handle holds the final forward state.
 * theta branches (src/Kobayashi.cpp:154-167) are piecewise: where a branch
 * fires theta = const + atan(gy/gx) is differentiated, where none fires the
 * kept theta passes its adjoint through. */
typedef struct kg_grads {
    double loss;
    double d_tau, d_eps_bar, d_K, d_delta, d_theta0, d_alpha, d_gamma, d_teq, d_noise_a;
} kg_grads;

int kg_adjoint(kg_handle h, int nsteps, int checkpoint_every, const void *target, size_t bytes,
               kg_grads *out, void *d_T0, void *d_phi0);

/* ---- batched ensemble ------------------------------------------------------------------
 * `count` independent members of any sizes and parameters (each its own
 * kg_params, including seed and instance), stepped together: one launch per
 * step covers every tile of every member, or a hipGraph replays a captured
 * run of launches.  Member m of a batch evolves bit for bit like a plain
 * handle with the same size, precision, parameters and initial state.
 * No counterpart in the reference (one grid per process there); the
 * per-member semantics are those of kg_step. */
#define KG_BATCH_LAUNCH 0         /* one kernel launch per step */
#define KG_BATCH_GRAPH 1          /* a captured hipGraph of graph_steps launches per replay */
#define KG_BATCH_GRAPH_LISTS 2    /* the same, with the chunk lists kept on the device */

typedef struct kg_batch_s *kg_batch;

int kg_batch_create(int count, const int *nx, const int *ny, int precision,
                    const kg_params *params, kg_batch *out);
int kg_batch_destroy(kg_batch b);
int kg_batch_count(kg_batch b, int *count);
/* Parameters of one member (takes effect at the next step; fields unchanged). */
int kg_batch_set_params(kg_batch b, int member, const kg_params *p);
int kg_batch_get_params(kg_batch b, int member, kg_params *p);
/* member < 0 applies to every member (disk centred at (nx/2, ny/2) when cx < 0) */
int kg_batch_reset_disk(kg_batch b, int member, int cx, int cy, double r);
int kg_batch_reset_seeds(kg_batch b, int member, int count, double r, int margin, uint64_t seed);
int kg_batch_set_fields(kg_batch b, int member, const void *phi, const void *T,
                        const void *theta, size_t bytes);
int kg_batch_get_phase(kg_batch b, int member, void *dst, size_t bytes);
int kg_batch_get_temperature(kg_batch b, int member, void *dst, size_t bytes);
int kg_batch_get_angle(kg_batch b, int member, void *dst, size_t bytes);
int kg_batch_reduce(kg_batch b, int member, kg_stats *out);
/* kg_front_profile of one member */
int kg_batch_front_profile(kg_batch b, int member, double cx, double cy, int n, double level,
                           double *radii);
/* mode KG_BATCH_LAUNCH, KG_BATCH_GRAPH or KG_BATCH_GRAPH_LISTS (graph_steps
 * even, >= 2).  KG_BATCH_LAUNCH skips chunks by lists the host makes
 * (waiting every 7 steps for the GPU's answer); KG_BATCH_GRAPH steps every
 * tile and never waits; KG_BATCH_GRAPH_LISTS never waits and skips chunks by
 * lists made on the device, at the price of a fixed grid whose unused
 * blocks exit and two small kernels per step.  Chunk lists need at least
 * 256 chunks over the members; smaller batches step every tile. */
int kg_batch_set_mode(kg_batch b, int mode, int graph_steps);
int kg_batch_get_mode(kg_batch b, int *mode, int *graph_steps);
/* Chunk shape (in 64 x 30 tiles) into which every member is cut: the step
 * launch deals the chunks round-robin to the XCDs, and in KG_BATCH_LAUNCH
 * mode the chunk lists skip whole chunks.  One of 2x2 (the default), 2x4,
 * 4x2, 4x1, 8x1; results are bitwise the same in every shape. */
int kg_batch_set_chunk_shape(kg_batch b, int chunk_w, int chunk_h);
int kg_batch_get_chunk_shape(kg_batch b, int *chunk_w, int *chunk_h);
/* Times trial_steps steps of each launch configuration (launch form and
 * chunk shape) on a copy of the current state and keeps the fastest (the
 * state is restored afterwards; kg_batch_get_chunk_shape reports the shape
 * it picked).  The chunk lists take effect within two liveness periods
 * (14 steps) of the trial's start, so trials shorter than that time the
 * launch of every chunk. */
int kg_batch_autotune(kg_batch b, int trial_steps, int *mode, int *graph_steps);
int kg_batch_step(kg_batch b, int nsteps);
int kg_batch_step_timed(kg_batch b, int nsteps, double *ms);
int kg_batch_synchronize(kg_batch b);
int kg_batch_get_step_count(kg_batch b, uint64_t *steps);
/* Diagnostic, no counterpart in the reference: as kg_get_skip_stats, for a
 * batch with at least 256 chunks over its members.  A list made on the
 * device (KG_BATCH_GRAPH_LISTS) counts as the 7 launches it serves. */
int kg_batch_get_skip_stats(kg_batch b, uint64_t *listed_launches, uint64_t *stale_answers);
/* kg_adjoint for every member at once: targets[m] and d_T0[m] (optional;
 * the array or an entry may be NULL) are host fields of member m; out[m]
 * receives member m's loss and gradients.  One launch per kernel covers all
 * members; afterwards the batch holds the final forward state. */
int kg_batch_adjoint(kg_batch b, int nsteps, int checkpoint_every, const void *const *targets,
                     kg_grads *out, void *const *d_T0);

#ifdef __cplusplus
}
#endif

#endif /* KOBAYASHI_GPU_H */
