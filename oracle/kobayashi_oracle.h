/*
 * kobayashi_oracle.h -- CPU oracle for the Kobayashi (1993) phase-field step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's numerical core, used as the checker for the HIP implementation in
 * crystalgrowth_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * What it restates (reference = jklae/CrystalGrowth):
 *   kgo_step_*   src/Kobayashi.cpp:125-175 (_computeGradientLaplacian) followed by
 *                src/Kobayashi.cpp:177-221 (_evolution), in the same operation
 *                order, in float (the reference's type) or double (the FP64 oracle).
 *   kgo_init_plus  src/Kobayashi.cpp:98-123 (_vectorInit + _createNucleus).
 *   kgo_default_params  src/Kobayashi.cpp:61-63 and :76-84 (ctor + _parameterInit).
 * Extensions that the reference lacks, each reducing exactly to it at its
 * neutral setting: anisotropy offset theta0 (0), Philox noise amplitude a (0),
 * zero-flux boundaries (periodic), disk / seeded multi-nucleus initial states.
 *
 * How each function is pinned is listed in tests/test_oracle_pins.py; the
 * reference itself cannot be compiled here (its header includes the
 * un-vendored DXViewer's Win32App.h), so no bit patterns of the reference's
 * own binary exist to compare with.
 */
#ifndef KOBAYASHI_ORACLE_H
#define KOBAYASHI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { KGO_PERIODIC = 0, KGO_ZERO_FLUX = 1 };

typedef struct kgo_params {
    double dx, dy, dt;
    double tau, eps_bar, mu, K, delta, aniso, alpha, gamma, teq;
    double theta0;      /* anisotropy offset (0 = reference) */
    double noise_a;     /* noise amplitude a (0 = reference) */
    int32_t boundary;   /* KGO_PERIODIC (reference) or KGO_ZERO_FLUX */
    uint32_t instance;  /* ensemble member id, part of the noise counter */
    uint64_t seed;      /* Philox key */
} kgo_params;

/* Reference defaults: dx=dy=0.03, dt=1e-4 (src/main.cpp:16), tau=3e-4,
 * eps_bar=0.01, mu=1, K=1.6, delta=0.05, j=6, alpha=0.9, gamma=10, Teq=1. */
void kgo_default_params(kgo_params *p);

/* Threads the step loops use: 1, or the OpenMP team size of the build with
 * -DKGO_OPENMP (libkg_oracle_omp.so, bitwise the same results). */
int kgo_threads(void);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), one block. */
void kgo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Noise sample chi in [-1/2, 1/2) for cell (i, j) at step `step`:
 * counter = (i >> 2, j, step, instance), key = (seed lo, seed hi),
 * word = i & 3, chi = (word >> 8) * 2^-24 - 1/2 (exact in float). */
double kgo_noise(const kgo_params *p, int i, int j, uint32_t step);

/* One explicit-Euler step.  phi, T and theta are nx*ny arrays, x fastest
 * (src/Kobayashi.h:91).  theta is the persistent angle field (read-modify-
 * write, src/Kobayashi.cpp:154-167).  `step` is the noise counter.
 * `work` is scratch of 7*nx*ny elements (or NULL to allocate internally). */
void kgo_step_f32(int nx, int ny, const kgo_params *p, float *phi, float *T,
                  float *theta, uint32_t step, float *work);
void kgo_step_f64(int nx, int ny, const kgo_params *p, double *phi, double *T,
                  double *theta, uint32_t step, double *work);

/* nsteps steps with noise counters step0, step0+1, ... */
void kgo_run_f32(int nx, int ny, const kgo_params *p, float *phi, float *T,
                 float *theta, uint32_t step0, int nsteps);
void kgo_run_f64(int nx, int ny, const kgo_params *p, double *phi, double *T,
                 double *theta, uint32_t step0, int nsteps);

/* Pass-1 quantities of one step without modifying the state, for tests:
 * out7 = [gx, gy, lap_phi, lap_T, theta_new, eps, eps_deriv], each nx*ny. */
void kgo_pass1_f64(int nx, int ny, const kgo_params *p, const double *phi,
                   const double *T, const double *theta, double *out7);
void kgo_pass1_f32(int nx, int ny, const kgo_params *p, const float *phi,
                   const float *T, const float *theta, float *out7);

/* Initial states (all fields zeroed first, as _vectorInit does). */
void kgo_init_plus_f64(int nx, int ny, double *phi, double *T, double *theta);
void kgo_init_disk_f64(int nx, int ny, double *phi, double *T, double *theta,
                       int cx, int cy, double r);
/* count disks of radius r at Philox-drawn centres at least `margin` cells
 * from every edge; centres are returned in cxy (2*count ints) if non-NULL. */
void kgo_init_seeds_f64(int nx, int ny, double *phi, double *T, double *theta,
                        int count, double r, int margin, uint64_t seed, int *cxy);

#ifdef __cplusplus
}
#endif

#endif /* KOBAYASHI_ORACLE_H */
