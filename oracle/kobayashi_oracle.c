/*
 * kobayashi_oracle.c -- CPU oracle of the Kobayashi step (TEST INFRASTRUCTURE
 * ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline;
 * the product never links it).  See kobayashi_oracle.h for what it restates.
 *
 * Build (oracle/Makefile): gcc -O2 -ffp-contract=off -fPIC -shared, glibc libm.
 */
#ifdef KGO_OPENMP
#include <omp.h>
#endif
#include "kobayashi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int kgo_threads(void)
{
#ifdef KGO_OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void kgo_default_params(kgo_params *p)
{
    memset(p, 0, sizeof(*p));
    /* constructor, src/Kobayashi.cpp:61-63, and src/main.cpp:16 for dt */
    p->dx = (double)0.03f;
    p->dy = (double)0.03f;
    p->dt = (double)0.0001f;
    /* _parameterInit, src/Kobayashi.cpp:76-84 */
    p->tau = (double)0.0003f;
    p->eps_bar = (double)0.010f;
    p->mu = (double)1.0f;
    p->K = (double)1.6f;
    p->delta = (double)0.05f;
    p->aniso = (double)6.0f;
    p->alpha = (double)0.9f;
    p->gamma = (double)10.0f;
    p->teq = (double)1.0f;
    p->theta0 = 0.0;
    p->noise_a = 0.0;
    p->boundary = KGO_PERIODIC;
    p->instance = 0;
    p->seed = 0;
}

/* Philox4x32-10 as published by Salmon et al., "Parallel random numbers: as
 * easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds with the key bumped between
 * rounds.  Pinned by the published known-answer vectors in tests/golden. */
void kgo_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

double kgo_noise(const kgo_params *p, int i, int j, uint32_t step)
{
    const uint32_t ctr[4] = {(uint32_t)i >> 2, (uint32_t)j, step, p->instance};
    const uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    uint32_t out[4];
    kgo_philox4x32_10(ctr, key, out);
    const uint32_t w = out[(uint32_t)i & 3u];
    return (double)(w >> 8) * (1.0 / 16777216.0) - 0.5;
}

#define REAL float
#define SUF _f32
#define ATAN atanf
#define COS cosf
#define SIN sinf
#define PI_R 3.14159274f
#include "kobayashi_oracle_impl.inc"
#undef REAL
#undef SUF
#undef ATAN
#undef COS
#undef SIN
#undef PI_R

#define REAL double
#define SUF _f64
#define ATAN atan
#define COS cos
#define SIN sin
#define PI_R M_PI
#include "kobayashi_oracle_impl.inc"
#undef REAL
#undef SUF
#undef ATAN
#undef COS
#undef SIN
#undef PI_R

static void zero3(size_t n, double *phi, double *T, double *theta)
{
    memset(phi, 0, n * sizeof(double));
    memset(T, 0, n * sizeof(double));
    memset(theta, 0, n * sizeof(double));
}

/* _vectorInit + _createNucleus(nx/2, ny/2): src/Kobayashi.cpp:98-123 */
void kgo_init_plus_f64(int nx, int ny, double *phi, double *T, double *theta)
{
    zero3((size_t)nx * ny, phi, T, theta);
    const int x = nx / 2, y = ny / 2;
    phi[x + (size_t)nx * y] = 1.0;
    phi[(x - 1) + (size_t)nx * y] = 1.0;
    phi[(x + 1) + (size_t)nx * y] = 1.0;
    phi[x + (size_t)nx * (y - 1)] = 1.0;
    phi[x + (size_t)nx * (y + 1)] = 1.0;
}

static void disk(int nx, int ny, double *phi, int cx, int cy, double r)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const double di = (double)(i - cx), dj = (double)(j - cy);
            if (di * di + dj * dj <= r * r)
                phi[i + (size_t)nx * j] = 1.0;
        }
}

void kgo_init_disk_f64(int nx, int ny, double *phi, double *T, double *theta,
                       int cx, int cy, double r)
{
    zero3((size_t)nx * ny, phi, T, theta);
    disk(nx, ny, phi, cx, cy, r);
}

void kgo_init_seeds_f64(int nx, int ny, double *phi, double *T, double *theta,
                        int count, double r, int margin, uint64_t seed, int *cxy)
{
    zero3((size_t)nx * ny, phi, T, theta);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t spanx = (uint32_t)(nx - 2 * margin), spany = (uint32_t)(ny - 2 * margin);
    for (int k = 0; k < count; k++) {
        const uint32_t ctr[4] = {(uint32_t)k, 0u, 0u, 0xFFFFFFFFu};
        uint32_t out[4];
        kgo_philox4x32_10(ctr, key, out);
        const int cx = margin + (int)(out[0] % spanx);
        const int cy = margin + (int)(out[1] % spany);
        if (cxy) {
            cxy[2 * k] = cx;
            cxy[2 * k + 1] = cy;
        }
        disk(nx, ny, phi, cx, cy, r);
    }
}
