// kg_math.h -- transcendental functions of the step kernels.
//
// FP32: polynomial atan, sin and cos written for this solver.  Coefficients
// are near-minimax fits made by csrc/fit_minimax.py (Lawson's iteratively
// reweighted least squares, float64, rounded to float); with float
// evaluation the maximum relative errors are 1.6e-7 (atan) and 6.5e-8 (sin),
// and the absolute error of cos is 6.5e-8.
// FP64: the device library's double functions (accurate to ~1 ulp), so the
// FP64 path stays a faithful oracle-grade computation.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace kg {

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// atan(q) for q in [0, 1]: q * A(q^2)
__device__ __forceinline__ float atan_unit(float q)
{
    const float q2 = q * q;
    float p = 2.849889686e-03f;
    p = fmaf(p, q2, -1.606862992e-02f);
    p = fmaf(p, q2, 4.269151762e-02f);
    p = fmaf(p, q2, -7.504294813e-02f);
    p = fmaf(p, q2, 1.064093411e-01f);
    p = fmaf(p, q2, -1.420364380e-01f);
    p = fmaf(p, q2, 1.999261975e-01f);
    p = fmaf(p, q2, -3.333307207e-01f);
    p = fmaf(p, q2, 1.0f);
    return q * p;
}

// atan(x), x any float (NaN propagates, +-inf -> +-pi/2)
__device__ __forceinline__ float kg_atan(float x)
{
    const float ax = fabsf(x);
    const bool big = ax > 1.0f;
    float p = atan_unit(big ? fast_rcp(ax) : ax);
    p = big ? 1.57079637f - p : p;
    return copysignf(p, x);
}

// atan(y / x) in (-pi/2, pi/2) without forming the quotient (one reciprocal)
__device__ __forceinline__ float kg_atan_ratio(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    float p = atan_unit(mn * fast_rcp(mx));
    p = ay > ax ? 1.57079637f - p : p;
    return ((x < 0.0f) != (y < 0.0f)) ? -p : p;
}

// Largest |x| kg_sincos_reduced serves (its three-part reduction stays exact
// well beyond it); callers that cannot bound |x| use kg_sincos.
#define KG_SINCOS_REDUCED_MAX 1.0e4f

// sin and cos of x for |x| <= KG_SINCOS_REDUCED_MAX: Cody-Waite reduction by
// pi/2 (three-part constant, one fma per part), then polynomials on
// [-pi/4, pi/4].
__device__ __forceinline__ void kg_sincos_reduced(float x, float *s, float *c)
{
    const float k = rintf(x * 0.636619772f);
    float r = fmaf(-k, 1.570796371e+00f, x);
    r = fmaf(-k, -4.371138829e-08f, r);
    r = fmaf(-k, -1.776356839e-15f, r);
    const float r2 = r * r;
    float sp = 2.718122005e-06f;
    sp = fmaf(sp, r2, -1.983931288e-04f);
    sp = fmaf(sp, r2, 8.333329111e-03f);
    sp = fmaf(sp, r2, -1.666666716e-01f);
    sp = fmaf(sp * r2, r, r);
    float cp = -2.718010137e-07f;
    cp = fmaf(cp, r2, 2.479904106e-05f);
    cp = fmaf(cp, r2, -1.388888108e-03f);
    cp = fmaf(cp, r2, 4.166666791e-02f);
    cp = fmaf(cp, r2, -5.0e-01f);
    cp = fmaf(cp, r2, 1.0f);
    const int q = (int)k & 3;
    const float sv = (q & 1) ? cp : sp;
    const float cv = (q & 1) ? sp : cp;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

// sin and cos of any float x: kg_sincos_reduced, or the library call beyond
// its range (never reached by j (theta - theta0) with theta from the step)
__device__ __forceinline__ void kg_sincos(float x, float *s, float *c)
{
    if (fabsf(x) > KG_SINCOS_REDUCED_MAX) {
        sincosf(x, s, c);
        return;
    }
    kg_sincos_reduced(x, s, c);
}

__device__ __forceinline__ double kg_atan(double x) { return atan(x); }
__device__ __forceinline__ double kg_atan_ratio(double y, double x) { return atan(y / x); }
__device__ __forceinline__ void kg_sincos(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void kg_sincos_reduced(double x, double *s, double *c)
{
    sincos(x, s, c);
}

}  // namespace kg
