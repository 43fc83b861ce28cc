// kg_step.hip -- gfx950 kernels of the Kobayashi phase-field step.
//
// One explicit-Euler step = the reference's pass 1 (_computeGradientLaplacian,
// src/Kobayashi.cpp:125-175) and pass 2 (_evolution, :177-221).  The fused
// kernel does both in one launch: a 64x16 output tile per 256-thread block,
// phi staged in LDS with a 2-cell halo, T with a 1-cell halo, the pass-1
// quantities that pass 2 reads from neighbours (eps^2, eps eps' gx,
// eps eps' gy) computed once per cell of the tile plus a 1-cell halo and kept
// in LDS.  phi and T ping-pong between two buffers; theta is one buffer,
// written only where a branch of :154-167 fires and read only where none
// does, so blocks never race on it.
//
// Memory-bound: per cell-update the algorithmic traffic is one read and one
// write of phi and T (16 B in FP32).  Tiles whose phi is exactly 0 or exactly
// 1 everywhere (plus halo) reduce to the T diffusion update, bit for bit the
// same result as the full arithmetic (every cross term is a signed zero).
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include "kg_internal.h"
#include "kg_philox.h"

namespace kg {

namespace {

constexpr int TX = 64;   // output tile width (x, contiguous)
constexpr int TY = 16;   // output tile height
constexpr int NT = 256;  // threads per block (4 wave64)
constexpr int NXCD = 8;  // MI355X: workgroups go round-robin over 8 XCDs

template <typename R> struct Consts;
template <> struct Consts<float> {
    // PI_F of the reference's viewer library, taken as float(pi)
    static __device__ __forceinline__ float pi() { return 3.14159274f; }
};
template <> struct Consts<double> {
    static __device__ __forceinline__ double pi() { return 3.14159265358979323846; }
};

__device__ __forceinline__ float kg_atan(float x) { return atanf(x); }
__device__ __forceinline__ double kg_atan(double x) { return atan(x); }
__device__ __forceinline__ void kg_sincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void kg_sincos(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ float kg_div(float a, float b) { return __fdividef(a, b); }
__device__ __forceinline__ double kg_div(double a, double b) { return a / b; }
__device__ __forceinline__ bool is_pos_zero(float v) { return __float_as_uint(v) == 0u; }
__device__ __forceinline__ bool is_pos_zero(double v) { return __double_as_longlong(v) == 0ll; }

template <int BC>
__device__ __forceinline__ int map_index(int x, int n)
{
    if (BC == 1)
        return x < 0 ? 0 : (x >= n ? n - 1 : x);
    while (x < 0)
        x += n;
    while (x >= n)
        x -= n;
    return x;
}

// The angle update of src/Kobayashi.cpp:154-167.  Returns true when a branch
// fires (theta := th); false leaves the stored theta in force.  NaN
// gradients fire nothing, as in the reference.
template <typename R>
__device__ __forceinline__ bool theta_branch(R gx, R gy, R &th)
{
    const R E = (R)1.1920928955078125e-7;  // FLT_EPSILON in both precisions
    const R PI = Consts<R>::pi();
    const R t = kg_atan(kg_div(gy, gx));
    if (gx > E) {
        if (gy < -E) {
            th = (R)2.0 * PI + t;
            return true;
        }
        if (gy > E) {
            th = t;
            return true;
        }
        return false;
    }
    if (gx < -E) {
        th = PI + t;
        return true;
    }
    if (gx <= E && gx >= -E) {
        if (gy < -E) {
            th = (R)-0.5 * PI;
            return true;
        }
        if (gy > E) {
            th = (R)0.5 * PI;
            return true;
        }
    }
    return false;
}

template <typename R>
__device__ __forceinline__ R cell_noise(int i, int j, const StepArgs<R> &a)
{
    const u32x4 v = philox4x32_10(u32x4{(uint32_t)i >> 2, (uint32_t)j, a.step, a.instance},
                                  a.key0, a.key1);
    return noise_from_word<R>(select_word(v, (uint32_t)i & 3u));
}

// (2 (E + W + N + S) + NE + SW + NW + SE - 12 C), the order of :142-151.
template <typename R>
__device__ __forceinline__ R lap9_sum(const R *s, int c, int w)
{
    return (R)2.0 * (s[c + 1] + s[c - 1] + s[c + w] + s[c - w]) + s[c + w + 1] + s[c - w - 1] +
           s[c + w - 1] + s[c - w + 1] - (R)12.0 * s[c];
}

// Reaction + combination of :206-215 for one cell.
template <typename R>
__device__ __forceinline__ void update_cell(const StepArgs<R> &a, int i, int j, R phi, R T,
                                            R lap, R lapT, R e2c, R t1, R t2, R t3, R &phin,
                                            R &Tn)
{
    const R m = a.alpha_pi * kg_atan(a.gamma * (a.teq - T));
    R q = phi - (R)0.5 + m;
    if (a.noise_a != (R)0)
        q = q + a.noise_a * cell_noise<R>(i, j, a);
    const R sum = t1 + t2 + e2c * lap + t3 + phi * ((R)1.0 - phi) * q;
    phin = phi + sum * a.dt_tau;
    Tn = T + lapT * a.dt + a.K * (phin - phi);
}

__device__ __forceinline__ int xcd_tile(int ntiles, int per_xcd)
{
    // Block b runs on XCD b % 8; give each XCD a contiguous run of tiles so
    // that halo rows shared by neighbouring tiles stay in one L2.
    const int b = blockIdx.x;
    return (b % NXCD) * per_xcd + b / NXCD;
}

template <typename R, int BC>
__global__ __launch_bounds__(NT) void k_step_fused(StepArgs<R> a)
{
    constexpr int PW = TX + 4, PH = TY + 4;  // phi tile, halo 2
    constexpr int QW = TX + 2, QH = TY + 2;  // T and pass-1 tiles, halo 1
    __shared__ R sP[PH * PW];
    __shared__ R sT[QH * QW];
    __shared__ R sE2[QH * QW];
    __shared__ R sA[QH * QW];
    __shared__ R sB[QH * QW];

    const int tile = xcd_tile(a.ntiles, a.per_xcd);
    if (tile >= a.ntiles)
        return;
    const int nx = a.nx, ny = a.ny;
    const int nbx = (nx + TX - 1) / TX;
    const int x0 = (tile % nbx) * TX, y0 = (tile / nbx) * TY;
    const int tid = threadIdx.x;

    int all0 = 1, all1 = 1;
    for (int k = tid; k < PH * PW; k += NT) {
        const int ly = k / PW, lx = k - ly * PW;
        const int gx = map_index<BC>(x0 + lx - 2, nx), gy = map_index<BC>(y0 + ly - 2, ny);
        const R v = a.phi_in[(size_t)gy * nx + gx];
        sP[k] = v;
        all0 &= is_pos_zero(v);
        all1 &= (v == (R)1.0);
    }
    int finite = 1;
    for (int k = tid; k < QH * QW; k += NT) {
        const int ly = k / QW, lx = k - ly * QW;
        const int gx = map_index<BC>(x0 + lx - 1, nx), gy = map_index<BC>(y0 + ly - 1, ny);
        const R v = a.T_in[(size_t)gy * nx + gx];
        sT[k] = v;
        finite &= isfinite(v) ? 1 : 0;
    }
    const int flat0 = __syncthreads_and(all0 & finite);
    const int flat1 = __syncthreads_and(all1 & finite);
    const bool flat = flat0 || flat1;

    if (!flat) {
        for (int k = tid; k < QH * QW; k += NT) {
            const int ly = k / QW, lx = k - ly * QW;
            const int ux = x0 + lx - 1, uy = y0 + ly - 1;  // unmapped global coords
            int cx, cy, sx = lx, sy = ly;
            if (BC == 1) {
                // zero-flux: a ghost cell carries the pass-1 values of the
                // boundary cell it clamps to (pass 2 clamps its indices too)
                cx = ux < 0 ? 0 : (ux >= nx ? nx - 1 : ux);
                cy = uy < 0 ? 0 : (uy >= ny ? ny - 1 : uy);
                sx += cx - ux;
                sy += cy - uy;
            } else {
                cx = map_index<0>(ux, nx);
                cy = map_index<0>(uy, ny);
            }
            const int c = (sy + 1) * PW + (sx + 1);
            const R gx = (sP[c + 1] - sP[c - 1]) * a.inv_dx;
            const R gy = (sP[c + PW] - sP[c - PW]) * a.inv_dy;
            const size_t gidx = (size_t)cy * nx + cx;
            R th;
            if (!theta_branch(gx, gy, th)) {
                th = a.theta[gidx];
            } else if (lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && ux < nx && uy < ny) {
                a.theta[gidx] = th;
            }
            R sn, cs;
            kg_sincos(a.aniso * (th - a.theta0), &sn, &cs);
            const R eps = a.eps_bar * ((R)1.0 + a.delta * cs);
            const R f = eps * (a.neg_ebar_j_delta * sn);
            sE2[k] = eps * eps;
            sA[k] = f * gx;
            sB[k] = f * gy;
        }
        __syncthreads();
    }

    for (int k = tid; k < TX * TY; k += NT) {
        const int ly = k / TX, lx = k - ly * TX;
        const int i = x0 + lx, j = y0 + ly;
        if (i >= nx || j >= ny)
            continue;
        const int cp = (ly + 2) * PW + (lx + 2);
        const int cq = (ly + 1) * QW + (lx + 1);
        const R phi = sP[cp], T = sT[cq];
        const R lapT = lap9_sum(sT, cq, QW) * a.inv_3dx2;
        R phin, Tn;
        if (flat) {
            phin = phi;
            Tn = T + lapT * a.dt + a.K * (phin - phi);
        } else {
            const R lap = lap9_sum(sP, cp, PW) * a.inv_3dx2;
            const R gx = (sP[cp + 1] - sP[cp - 1]) * a.inv_dx;
            const R gy = (sP[cp + PW] - sP[cp - PW]) * a.inv_dy;
            const R gEx = (sE2[cq + 1] - sE2[cq - 1]) * a.inv_dx;
            const R gEy = (sE2[cq + QW] - sE2[cq - QW]) * a.inv_dy;
            const R t1 = (sA[cq + QW] - sA[cq - QW]) * a.inv_dy;
            const R t2 = -(sB[cq + 1] - sB[cq - 1]) * a.inv_dx;
            const R t3 = gEx * gx + gEy * gy;
            update_cell(a, i, j, phi, T, lap, lapT, sE2[cq], t1, t2, t3, phin, Tn);
        }
        const size_t o = (size_t)j * nx + i;
        a.phi_out[o] = phin;
        a.T_out[o] = Tn;
    }
}

// ---- two-pass variant: the reference's structure on the GPU -------------
template <typename R, int BC>
__global__ __launch_bounds__(256) void k_pass1(StepArgs<R> a, PassTemps<R> t)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    if (i >= nx || j >= ny)
        return;
    const int ip = map_index<BC>(i + 1, nx), im = map_index<BC>(i - 1, nx);
    const int jp = map_index<BC>(j + 1, ny), jm = map_index<BC>(j - 1, ny);
    const R *p = a.phi_in;
    const R *T = a.T_in;
    auto at = [nx](const R *f, int x, int y) { return f[(size_t)y * nx + x]; };
    const size_t c = (size_t)j * nx + i;
    const R gx = (at(p, ip, j) - at(p, im, j)) * a.inv_dx;
    const R gy = (at(p, i, jp) - at(p, i, jm)) * a.inv_dy;
    t.gx[c] = gx;
    t.gy[c] = gy;
    t.lap[c] = ((R)2.0 * (at(p, ip, j) + at(p, im, j) + at(p, i, jp) + at(p, i, jm)) +
                at(p, ip, jp) + at(p, im, jm) + at(p, im, jp) + at(p, ip, jm) -
                (R)12.0 * at(p, i, j)) * a.inv_3dx2;
    t.lapT[c] = ((R)2.0 * (at(T, ip, j) + at(T, im, j) + at(T, i, jp) + at(T, i, jm)) +
                 at(T, ip, jp) + at(T, im, jm) + at(T, im, jp) + at(T, ip, jm) -
                 (R)12.0 * at(T, i, j)) * a.inv_3dx2;
    R th;
    if (theta_branch(gx, gy, th))
        a.theta[c] = th;
    else
        th = a.theta[c];
    R sn, cs;
    kg_sincos(a.aniso * (th - a.theta0), &sn, &cs);
    t.eps[c] = a.eps_bar * ((R)1.0 + a.delta * cs);
    t.epsd[c] = a.neg_ebar_j_delta * sn;
}

template <typename R, int BC>
__global__ __launch_bounds__(256) void k_pass2(StepArgs<R> a, PassTemps<R> t)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int nx = a.nx, ny = a.ny;
    if (i >= nx || j >= ny)
        return;
    const int ip = map_index<BC>(i + 1, nx), im = map_index<BC>(i - 1, nx);
    const int jp = map_index<BC>(j + 1, ny), jm = map_index<BC>(j - 1, ny);
    const size_t c = (size_t)j * nx + i;
    const size_t cE = (size_t)j * nx + ip, cW = (size_t)j * nx + im;
    const size_t cN = (size_t)jp * nx + i, cS = (size_t)jm * nx + i;
    const R *e = t.eps;
    const R *ed = t.epsd;
    const R gEx = (e[cE] * e[cE] - e[cW] * e[cW]) * a.inv_dx;
    const R gEy = (e[cN] * e[cN] - e[cS] * e[cS]) * a.inv_dy;
    const R t1 = (e[cN] * ed[cN] * t.gx[cN] - e[cS] * ed[cS] * t.gx[cS]) * a.inv_dy;
    const R t2 = -(e[cE] * ed[cE] * t.gy[cE] - e[cW] * ed[cW] * t.gy[cW]) * a.inv_dx;
    const R t3 = gEx * t.gx[c] + gEy * t.gy[c];
    R phin, Tn;
    update_cell(a, i, j, a.phi_in[c], a.T_in[c], t.lap[c], t.lapT[c], e[c] * e[c], t1, t2, t3,
                phin, Tn);
    a.phi_out[c] = phin;
    a.T_out[c] = Tn;
}

// ---- initial states -------------------------------------------------------
template <typename R>
__global__ void k_init(R *phi, R *T, R *theta, int nx, int ny, int kind, int x, int y,
                       const int *disks, int ndisks, double r2)
{
    const size_t n = (size_t)nx * ny;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        R v = (R)0;
        if (kind == 0) {
            // src/Kobayashi.cpp:116-123
            if ((j == y && (i == x || i == x - 1 || i == x + 1)) || (i == x && (j == y - 1 || j == y + 1)))
                v = (R)1;
        } else {
            for (int d = 0; d < ndisks; ++d) {
                const double di = (double)(i - disks[2 * d]), dj = (double)(j - disks[2 * d + 1]);
                if (di * di + dj * dj <= r2)
                    v = (R)1;
            }
        }
        phi[c] = v;
        T[c] = (R)0;
        theta[c] = (R)0;
    }
}

// ---- deterministic FP64 reduction -----------------------------------------
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
constexpr int RED_FIELDS = 7;

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partials(const R *phi, const R *T,
                                                                 size_t n, double K,
                                                                 double *partials)
{
    double s[RED_FIELDS] = {0.0, 0.0, 0.0, INFINITY, -INFINITY, 0.0, 0.0};
    for (size_t c = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; c < n;
         c += (size_t)RED_BLOCKS * RED_THREADS) {
        const double p = (double)phi[c], t = (double)T[c];
        if (!isfinite(p) || !isfinite(t)) {
            s[6] += 1.0;
            continue;
        }
        s[0] += p;
        s[1] += t;
        s[2] += t - K * p;
        s[3] = fmin(s[3], p);
        s[4] = fmax(s[4], p);
        s[5] += p > 0.5 ? 1.0 : 0.0;
    }
    __shared__ double sh[RED_FIELDS][RED_THREADS];
    for (int f = 0; f < RED_FIELDS; ++f)
        sh[f][threadIdx.x] = s[f];
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int f = 0; f < RED_FIELDS; ++f) {
                const double x = sh[f][threadIdx.x], y = sh[f][threadIdx.x + w];
                sh[f][threadIdx.x] = f == 3 ? fmin(x, y) : (f == 4 ? fmax(x, y) : x + y);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < RED_FIELDS)
        partials[(size_t)threadIdx.x * RED_BLOCKS + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double *partials,
                                                              ReduceOut *out)
{
    __shared__ double sh[RED_THREADS];
    double res[RED_FIELDS];
    for (int f = 0; f < RED_FIELDS; ++f) {
        double v = f == 3 ? INFINITY : (f == 4 ? -INFINITY : 0.0);
        for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS) {
            const double x = partials[(size_t)f * RED_BLOCKS + b];
            v = f == 3 ? fmin(v, x) : (f == 4 ? fmax(v, x) : v + x);
        }
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                const double x = sh[threadIdx.x], y = sh[threadIdx.x + w];
                sh[threadIdx.x] = f == 3 ? fmin(x, y) : (f == 4 ? fmax(x, y) : x + y);
            }
            __syncthreads();
        }
        res[f] = sh[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out->sum_phi = res[0];
        out->sum_T = res[1];
        out->sum_T_minus_K_phi = res[2];
        out->min_phi = res[3];
        out->max_phi = res[4];
        out->above_half = res[5];
        out->nonfinite = res[6];
    }
}

// ---- display ramp (src/Kobayashi.cpp:309-345) -----------------------------
template <typename R>
__global__ void k_colormap(const R *phi_f, int nx, int ny, int order, float4 *rgba)
{
    const size_t n = (size_t)nx * ny;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        size_t src = c;
        if (order == 1) {
            // constant buffer entry i shows _phi[_INDEX(i / n, i % n)]
            const size_t side = (size_t)nx;
            src = (c / side) + (size_t)nx * (c % side);
        }
        const float phi = (float)phi_f[src];
        const float c0[3] = {0.0f, 0.0f, 0.0f};
        const float c1[3] = {0.2505490f, 0.5000000f, 0.9882353f};
        const float c2[3] = {0.3607843f, 1.0000000f, 0.9882353f};
        const float c3[3] = {0.9005490f, 1.0000000f, 0.9882353f};
        const float b1 = 0.9f, b2 = 0.99f, b3 = 1.0f;
        const float *lo, *hi;
        float ratio;
        if (phi <= b1) {
            ratio = phi * (1.0f / b1);
            lo = c0;
            hi = c1;
        } else if (phi > b1 && phi <= b2) {
            ratio = (phi - b1) * (1.0f / (b2 - b1));
            lo = c1;
            hi = c2;
        } else {
            ratio = (phi - b2) * (1.0f / (b3 - b2));
            lo = c2;
            hi = c3;
        }
        float4 o;
        o.x = lo[0] * (1.0f - ratio) + hi[0] * ratio;
        o.y = lo[1] * (1.0f - ratio) + hi[1] * ratio;
        o.z = lo[2] * (1.0f - ratio) + hi[2] * ratio;
        o.w = 1.0f;
        rgba[c] = o;
    }
}

unsigned grid_1d(size_t n, unsigned threads)
{
    size_t b = (n + threads - 1) / threads;
    if (b > 65536)
        b = 65536;
    if (b < 1)
        b = 1;
    return (unsigned)b;
}

}  // namespace

void fused_geometry(int nx, int ny, int *ntiles, int *per_xcd, int *nblocks)
{
    const int tiles = ((nx + TX - 1) / TX) * ((ny + TY - 1) / TY);
    const int per = (tiles + NXCD - 1) / NXCD;
    *ntiles = tiles;
    *per_xcd = per;
    *nblocks = per * NXCD;
}

template <typename R>
hipError_t launch_step_fused(const StepArgs<R> &a, hipStream_t s)
{
    int ntiles, per, nblocks;
    fused_geometry(a.nx, a.ny, &ntiles, &per, &nblocks);
    StepArgs<R> b = a;
    b.ntiles = ntiles;
    b.per_xcd = per;
    if (a.boundary == 1)
        hipLaunchKernelGGL((k_step_fused<R, 1>), dim3(nblocks), dim3(NT), 0, s, b);
    else
        hipLaunchKernelGGL((k_step_fused<R, 0>), dim3(nblocks), dim3(NT), 0, s, b);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_step_two_pass(const StepArgs<R> &a, const PassTemps<R> &t, hipStream_t s)
{
    const dim3 grid((a.nx + 63) / 64, (a.ny + 3) / 4);
    if (a.boundary == 1) {
        hipLaunchKernelGGL((k_pass1<R, 1>), grid, dim3(256), 0, s, a, t);
        hipLaunchKernelGGL((k_pass2<R, 1>), grid, dim3(256), 0, s, a, t);
    } else {
        hipLaunchKernelGGL((k_pass1<R, 0>), grid, dim3(256), 0, s, a, t);
        hipLaunchKernelGGL((k_pass2<R, 0>), grid, dim3(256), 0, s, a, t);
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_init(R *phi, R *T, R *theta, int nx, int ny, int kind, int x, int y,
                       const int *disks, int ndisks, double r2, hipStream_t s)
{
    const size_t n = (size_t)nx * ny;
    hipLaunchKernelGGL((k_init<R>), dim3(grid_1d(n, 256)), dim3(256), 0, s, phi, T, theta, nx, ny,
                       kind, x, y, disks, ndisks, r2);
    return hipGetLastError();
}

size_t reduce_partials_size() { return (size_t)RED_FIELDS * RED_BLOCKS; }

template <typename R>
hipError_t launch_reduce(const R *phi, const R *T, size_t n, double K, double *partials,
                         ReduceOut *out, hipStream_t s)
{
    hipLaunchKernelGGL((k_reduce_partials<R>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, phi, T,
                       n, K, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, s, partials, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_colormap(const R *phi, int nx, int ny, int order, float *rgba, hipStream_t s)
{
    const size_t n = (size_t)nx * ny;
    hipLaunchKernelGGL((k_colormap<R>), dim3(grid_1d(n, 256)), dim3(256), 0, s, phi, nx, ny, order,
                       reinterpret_cast<float4 *>(rgba));
    return hipGetLastError();
}

#define KG_INSTANTIATE(R)                                                                        \
    template hipError_t launch_step_fused<R>(const StepArgs<R> &, hipStream_t);                  \
    template hipError_t launch_step_two_pass<R>(const StepArgs<R> &, const PassTemps<R> &,       \
                                                hipStream_t);                                    \
    template hipError_t launch_init<R>(R *, R *, R *, int, int, int, int, int, const int *, int, \
                                       double, hipStream_t);                                     \
    template hipError_t launch_reduce<R>(const R *, const R *, size_t, double, double *,         \
                                         ReduceOut *, hipStream_t);                              \
    template hipError_t launch_colormap<R>(const R *, int, int, int, float *, hipStream_t);
KG_INSTANTIATE(float)
KG_INSTANTIATE(double)

}  // namespace kg
