// kg_adjoint.hip -- reverse-mode (adjoint) of the Kobayashi step.
//
// Loss L = sum over cells of (phi_N - target)^2 after N steps, accumulated in
// FP64.  One reverse step maps the adjoint of the step's outputs
// (lphi', lT', ltheta') to the adjoint of its inputs (lphi, lT, ltheta_old)
// and adds the step's contribution to dL/d(parameter).  Derivation, with
// s = t1 + t2 + e^2 lap + t3 + r the bracket of src/Kobayashi.cpp:211-214,
// phi' = phi + s dt/tau, T' = T + lapT dt + K (phi' - phi):
//   ls        = (lphi' + K lT') dt/tau                          (adjoint of s)
//   l(e^2)(c) = sum over readers k of s(k):  +-ls(k) g(k)/dx (gEx, gEy) + ls(c) lap(c)
//   lA(c)     = ls(S-reader) / dy - ls(N-reader) / dy   (t1),   lB similarly (t2)
//   ltheta    = l(e^2) 2 e e' + (lA gx + lB gy)(e'^2 + e e'') + ltheta'
//   theta = atan branch: lgx += ltheta (-gy/|g|^2), lgy += ltheta gx/|g|^2
//   theta not rewritten (src/Kobayashi.cpp:154-167): ltheta_old = ltheta
//   lphi, lT  = gathers of lgx, lgy, ls e^2 (9-point) and lT' dt (9-point)
//               through the inverse neighbour maps, plus the local terms.
// "Readers" enumerate the inverse of the boundary map: periodic has one
// reader per direction; zero-flux boundary cells are also read by themselves.
// Every kernel is a gather (no atomics), so results are deterministic.
//
// Checkpoint-and-recompute: the forward run stores (phi, T, theta) every
// `every` steps; the reverse pass recomputes each segment's states with the
// same fused kernel (bitwise identical) and walks them backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "kg_host.h"
#include "kg_math.h"
#include "kg_philox.h"

namespace kg {

namespace {

constexpr int ADJ_THREADS = 256;
#ifndef KG_ADJ_CHUNK
#define KG_ADJ_CHUNK 256
#endif
constexpr int ADJ_CHUNK = KG_ADJ_CHUNK;  // cells per block
enum { P_TAU, P_EPS_BAR, P_K, P_DELTA, P_THETA0, P_ALPHA, P_GAMMA, P_TEQ, P_NOISE, NPAR };

template <typename R>
struct AdjTemps {
    R *gx, *gy, *lap, *lapT, *th, *eps, *epsd, *cs, *sn;   // pass 1 of the forward step
    R *ls, *gEx, *gEy, *lphi_loc, *lT_loc;                  // reverse of pass 2
    R *lgx, *lgy, *lth_old;                                 // reverse of pass 1 (per cell)
    signed char *code;                                      // theta: 0 kept, 1 atan, 2 axis
};
constexpr int NTEMP = 17;

template <typename R>
struct AdjIO {
    const R *phi, *T, *theta_old;   // the step's input state
    const R *lphi1, *lT1, *lth1;    // adjoint of the step's outputs
    R *lphi0, *lT0, *lth0;          // adjoint of the step's inputs
};

// One grid of the (possibly batched) adjoint: its step parameters and the
// element offset of its fields in every buffer.  Blocks take chunks (member,
// first cell, cell count, accumulator slot) from a table.
template <typename R>
struct AdjMember {
    StepArgs<R> a;
    size_t off;
    R tau;
};

template <typename R>
__device__ __forceinline__ AdjIO<R> shift(AdjIO<R> io, size_t o)
{
    io.phi += o; io.T += o; io.theta_old += o;
    io.lphi1 += o; io.lT1 += o; io.lth1 += o;
    io.lphi0 += o; io.lT0 += o; io.lth0 += o;
    return io;
}

template <typename R>
__device__ __forceinline__ AdjTemps<R> shift(AdjTemps<R> t, size_t o)
{
    R **f[] = {&t.gx, &t.gy, &t.lap, &t.lapT, &t.th, &t.eps, &t.epsd, &t.cs, &t.sn,
               &t.ls, &t.gEx, &t.gEy, &t.lphi_loc, &t.lT_loc, &t.lgx, &t.lgy, &t.lth_old};
    for (R **q : f)
        *q += o;
    t.code += o;
    return t;
}

template <int BC>
__device__ __forceinline__ int wrap(int x, int n)
{
    if (BC == 1)
        return x < 0 ? 0 : (x >= n ? n - 1 : x);
    return x < 0 ? x + n : (x >= n ? x - n : x);
}

// Cells k with map(k + o) == c along one axis (o in {-1, 0, 1}).
template <int BC>
__device__ __forceinline__ int readers(int c, int o, int n, int ks[2])
{
    if (o == 0) {
        ks[0] = c;
        return 1;
    }
    if (BC == 0) {
        int k = c - o;
        ks[0] = k < 0 ? k + n : (k >= n ? k - n : k);
        return 1;
    }
    int m = 0;
    const int k = c - o;
    if (k >= 0 && k < n)
        ks[m++] = k;
    if ((o == 1 && c == n - 1) || (o == -1 && c == 0))
        ks[m++] = c;
    return m;
}

template <typename R> __device__ __forceinline__ R pi_of();
template <> __device__ __forceinline__ float pi_of<float>() { return 3.14159274f; }
template <> __device__ __forceinline__ double pi_of<double>() { return 3.14159265358979323846; }

// theta update of src/Kobayashi.cpp:154-167 with the branch class:
// 0 = theta kept, 1 = offset + atan(gy/gx), 2 = +-pi/2.
template <typename R>
__device__ __forceinline__ int theta_class(R gx, R gy, R &th)
{
    const R E = (R)1.1920928955078125e-7;
    const R PI = pi_of<R>();
    const R t = kg_atan_ratio(gy, gx);
    if (gx > E) {
        if (gy < -E) { th = (R)2.0 * PI + t; return 1; }
        if (gy > E) { th = t; return 1; }
        return 0;
    }
    if (gx < -E) { th = PI + t; return 1; }
    if (gx <= E && gx >= -E) {
        if (gy < -E) { th = (R)-0.5 * PI; return 2; }
        if (gy > E) { th = (R)0.5 * PI; return 2; }
    }
    return 0;
}

template <typename R>
__device__ __forceinline__ void block_add(double (&v)[NPAR], double *acc, int slot, int nslots)
{
    __shared__ double sh[NPAR][ADJ_THREADS];
    for (int p = 0; p < NPAR; ++p)
        sh[p][threadIdx.x] = v[p];
    __syncthreads();
    for (int w = ADJ_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int p = 0; p < NPAR; ++p)
                sh[p][threadIdx.x] += sh[p][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < NPAR)
        acc[(size_t)threadIdx.x * nslots + slot] += sh[threadIdx.x][0];
}

// A1: pass-1 quantities of the forward step (same arithmetic as the kernels)
template <typename R, int BC>
__device__ __forceinline__ void adj_pass1(const StepArgs<R> &a, const AdjIO<R> &io,
                                          const AdjTemps<R> &t, size_t c0, size_t c1)
{
    const int nx = a.nx, ny = a.ny;
    for (size_t c = c0 + threadIdx.x; c < c1; c += ADJ_THREADS) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        const int ip = wrap<BC>(i + 1, nx), im = wrap<BC>(i - 1, nx);
        const int jp = wrap<BC>(j + 1, ny), jm = wrap<BC>(j - 1, ny);
        const R *p = io.phi, *T = io.T;
        auto at = [nx](const R *f, int x, int y) { return f[(size_t)y * nx + x]; };
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
        const int code = theta_class(gx, gy, th);
        if (code == 0)
            th = io.theta_old[c];
        t.code[c] = (signed char)code;
        t.th[c] = th;
        R sn, cs;
        kg_sincos(a.aniso * (th - a.theta0), &sn, &cs);
        t.cs[c] = cs;
        t.sn[c] = sn;
        t.eps[c] = a.eps_bar * ((R)1.0 + a.delta * cs);
        t.epsd[c] = a.neg_ebar_j_delta * sn;
    }
}

// A2: reverse of pass 2 at each cell (forward neighbour maps)
template <typename R, int BC>
__device__ __forceinline__ void adj_pass2(const StepArgs<R> &a, const AdjIO<R> &io,
                                          const AdjTemps<R> &t, R tau, size_t c0, size_t c1,
                                          double (&v)[NPAR])
{
    const int nx = a.nx, ny = a.ny;
    for (size_t c = c0 + threadIdx.x; c < c1; c += ADJ_THREADS) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        const size_t cE = (size_t)j * nx + wrap<BC>(i + 1, nx);
        const size_t cW = (size_t)j * nx + wrap<BC>(i - 1, nx);
        const size_t cN = (size_t)wrap<BC>(j + 1, ny) * nx + i;
        const size_t cS = (size_t)wrap<BC>(j - 1, ny) * nx + i;
        const R *e = t.eps, *ed = t.epsd;
        const R gEx = (e[cE] * e[cE] - e[cW] * e[cW]) * a.inv_dx;
        const R gEy = (e[cN] * e[cN] - e[cS] * e[cS]) * a.inv_dy;
        const R t1 = (e[cN] * ed[cN] * t.gx[cN] - e[cS] * ed[cS] * t.gx[cS]) * a.inv_dy;
        const R t2 = -(e[cE] * ed[cE] * t.gy[cE] - e[cW] * ed[cW] * t.gy[cW]) * a.inv_dx;
        const R t3 = gEx * t.gx[c] + gEy * t.gy[c];
        const R phi = io.phi[c], T = io.T[c];
        const R u = a.gamma * (a.teq - T);
        const R atn = kg_atan(u);
        const R m = a.alpha_pi * atn;
        R chi = (R)0;
        if (a.noise_a != (R)0) {
            const u32x4 w = philox4x32_10(u32x4{(uint32_t)i >> 2, (uint32_t)j, a.step, a.instance},
                                          a.key0, a.key1);
            chi = noise_from_word<R>(select_word(w, (uint32_t)i & 3u));
        }
        const R q = phi - (R)0.5 + m + a.noise_a * chi;
        const R w1 = phi * ((R)1.0 - phi);
        const R s = t1 + t2 + e[c] * e[c] * t.lap[c] + t3 + w1 * q;
        const R lp1 = io.lphi1[c], lT1 = io.lT1[c];
        const R ls = (lp1 + a.K * lT1) * a.dt_tau;
        const R d = (R)1.0 + u * u;
        v[P_TAU] += (double)ls * (double)(-s / tau);
        v[P_K] += (double)lT1 * (double)(s * a.dt_tau);
        v[P_ALPHA] += (double)ls * (double)(w1 * atn) / (double)pi_of<R>();
        v[P_GAMMA] += (double)ls * (double)(w1 * a.alpha_pi * (a.teq - T) / d);
        v[P_TEQ] += (double)ls * (double)(w1 * a.alpha_pi * a.gamma / d);
        v[P_NOISE] += (double)ls * (double)(w1 * chi);
        t.ls[c] = ls;
        t.gEx[c] = gEx;
        t.gEy[c] = gEy;
        // phi enters phi' directly (coefficient 1) and through s; T' = T + lapT dt
        // + K s dt/tau has no other phi dependence (its K lT' is inside ls)
        t.lphi_loc[c] = lp1 + ls * (((R)1.0 - (R)2.0 * phi) * q + w1);
        t.lT_loc[c] = lT1 - ls * w1 * a.alpha_pi * a.gamma / d;
    }
}

// A3: reverse of the pass-1 quantities at each cell (inverse neighbour maps)
template <typename R, int BC>
__device__ __forceinline__ void adj_cell(const StepArgs<R> &a, const AdjIO<R> &io,
                                         const AdjTemps<R> &t, size_t c0, size_t c1,
                                         double (&v)[NPAR])
{
    const int nx = a.nx, ny = a.ny;
    for (size_t c = c0 + threadIdx.x; c < c1; c += ADJ_THREADS) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        R le2 = t.ls[c] * t.lap[c], lA = (R)0, lB = (R)0;
        int ks[2];
        int m = readers<BC>(i, 1, nx, ks);          // k with E(k) = c
        for (int q = 0; q < m; ++q) {
            const size_t k = (size_t)j * nx + ks[q];
            le2 += t.ls[k] * t.gx[k] * a.inv_dx;
            lB -= t.ls[k] * a.inv_dx;
        }
        m = readers<BC>(i, -1, nx, ks);             // k with W(k) = c
        for (int q = 0; q < m; ++q) {
            const size_t k = (size_t)j * nx + ks[q];
            le2 -= t.ls[k] * t.gx[k] * a.inv_dx;
            lB += t.ls[k] * a.inv_dx;
        }
        m = readers<BC>(j, 1, ny, ks);              // k with N(k) = c
        for (int q = 0; q < m; ++q) {
            const size_t k = (size_t)ks[q] * nx + i;
            le2 += t.ls[k] * t.gy[k] * a.inv_dy;
            lA += t.ls[k] * a.inv_dy;
        }
        m = readers<BC>(j, -1, ny, ks);             // k with S(k) = c
        for (int q = 0; q < m; ++q) {
            const size_t k = (size_t)ks[q] * nx + i;
            le2 -= t.ls[k] * t.gy[k] * a.inv_dy;
            lA -= t.ls[k] * a.inv_dy;
        }
        const R eps = t.eps[c], epsd = t.epsd[c], cs = t.cs[c], sn = t.sn[c];
        const R gx = t.gx[c], gy = t.gy[c];
        const R F = eps * epsd;
        const R epsd_th = a.neg_ebar_j_delta * a.aniso * cs;  // d eps' / d theta
        const R F_th = epsd * epsd + eps * epsd_th;
        const R G = lA * gx + lB * gy;
        const R lth = le2 * (R)2.0 * eps * epsd + G * F_th + io.lth1[c];
        R lgx = lA * F + t.ls[c] * t.gEx[c];
        R lgy = lB * F + t.ls[c] * t.gEy[c];
        R lth_old = (R)0;
        const int code = t.code[c];
        if (code == 1) {
            const R g2 = gx * gx + gy * gy;
            lgx += lth * (-gy / g2);
            lgy += lth * (gx / g2);
        } else if (code == 0) {
            lth_old = lth;
        }
        t.lgx[c] = lgx;
        t.lgy[c] = lgy;
        t.lth_old[c] = lth_old;
        const R de_dd = a.eps_bar * cs, ded_dd = -a.eps_bar * a.aniso * sn;
        v[P_DELTA] += (double)(le2 * (R)2.0 * eps * de_dd) + (double)(G * (de_dd * epsd + eps * ded_dd));
        v[P_EPS_BAR] += (double)(le2 * (R)2.0 * eps * eps / a.eps_bar) + (double)(G * (R)2.0 * F / a.eps_bar);
        v[P_THETA0] -= (double)(le2 * (R)2.0 * eps * epsd) + (double)(G * F_th);
    }
}

// A4: gather of the input adjoints
template <typename R, int BC>
__device__ __forceinline__ void adj_gather(const StepArgs<R> &a, const AdjIO<R> &io,
                                           const AdjTemps<R> &t, size_t c0, size_t c1)
{
    const int nx = a.nx, ny = a.ny;
    for (size_t c = c0 + threadIdx.x; c < c1; c += ADJ_THREADS) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        R lp = t.lphi_loc[c], lT = t.lT_loc[c];
        int kx[2], ky[2];
        int m = readers<BC>(i, 1, nx, kx);
        for (int q = 0; q < m; ++q)
            lp += t.lgx[(size_t)j * nx + kx[q]] * a.inv_dx;
        m = readers<BC>(i, -1, nx, kx);
        for (int q = 0; q < m; ++q)
            lp -= t.lgx[(size_t)j * nx + kx[q]] * a.inv_dx;
        m = readers<BC>(j, 1, ny, ky);
        for (int q = 0; q < m; ++q)
            lp += t.lgy[(size_t)ky[q] * nx + i] * a.inv_dy;
        m = readers<BC>(j, -1, ny, ky);
        for (int q = 0; q < m; ++q)
            lp -= t.lgy[(size_t)ky[q] * nx + i] * a.inv_dy;
        for (int oy = -1; oy <= 1; ++oy) {
            const int my = readers<BC>(j, oy, ny, ky);
            for (int ox = -1; ox <= 1; ++ox) {
                const R w = (ox == 0 && oy == 0) ? (R)-12.0 : ((ox == 0 || oy == 0) ? (R)2.0 : (R)1.0);
                const int mx = readers<BC>(i, ox, nx, kx);
                for (int qy = 0; qy < my; ++qy)
                    for (int qx = 0; qx < mx; ++qx) {
                        const size_t k = (size_t)ky[qy] * nx + kx[qx];
                        lp += w * a.inv_3dx2 * t.ls[k] * t.eps[k] * t.eps[k];
                        lT += w * a.inv_3dx2 * a.dt * io.lT1[k];
                    }
            }
        }
        io.lphi0[c] = lp;
        io.lT0[c] = lT;
        io.lth0[c] = t.lth_old[c];
    }
}

// The four reverse kernels: each block takes one chunk of one member.
template <typename R, int KIND>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj(const AdjMember<R> *__restrict__ members,
                                                     const int4 *__restrict__ chunks, AdjIO<R> io,
                                                     AdjTemps<R> t, double *acc, int nslots,
                                                     uint32_t step)
{
    const int4 ch = chunks[blockIdx.x];
    AdjMember<R> m = members[ch.x];
    m.a.step = step;  // the noise counter of the step being reversed
    const AdjIO<R> mio = shift(io, m.off);
    const AdjTemps<R> mt = shift(t, m.off);
    const size_t c0 = (size_t)ch.y, c1 = c0 + (size_t)ch.z;
    double v[NPAR] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool zf = m.a.boundary == 1;
    if (KIND == 0) {
        if (zf) adj_pass1<R, 1>(m.a, mio, mt, c0, c1); else adj_pass1<R, 0>(m.a, mio, mt, c0, c1);
    } else if (KIND == 1) {
        if (zf) adj_pass2<R, 1>(m.a, mio, mt, m.tau, c0, c1, v); else adj_pass2<R, 0>(m.a, mio, mt, m.tau, c0, c1, v);
    } else if (KIND == 2) {
        if (zf) adj_cell<R, 1>(m.a, mio, mt, c0, c1, v); else adj_cell<R, 0>(m.a, mio, mt, c0, c1, v);
    } else {
        if (zf) adj_gather<R, 1>(m.a, mio, mt, c0, c1); else adj_gather<R, 0>(m.a, mio, mt, c0, c1);
    }
    if (KIND == 1 || KIND == 2)
        block_add<R>(v, acc, ch.w, nslots);
}

// loss partial of each chunk = sum (phi - target)^2 in FP64; lphi = 2 (phi -
// target), lT = ltheta = 0
template <typename R>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_seed(const AdjMember<R> *__restrict__ members,
                                                          const int4 *__restrict__ chunks,
                                                          const R *phi, const R *target, R *lphi,
                                                          R *lT, R *lth, double *partials)
{
    const int4 ch = chunks[blockIdx.x];
    const size_t o = members[ch.x].off;
    double s = 0.0;
    for (size_t c = (size_t)ch.y + threadIdx.x; c < (size_t)ch.y + ch.z; c += ADJ_THREADS) {
        const R d = phi[o + c] - target[o + c];
        s += (double)d * (double)d;
        lphi[o + c] = (R)2.0 * d;
        lT[o + c] = (R)0;
        lth[o + c] = (R)0;
    }
    __shared__ double sh[ADJ_THREADS];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = ADJ_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[blockIdx.x] = sh[0];
}

// The gather of one reverse step (A4, step s) and the pass-1 quantities of
// the next one (A1, step s - 1) in one launch: per cell the two are
// independent -- A1 reads the state of step s - 1 and writes its own temps,
// A4 reads temps of step s -- except eps, which both touch; the driver gives
// A1 the other of two eps buffers (t1.eps != t.eps).
template <typename R>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_gather_pass1(
    const AdjMember<R> *__restrict__ members, const int4 *__restrict__ chunks, AdjIO<R> io,
    AdjTemps<R> t, AdjIO<R> io1, AdjTemps<R> t1)
{
    const int4 ch = chunks[blockIdx.x];
    const AdjMember<R> m = members[ch.x];
    const AdjIO<R> mio = shift(io, m.off), mio1 = shift(io1, m.off);
    const AdjTemps<R> mt = shift(t, m.off), mt1 = shift(t1, m.off);
    const size_t c0 = (size_t)ch.y, c1 = c0 + (size_t)ch.z;
    if (m.a.boundary == 1) {
        adj_gather<R, 1>(m.a, mio, mt, c0, c1);
        adj_pass1<R, 1>(m.a, mio1, mt1, c0, c1);
    } else {
        adj_gather<R, 0>(m.a, mio, mt, c0, c1);
        adj_pass1<R, 0>(m.a, mio1, mt1, c0, c1);
    }
}

// The fused launch pays where a reverse step is bound by launches and
// dependent round trips, not by bandwidth: up to about one round of resident
// blocks.  Measured on MI355X (FP32): 512^2 (1024 chunks) 15.1 -> 13.2 us per
// gather + pass-1 pair, the whole adjoint 0.0472 -> 0.0450 s; 64 members of
// 128^2..512^2 (about 25 000 chunks) 0.296 -> 0.316 s per batched gradient,
// slower.  Larger launches keep the separate kernels.
constexpr int ADJ_FUSE_MAX_CHUNKS = 2048;

// One reverse step, step counter `step`, on temps t.  first: the step's
// pass-1 quantities are not in t yet (A1 runs here); next != nullptr: the
// state of the step before, whose A1 joins this step's gather, writing the
// temps *next_t (the other eps buffer).
template <typename R>
hipError_t reverse_step(const AdjMember<R> *members, const int4 *chunks, int nchunks,
                        const AdjIO<R> &io, const AdjTemps<R> &t, double *acc, uint32_t step,
                        bool first, const AdjIO<R> *next, const AdjTemps<R> *next_t, hipStream_t s)
{
    const dim3 g(nchunks), b(ADJ_THREADS);
    if (first)
        hipLaunchKernelGGL((k_adj<R, 0>), g, b, 0, s, members, chunks, io, t, acc, nchunks, step);
    hipLaunchKernelGGL((k_adj<R, 1>), g, b, 0, s, members, chunks, io, t, acc, nchunks, step);
    hipLaunchKernelGGL((k_adj<R, 2>), g, b, 0, s, members, chunks, io, t, acc, nchunks, step);
    if (next)
        hipLaunchKernelGGL((k_adj_gather_pass1<R>), g, b, 0, s, members, chunks, io, t, *next,
                           *next_t);
    else
        hipLaunchKernelGGL((k_adj<R, 3>), g, b, 0, s, members, chunks, io, t, acc, nchunks, step);
    return hipGetLastError();
}

}  // namespace

struct AdjointWork {
    size_t total = 0;        // elements per field buffer
    size_t esize = 0;
    int ckpt_cap = 0, seg_cap = 0, nmembers = 0, nchunks = 0;
    char *ckpt = nullptr;    // ckpt_cap states of 3 fields
    char *seg = nullptr;     // seg_cap + 1 states of 3 fields
    char *lam = nullptr;     // 2 x (lphi, lT, ltheta)
    char *tmp = nullptr;     // NTEMP fields, target, a second eps
    signed char *code = nullptr;
    double *acc = nullptr;   // NPAR x nchunks
    double *partials = nullptr;
    void *members = nullptr; // AdjMember<R>[nmembers]
    int4 *chunks = nullptr;
    void *tables = nullptr;  // batch recompute: seg_cap tables of nmembers StepArgs
};

void free_adjoint(AdjointWork *w)
{
    if (!w)
        return;
    void *p[] = {w->ckpt, w->seg, w->lam, w->tmp, w->code, w->acc, w->partials, w->members,
                 w->chunks, w->tables};
    for (void *q : p)
        if (q)
            (void)hipFree(q);
    delete w;
}

namespace {

// Buffers for `total` elements per field, sized for nckpt checkpoints and
// segments of seglen steps; reused while large enough.
int ensure_work(AdjointWork *&w, size_t total, size_t es, int nckpt, int seglen, int nmembers,
                int nchunks, size_t args_bytes, bool tables)
{
    const size_t fb = total * es;
    if (w && (w->total != total || w->esize != es || w->ckpt_cap < nckpt || w->seg_cap < seglen ||
              w->nmembers != nmembers || w->nchunks != nchunks || (tables && !w->tables))) {
        free_adjoint(w);
        w = nullptr;
    }
    if (w)
        return KG_OK;
    w = new AdjointWork();
    w->total = total;
    w->esize = es;
    w->ckpt_cap = nckpt;
    w->seg_cap = seglen;
    w->nmembers = nmembers;
    w->nchunks = nchunks;
    bool ok = hipMalloc(&w->ckpt, 3 * fb * (size_t)nckpt) == hipSuccess &&
              hipMalloc(&w->seg, 3 * fb * (size_t)(seglen + 1)) == hipSuccess &&
              hipMalloc(&w->lam, 6 * fb) == hipSuccess &&
              hipMalloc(&w->tmp, (NTEMP + 2) * fb) == hipSuccess &&
              hipMalloc(&w->code, total) == hipSuccess &&
              hipMalloc(&w->acc, (size_t)NPAR * nchunks * sizeof(double)) == hipSuccess &&
              hipMalloc(&w->partials, (size_t)nchunks * sizeof(double)) == hipSuccess &&
              hipMalloc(&w->members, args_bytes) == hipSuccess &&
              hipMalloc(&w->chunks, (size_t)nchunks * sizeof(int4)) == hipSuccess &&
              (!tables || hipMalloc(&w->tables, (size_t)seglen * nmembers * sizeof(StepArgs<double>)) ==
                              hipSuccess);
    if (!ok) {
        free_adjoint(w);
        w = nullptr;
        return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the adjoint buffers failed");
    }
    return KG_OK;
}

// What the driver needs from a single grid or a batch.
template <typename R>
struct AdjJob {
    hipStream_t stream;
    size_t total;                       // elements per field buffer
    std::vector<int> nx, ny;
    std::vector<size_t> off;
    std::vector<kg_params> params;
    uint64_t s0;                        // noise counter of the first step
    R *cur_phi, *cur_T, *cur_theta;     // current state buffers (updated by fwd)
    std::function<int()> fwd;           // one step on the job's own buffers
    // one recompute step i: state seg[i] -> seg[i+1] with step counter `step`
    std::function<int(int i, R *const *in, R *const *out, uint32_t step)> recompute;
};

template <typename R>
int adjoint_core(AdjJob<R> &job, AdjointWork *&work, int nsteps, int every, const R *target_host,
                 std::vector<kg_grads> &grads, R *d_T0_dev_copy_host, R *d_phi0_host,
                 bool tables)
{
    const int M = (int)job.nx.size();
    const size_t total = job.total, es = sizeof(R), fb = total * es;
    const int nseg = (nsteps + every - 1) / every;
    std::vector<int4> chunks;
    for (int m = 0; m < M; ++m) {
        const size_t n = (size_t)job.nx[m] * job.ny[m];
        for (size_t c = 0; c < n; c += ADJ_CHUNK)
            chunks.push_back(make_int4(m, (int)c, (int)std::min((size_t)ADJ_CHUNK, n - c), 0));
    }
    const int nch = (int)chunks.size();
    for (int k = 0; k < nch; ++k)
        chunks[k].w = k;
    if (int rc = ensure_work(work, total, es, nseg, every, M, nch, (size_t)M * sizeof(AdjMember<R>),
                             tables))
        return rc;
    AdjointWork &w = *work;
    const hipStream_t s = job.stream;
    std::vector<AdjMember<R>> mem(M);
    for (int m = 0; m < M; ++m) {
        mem[m].a = make_args<R>(job.params[m], job.nx[m], job.ny[m]);
        mem[m].off = job.off[m];
        mem[m].tau = (R)job.params[m].tau;
    }
    KG_HIP(hipMemcpyAsync(w.members, mem.data(), M * sizeof(AdjMember<R>), hipMemcpyHostToDevice, s));
    KG_HIP(hipMemcpyAsync(w.chunks, chunks.data(), nch * sizeof(int4), hipMemcpyHostToDevice, s));
    KG_HIP(hipStreamSynchronize(s));
    auto st = [&](char *base, int k, int f) { return reinterpret_cast<R *>(base + (3 * (size_t)k + f) * fb); };
    R *tgt = reinterpret_cast<R *>(w.tmp + (size_t)NTEMP * fb);
    KG_HIP(hipMemcpyAsync(tgt, target_host, fb, hipMemcpyHostToDevice, s));

    // forward with checkpoints
    for (int k = 0; k < nseg; ++k) {
        KG_HIP(hipMemcpyAsync(st(w.ckpt, k, 0), job.cur_phi, fb, hipMemcpyDeviceToDevice, s));
        KG_HIP(hipMemcpyAsync(st(w.ckpt, k, 1), job.cur_T, fb, hipMemcpyDeviceToDevice, s));
        KG_HIP(hipMemcpyAsync(st(w.ckpt, k, 2), job.cur_theta, fb, hipMemcpyDeviceToDevice, s));
        const int len = std::min(every, nsteps - k * every);
        for (int i = 0; i < len; ++i)
            if (int rc = job.fwd())
                return rc;
    }

    // loss and the seed adjoint
    R *lam[2][3];
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 3; ++f)
            lam[b][f] = reinterpret_cast<R *>(w.lam + (3 * (size_t)b + f) * fb);
    const AdjMember<R> *dmem = static_cast<const AdjMember<R> *>(w.members);
    hipLaunchKernelGGL((k_adj_seed<R>), dim3(nch), dim3(ADJ_THREADS), 0, s, dmem, w.chunks,
                       (const R *)job.cur_phi, (const R *)tgt, lam[0][0], lam[0][1], lam[0][2],
                       w.partials);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemsetAsync(w.acc, 0, (size_t)NPAR * nch * sizeof(double), s));

    AdjTemps<R> t;
    R *tp = reinterpret_cast<R *>(w.tmp);
    R **fields[] = {&t.gx, &t.gy, &t.lap, &t.lapT, &t.th, &t.eps, &t.epsd, &t.cs, &t.sn,
                    &t.ls, &t.gEx, &t.gEy, &t.lphi_loc, &t.lT_loc, &t.lgx, &t.lgy, &t.lth_old};
    for (int f = 0; f < NTEMP; ++f)
        *fields[f] = tp + (size_t)f * total;
    t.code = w.code;
    // eps of the step being reversed and of the one before (the fused
    // gather + pass-1 launch writes the latter while reading the former)
    R *eps_buf[2] = {t.eps, tp + (size_t)(NTEMP + 1) * total};
    int eb = 0;

    int lc = 0;
    bool first = true;  // the step about to be reversed still needs its pass 1
    for (int k = nseg - 1; k >= 0; --k) {
        const int len = std::min(every, nsteps - k * every);
        for (int f = 0; f < 3; ++f)
            KG_HIP(hipMemcpyAsync(st(w.seg, 0, f), st(w.ckpt, k, f), fb, hipMemcpyDeviceToDevice, s));
        for (int i = 0; i < len; ++i) {  // recompute the segment's states
            KG_HIP(hipMemcpyAsync(st(w.seg, i + 1, 2), st(w.seg, i, 2), fb, hipMemcpyDeviceToDevice, s));
            R *in[3] = {st(w.seg, i, 0), st(w.seg, i, 1), st(w.seg, i + 1, 2)};
            R *out[3] = {st(w.seg, i + 1, 0), st(w.seg, i + 1, 1), st(w.seg, i + 1, 2)};
            if (int rc = job.recompute(i, in, out, (uint32_t)(job.s0 + (uint64_t)k * every + i)))
                return rc;
        }
        for (int i = len - 1; i >= 0; --i) {
            AdjIO<R> io{st(w.seg, i, 0), st(w.seg, i, 1), st(w.seg, i, 2),
                        lam[lc][0], lam[lc][1], lam[lc][2],
                        lam[lc ^ 1][0], lam[lc ^ 1][1], lam[lc ^ 1][2]};
            t.eps = eps_buf[eb];
            // the step before, within this segment: its A1 joins this gather
            const bool fuse = i > 0 && nch <= ADJ_FUSE_MAX_CHUNKS;
            AdjIO<R> io_prev = io;
            AdjTemps<R> t_prev = t;
            if (fuse) {
                io_prev.phi = st(w.seg, i - 1, 0);
                io_prev.T = st(w.seg, i - 1, 1);
                io_prev.theta_old = st(w.seg, i - 1, 2);
                t_prev.eps = eps_buf[eb ^ 1];
            }
            const hipError_t e = reverse_step<R>(
                dmem, w.chunks, nch, io, t, w.acc, (uint32_t)(job.s0 + (uint64_t)k * every + i),
                first, fuse ? &io_prev : nullptr, fuse ? &t_prev : nullptr, s);
            if (e != hipSuccess)
                return hip_fail(e, "adjoint step launch");
            lc ^= 1;
            first = !fuse;  // the next step's pass 1 is still to run unless it ran here
            if (fuse)
                eb ^= 1;
        }
    }

    std::vector<double> acc((size_t)NPAR * nch), parts(nch);
    KG_HIP(hipMemcpyAsync(acc.data(), w.acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    KG_HIP(hipMemcpyAsync(parts.data(), w.partials, parts.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    if (d_T0_dev_copy_host)
        KG_HIP(hipMemcpyAsync(d_T0_dev_copy_host, lam[lc][1], fb, hipMemcpyDeviceToHost, s));
    if (d_phi0_host)
        KG_HIP(hipMemcpyAsync(d_phi0_host, lam[lc][0], fb, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
    grads.assign(M, kg_grads{});
    for (int k = 0; k < nch; ++k) {
        kg_grads &g = grads[chunks[k].x];
        double *q = &g.d_tau;
        const int map[NPAR] = {P_TAU, P_EPS_BAR, P_K, P_DELTA, P_THETA0, P_ALPHA, P_GAMMA, P_TEQ,
                               P_NOISE};
        for (int p = 0; p < NPAR; ++p)
            q[p] += acc[(size_t)map[p] * nch + k];
        g.loss += parts[k];
    }
    return KG_OK;
}

template <typename R>
int adjoint_single(kg_solver_s &h, int nsteps, int every, const void *target, kg_grads *out,
                   void *d_T0, void *d_phi0)
{
    AdjJob<R> job;
    job.stream = h.stream;
    job.total = h.cells();
    job.nx = {h.nx};
    job.ny = {h.ny};
    job.off = {0};
    job.params = {h.params};
    job.s0 = h.steps;
    job.cur_phi = static_cast<R *>(h.phi[h.cur]);
    job.cur_T = static_cast<R *>(h.T[h.cur]);
    job.cur_theta = static_cast<R *>(h.theta);
    // the forward steps below run without tile status: it no longer describes
    // the handle's buffers afterwards
    h.stat_valid = false;
    StepArgs<R> base = make_args<R>(h.params, h.nx, h.ny);
    base.wide_angles = wide_angles(h.params, h.theta_given);
    job.fwd = [&]() -> int {
        StepArgs<R> a = base;
        a.phi_in = static_cast<const R *>(h.phi[h.cur]);
        a.T_in = static_cast<const R *>(h.T[h.cur]);
        a.phi_out = static_cast<R *>(h.phi[h.cur ^ 1]);
        a.T_out = static_cast<R *>(h.T[h.cur ^ 1]);
        a.theta = static_cast<R *>(h.theta);
        a.step = (uint32_t)h.steps;
        KG_HIP(launch_step_fused<R>(a, h.stream));
        h.cur ^= 1;
        h.steps += 1;
        job.cur_phi = static_cast<R *>(h.phi[h.cur]);
        job.cur_T = static_cast<R *>(h.T[h.cur]);
        return KG_OK;
    };
    job.recompute = [&](int, R *const *in, R *const *o, uint32_t step) -> int {
        StepArgs<R> a = base;
        a.phi_in = in[0];
        a.T_in = in[1];
        a.phi_out = o[0];
        a.T_out = o[1];
        a.theta = o[2];
        a.step = step;
        KG_HIP(launch_step_fused<R>(a, h.stream));
        return KG_OK;
    };
    std::vector<kg_grads> g;
    if (int rc = adjoint_core<R>(job, h.adj, nsteps, every, static_cast<const R *>(target), g,
                                 static_cast<R *>(d_T0), static_cast<R *>(d_phi0), false))
        return rc;
    *out = g[0];
    return KG_OK;
}

template <typename R>
int adjoint_batch(kg_batch_s &b, int nsteps, int every, const void *const *targets, kg_grads *out,
                  void *const *d_T0)
{
    const int M = b.count;
    AdjJob<R> job;
    job.stream = b.stream;
    job.total = b.total;
    job.nx = b.nx;
    job.ny = b.ny;
    job.off = b.off;
    job.params = b.params;
    job.s0 = b.steps;
    job.cur_phi = static_cast<R *>(b.phi[b.cur()]);
    job.cur_T = static_cast<R *>(b.T[b.cur()]);
    job.cur_theta = static_cast<R *>(b.theta);
    job.fwd = [&]() -> int {
        if (int rc = kg_batch_step(&b, 1))
            return rc;
        job.cur_phi = static_cast<R *>(b.phi[b.cur()]);
        job.cur_T = static_cast<R *>(b.T[b.cur()]);
        return KG_OK;
    };
    std::vector<char> built(every, 0);
    job.recompute = [&](int i, R *const *in, R *const *o, uint32_t step) -> int {
        StepArgs<R> *table = static_cast<StepArgs<R> *>(b.adj->tables) + (size_t)i * M;
        if (!built[i]) {
            std::vector<StepArgs<R>> v(M);
            for (int m = 0; m < M; ++m) {
                StepArgs<R> a = make_args<R>(b.params[m], b.nx[m], b.ny[m]);
                a.wide_angles = wide_angles(b.params[m], b.theta_given[m]);
                a.phi_in = in[0] + b.off[m];
                a.T_in = in[1] + b.off[m];
                a.phi_out = o[0] + b.off[m];
                a.T_out = o[1] + b.off[m];
                a.theta = o[2] + b.off[m];
                v[m] = a;
            }
            KG_HIP(hipMemcpy(table, v.data(), M * sizeof(StepArgs<R>), hipMemcpyHostToDevice));
            built[i] = 1;
        }
        KG_HIP(launch_step_batch_direct<R>(table, b.tiles, b.ntiles, step, b.stream));
        return KG_OK;
    };
    std::vector<R> tgt(b.total, (R)0), dT(d_T0 ? b.total : 0);
    for (int m = 0; m < M; ++m)
        std::memcpy(tgt.data() + b.off[m], targets[m], b.cells(m) * sizeof(R));
    std::vector<kg_grads> g;
    const int rc0 = adjoint_core<R>(job, b.adj, nsteps, every, tgt.data(), g,
                                    d_T0 ? dT.data() : nullptr, nullptr, true);
    b.stat_valid = false;  // the core may restore states into the batch buffers
    if (rc0)
        return rc0;
    for (int m = 0; m < M; ++m) {
        out[m] = g[m];
        if (d_T0 && d_T0[m])
            std::memcpy(d_T0[m], dT.data() + b.off[m], b.cells(m) * sizeof(R));
    }
    return KG_OK;
}

}  // namespace

}  // namespace kg

extern "C" int kg_batch_adjoint(kg_batch b, int nsteps, int checkpoint_every,
                                const void *const *targets, kg_grads *out, void *const *d_T0)
{
    if (!b)
        return kg::fail(KG_ERR_INVALID_ARGUMENT, "null batch");
    if (!targets || !out)
        return kg::fail(KG_ERR_INVALID_ARGUMENT, "null targets or output");
    for (int m = 0; m < b->count; ++m)
        if (!targets[m])
            return kg::fail(KG_ERR_INVALID_ARGUMENT, "null target of a member");
    if (nsteps < 1 || checkpoint_every < 1)
        return kg::fail(KG_ERR_INVALID_ARGUMENT, "nsteps and checkpoint_every must be >= 1");
    KG_HIP(hipSetDevice(b->device));
    const int every = std::min(checkpoint_every, nsteps);
    return b->precision == KG_F64
               ? kg::adjoint_batch<double>(*b, nsteps, every, targets, out, d_T0)
               : kg::adjoint_batch<float>(*b, nsteps, every, targets, out, d_T0);
}

extern "C" int kg_adjoint(kg_handle h, int nsteps, int checkpoint_every, const void *target,
                          size_t bytes, kg_grads *out, void *d_T0, void *d_phi0)
{
    if (!h)
        return kg::fail(KG_ERR_INVALID_ARGUMENT, "null handle");
    if (!target || !out)
        return kg::fail(KG_ERR_INVALID_ARGUMENT, "null target or output");
    if (nsteps < 1 || checkpoint_every < 1)
        return kg::fail(KG_ERR_INVALID_ARGUMENT, "nsteps and checkpoint_every must be >= 1");
    if (bytes != h->field_bytes())
        return kg::fail(KG_ERR_SIZE, "target size does not match the field size");
    KG_HIP(hipSetDevice(h->device));
    const int every = std::min(checkpoint_every, nsteps);
    return h->precision == KG_F64
               ? kg::adjoint_single<double>(*h, nsteps, every, target, out, d_T0, d_phi0)
               : kg::adjoint_single<float>(*h, nsteps, every, target, out, d_T0, d_phi0);
}
