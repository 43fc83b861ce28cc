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

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kg_host.h"
#include "kg_math.h"
#include "kg_philox.h"

namespace kg {

namespace {

constexpr int ADJ_BLOCKS = 1024;
constexpr int ADJ_THREADS = 256;
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
__device__ __forceinline__ void block_add(double (&v)[NPAR], double *acc)
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
        acc[(size_t)threadIdx.x * ADJ_BLOCKS + blockIdx.x] += sh[threadIdx.x][0];
}

// A1: pass-1 quantities of the forward step (same arithmetic as the kernels)
template <typename R, int BC>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_pass1(StepArgs<R> a, AdjIO<R> io,
                                                           AdjTemps<R> t)
{
    const int nx = a.nx, ny = a.ny;
    const size_t n = (size_t)nx * ny;
    for (size_t c = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x; c < n;
         c += (size_t)ADJ_BLOCKS * ADJ_THREADS) {
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
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_pass2(StepArgs<R> a, AdjIO<R> io,
                                                           AdjTemps<R> t, R tau, double *acc)
{
    const int nx = a.nx, ny = a.ny;
    const size_t n = (size_t)nx * ny;
    double v[NPAR] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t c = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x; c < n;
         c += (size_t)ADJ_BLOCKS * ADJ_THREADS) {
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
        t.lphi_loc[c] = lp1 - a.K * lT1 + ls * (((R)1.0 - (R)2.0 * phi) * q + w1);
        t.lT_loc[c] = lT1 - ls * w1 * a.alpha_pi * a.gamma / d;
    }
    block_add<R>(v, acc);
}

// A3: reverse of the pass-1 quantities at each cell (inverse neighbour maps)
template <typename R, int BC>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_cell(StepArgs<R> a, AdjIO<R> io,
                                                          AdjTemps<R> t, double *acc)
{
    const int nx = a.nx, ny = a.ny;
    const size_t n = (size_t)nx * ny;
    double v[NPAR] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t c = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x; c < n;
         c += (size_t)ADJ_BLOCKS * ADJ_THREADS) {
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
    block_add<R>(v, acc);
}

// A4: gather of the input adjoints
template <typename R, int BC>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_gather(StepArgs<R> a, AdjIO<R> io,
                                                            AdjTemps<R> t)
{
    const int nx = a.nx, ny = a.ny;
    const size_t n = (size_t)nx * ny;
    for (size_t c = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x; c < n;
         c += (size_t)ADJ_BLOCKS * ADJ_THREADS) {
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

// loss = sum (phi - target)^2 in FP64; lphi = 2 (phi - target), lT = ltheta = 0
template <typename R>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_seed(const R *phi, const R *target, size_t n,
                                                          R *lphi, R *lT, R *lth, double *partials)
{
    double s = 0.0;
    for (size_t c = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x; c < n;
         c += (size_t)ADJ_BLOCKS * ADJ_THREADS) {
        const R d = phi[c] - target[c];
        s += (double)d * (double)d;
        lphi[c] = (R)2.0 * d;
        lT[c] = (R)0;
        lth[c] = (R)0;
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

template <typename R, int BC>
hipError_t reverse_step(const StepArgs<R> &a, const AdjIO<R> &io, const AdjTemps<R> &t, R tau,
                        double *acc, hipStream_t s)
{
    const dim3 g(ADJ_BLOCKS), b(ADJ_THREADS);
    hipLaunchKernelGGL((k_adj_pass1<R, BC>), g, b, 0, s, a, io, t);
    hipLaunchKernelGGL((k_adj_pass2<R, BC>), g, b, 0, s, a, io, t, tau, acc);
    hipLaunchKernelGGL((k_adj_cell<R, BC>), g, b, 0, s, a, io, t, acc);
    hipLaunchKernelGGL((k_adj_gather<R, BC>), g, b, 0, s, a, io, t);
    return hipGetLastError();
}

}  // namespace

struct AdjointWork {
    size_t n = 0;
    size_t esize = 0;
    int ckpt_cap = 0, seg_cap = 0;
    char *ckpt = nullptr;    // ckpt_cap states of 3 fields
    char *seg = nullptr;     // seg_cap + 1 states of 3 fields
    char *lam = nullptr;     // 2 x (lphi, lT, ltheta)
    char *tmp = nullptr;     // NTEMP fields + target
    signed char *code = nullptr;
    double *acc = nullptr;   // NPAR x ADJ_BLOCKS
    double *partials = nullptr;
};

void free_adjoint(AdjointWork *w)
{
    if (!w)
        return;
    void *p[] = {w->ckpt, w->seg, w->lam, w->tmp, w->code, w->acc, w->partials};
    for (void *q : p)
        if (q)
            (void)hipFree(q);
    delete w;
}

namespace {

int ensure_work(kg_solver_s &h, int nckpt, int seglen)
{
    AdjointWork *w = h.adj;
    const size_t n = h.cells(), es = h.esize(), fb = n * es;
    if (w && (w->ckpt_cap < nckpt || w->seg_cap < seglen)) {
        free_adjoint(w);
        w = h.adj = nullptr;
    }
    if (w)
        return KG_OK;
    w = new AdjointWork();
    w->n = n;
    w->esize = es;
    w->ckpt_cap = nckpt;
    w->seg_cap = seglen;
    bool ok = hipMalloc(&w->ckpt, 3 * fb * (size_t)nckpt) == hipSuccess &&
              hipMalloc(&w->seg, 3 * fb * (size_t)(seglen + 1)) == hipSuccess &&
              hipMalloc(&w->lam, 6 * fb) == hipSuccess &&
              hipMalloc(&w->tmp, (NTEMP + 1) * fb) == hipSuccess &&
              hipMalloc(&w->code, n) == hipSuccess &&
              hipMalloc(&w->acc, (size_t)NPAR * ADJ_BLOCKS * sizeof(double)) == hipSuccess &&
              hipMalloc(&w->partials, ADJ_BLOCKS * sizeof(double)) == hipSuccess;
    if (!ok) {
        free_adjoint(w);
        return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the adjoint buffers failed");
    }
    h.adj = w;
    return KG_OK;
}

template <typename R>
int adjoint_impl(kg_solver_s &h, int nsteps, int every, const void *target, kg_grads *out,
                 void *d_T0, void *d_phi0)
{
    const size_t n = h.cells(), fb = h.field_bytes();
    const int nseg = (nsteps + every - 1) / every;
    if (int rc = ensure_work(h, nseg, every))
        return rc;
    AdjointWork &w = *h.adj;
    const hipStream_t s = h.stream;
    auto st = [&](char *base, int k, int f) { return reinterpret_cast<R *>(base + (3 * (size_t)k + f) * fb); };
    R *tgt = reinterpret_cast<R *>(w.tmp + (size_t)NTEMP * fb);
    KG_HIP(hipMemcpyAsync(tgt, target, fb, hipMemcpyHostToDevice, s));
    const uint64_t s0 = h.steps;
    StepArgs<R> base = make_args<R>(h.params, h.nx, h.ny);

    // forward with checkpoints, on the handle's own buffers
    for (int k = 0; k < nseg; ++k) {
        KG_HIP(hipMemcpyAsync(st(w.ckpt, k, 0), h.phi[h.cur], fb, hipMemcpyDeviceToDevice, s));
        KG_HIP(hipMemcpyAsync(st(w.ckpt, k, 1), h.T[h.cur], fb, hipMemcpyDeviceToDevice, s));
        KG_HIP(hipMemcpyAsync(st(w.ckpt, k, 2), h.theta, fb, hipMemcpyDeviceToDevice, s));
        const int len = std::min(every, nsteps - k * every);
        for (int i = 0; i < len; ++i) {
            StepArgs<R> a = base;
            a.phi_in = static_cast<const R *>(h.phi[h.cur]);
            a.T_in = static_cast<const R *>(h.T[h.cur]);
            a.phi_out = static_cast<R *>(h.phi[h.cur ^ 1]);
            a.T_out = static_cast<R *>(h.T[h.cur ^ 1]);
            a.theta = static_cast<R *>(h.theta);
            a.step = (uint32_t)h.steps;
            KG_HIP(launch_step_fused<R>(a, s));
            h.cur ^= 1;
            h.steps += 1;
        }
    }

    // loss and the seed adjoint
    R *lam[2][3];
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 3; ++f)
            lam[b][f] = reinterpret_cast<R *>(w.lam + (3 * (size_t)b + f) * fb);
    hipLaunchKernelGGL((k_adj_seed<R>), dim3(ADJ_BLOCKS), dim3(ADJ_THREADS), 0, s,
                       static_cast<const R *>(h.phi[h.cur]), (const R *)tgt, n, lam[0][0],
                       lam[0][1], lam[0][2], w.partials);
    KG_HIP(hipGetLastError());
    KG_HIP(hipMemsetAsync(w.acc, 0, (size_t)NPAR * ADJ_BLOCKS * sizeof(double), s));

    AdjTemps<R> t;
    R *tp = reinterpret_cast<R *>(w.tmp);
    R **fields[] = {&t.gx, &t.gy, &t.lap, &t.lapT, &t.th, &t.eps, &t.epsd, &t.cs, &t.sn,
                    &t.ls, &t.gEx, &t.gEy, &t.lphi_loc, &t.lT_loc, &t.lgx, &t.lgy, &t.lth_old};
    for (int f = 0; f < NTEMP; ++f)
        *fields[f] = tp + (size_t)f * n;
    t.code = w.code;

    int lc = 0;
    for (int k = nseg - 1; k >= 0; --k) {
        const int len = std::min(every, nsteps - k * every);
        for (int f = 0; f < 3; ++f)
            KG_HIP(hipMemcpyAsync(st(w.seg, 0, f), st(w.ckpt, k, f), fb, hipMemcpyDeviceToDevice, s));
        for (int i = 0; i < len; ++i) {  // recompute the segment's states
            KG_HIP(hipMemcpyAsync(st(w.seg, i + 1, 2), st(w.seg, i, 2), fb, hipMemcpyDeviceToDevice, s));
            StepArgs<R> a = base;
            a.phi_in = st(w.seg, i, 0);
            a.T_in = st(w.seg, i, 1);
            a.phi_out = st(w.seg, i + 1, 0);
            a.T_out = st(w.seg, i + 1, 1);
            a.theta = st(w.seg, i + 1, 2);
            a.step = (uint32_t)(s0 + (uint64_t)k * every + i);
            KG_HIP(launch_step_fused<R>(a, s));
        }
        for (int i = len - 1; i >= 0; --i) {
            StepArgs<R> a = base;
            a.step = (uint32_t)(s0 + (uint64_t)k * every + i);
            AdjIO<R> io{st(w.seg, i, 0), st(w.seg, i, 1), st(w.seg, i, 2),
                        lam[lc][0], lam[lc][1], lam[lc][2],
                        lam[lc ^ 1][0], lam[lc ^ 1][1], lam[lc ^ 1][2]};
            const hipError_t e = h.params.boundary == KG_ZERO_FLUX
                                     ? reverse_step<R, 1>(a, io, t, (R)h.params.tau, w.acc, s)
                                     : reverse_step<R, 0>(a, io, t, (R)h.params.tau, w.acc, s);
            if (e != hipSuccess)
                return hip_fail(e, "adjoint step launch");
            lc ^= 1;
        }
    }

    std::vector<double> acc((size_t)NPAR * ADJ_BLOCKS), parts(ADJ_BLOCKS);
    KG_HIP(hipMemcpyAsync(acc.data(), w.acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    KG_HIP(hipMemcpyAsync(parts.data(), w.partials, parts.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    if (d_T0)
        KG_HIP(hipMemcpyAsync(d_T0, lam[lc][1], fb, hipMemcpyDeviceToHost, s));
    if (d_phi0)
        KG_HIP(hipMemcpyAsync(d_phi0, lam[lc][0], fb, hipMemcpyDeviceToHost, s));
    KG_HIP(hipStreamSynchronize(s));
    double g[NPAR];
    for (int p = 0; p < NPAR; ++p) {
        double v = 0.0;
        for (int b = 0; b < ADJ_BLOCKS; ++b)
            v += acc[(size_t)p * ADJ_BLOCKS + b];
        g[p] = v;
    }
    double loss = 0.0;
    for (double v : parts)
        loss += v;
    out->loss = loss;
    out->d_tau = g[P_TAU];
    out->d_eps_bar = g[P_EPS_BAR];
    out->d_K = g[P_K];
    out->d_delta = g[P_DELTA];
    out->d_theta0 = g[P_THETA0];
    out->d_alpha = g[P_ALPHA];
    out->d_gamma = g[P_GAMMA];
    out->d_teq = g[P_TEQ];
    out->d_noise_a = g[P_NOISE];
    return KG_OK;
}

}  // namespace

}  // namespace kg

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
               ? kg::adjoint_impl<double>(*h, nsteps, every, target, out, d_T0, d_phi0)
               : kg::adjoint_impl<float>(*h, nsteps, every, target, out, d_T0, d_phi0);
}
