// kg_api.hip -- the C ABI of include/kobayashi_gpu.h over the gfx950 kernels.
//
// A handle owns: device fields phi[2], T[2] (ping-pong), theta (in place),
// pass-1 temporaries for the two-pass kernel (allocated on first use), FP64
// reduction scratch, one non-blocking HIP stream at default priority and two
// timing events.  Host code here validates every argument before any launch,
// so a kernel never sees a size or pointer it was not written for.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kobayashi_gpu.h"
#include "kg_host.h"
#include "kg_internal.h"
#include "kg_philox.h"

namespace {
thread_local std::string g_last_error;
constexpr long long kMaxCells = 1ll << 32;  // 4 Gi cells (64 GiB per FP32 field pair)
}  // namespace

namespace kg {

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(KG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void seed_centres(int nx, int ny, int count, int margin, uint64_t seed, int *cxy)
{
    const uint32_t spanx = (uint32_t)(nx - 2 * margin), spany = (uint32_t)(ny - 2 * margin);
    for (int k = 0; k < count; ++k) {
        // counter word 3 = 0xFFFFFFFF keeps these draws apart from the noise
        // counters (word 3 = ensemble member id)
        const u32x4 v = philox4x32_10(u32x4{(uint32_t)k, 0u, 0u, 0xFFFFFFFFu}, (uint32_t)seed,
                                      (uint32_t)(seed >> 32));
        cxy[2 * k] = margin + (int)(v.x % spanx);
        cxy[2 * k + 1] = margin + (int)(v.y % spany);
    }
}

}  // namespace kg

using kg::fail;
using kg::hip_fail;


namespace {

bool finite_pos(double v) { return std::isfinite(v) && v > 0.0; }

int check_handle(kg_handle h)
{
    if (!h)
        return fail(KG_ERR_INVALID_ARGUMENT, "null handle");
    return KG_OK;
}

template <typename R>
kg::StepArgs<R> make_args(const kg_solver_s &h)
{
    kg::StepArgs<R> a = kg::make_args<R>(h.params, h.nx, h.ny);
    a.wide_angles = kg::wide_angles(h.params, h.theta_given);
    // the dense launch streams its outputs past the caches on grids larger
    // than the last-level cache (measured on 4096^2: 54.2 -> 53.4 us per
    // step); the status launches re-read much of what they wrote from that
    // cache (35.4 -> 36.3 us with the hint), so they store normally
    a.stream_out = h.kernel == KG_KERNEL_FUSED_DENSE ? kg::stream_outputs(h.cells(), h.esize()) : 0;
    a.phi_in = static_cast<const R *>(h.phi[h.cur]);
    a.T_in = static_cast<const R *>(h.T[h.cur]);
    a.phi_out = static_cast<R *>(h.phi[h.cur ^ 1]);
    a.T_out = static_cast<R *>(h.T[h.cur ^ 1]);
    a.theta = static_cast<R *>(h.theta);
    if (h.stat_ok && h.kernel == KG_KERNEL_FUSED) {
        a.stat_in = h.stat + (size_t)h.cur * h.stat_tiles;
        a.stat_out = h.stat + (size_t)(h.cur ^ 1) * h.stat_tiles;
        a.tile_base = 0;
        a.tiles_x = h.tiles_x;
        a.tiles_y = h.tiles_y;
    }
    return a;
}

// The tile status no longer describes the buffers (fields written by
// something other than the fused step): it is cleared before the next step.
void invalidate_status(kg_solver_s &h) { h.stat_valid = false; }

int ensure_status(kg_solver_s &h)
{
    if (!h.stat_ok || h.stat_valid || h.kernel != KG_KERNEL_FUSED)
        return KG_OK;
    KG_HIP(kg::launch_fill_u32(h.stat, 2 * (size_t)h.stat_tiles, 0u, h.stream));
    h.stat_valid = true;
    // liveness answers about the old fields are void
    h.live.n = -1;
    for (bool &q : h.live_pending)
        q = false;
    return KG_OK;
}

// Chunks the fused step works on.  At every step that is a multiple of
// SKIP_PERIOD the host (1) takes the liveness answer it asked for SKIP_AGE
// periods ago, if any, as the list of chunks the next SKIP_PERIOD launches
// cover (no answer, an answer not all written yet, or more live chunks than
// a LiveList holds: all chunks), and (2) asks again: the step launch itself
// carries the request (leading blocks read the input status words and write
// the flags, tagged with the request's number, to pinned host memory), and
// an event follows the launch.  A chunk is skipped when every tile within
// SKIP_MARGIN_X tile columns and SKIP_MARGIN_Y tile rows of it (64 and 30
// cells) has the status "phi = +0, T = +0" in the state the request read;
// nonzero values cannot reach the chunk in the (SKIP_AGE + 1) * SKIP_PERIOD
// steps an answer is used for, and both of its buffers hold those zeros
// again before the answer is used (kg_step.hip checks the margin and states
// the argument), so skipping changes no bit.  The event skips the
// system-scope fence (a write-back of the L2 that would cost microseconds):
// the flags do not need it, being written with system-scope stores, and a
// flag not yet visible shows by its tag.  Waiting on an answer bounds how far
// the host runs ahead of the GPU to about SKIP_AGE + 1 periods.
// Sets slot to the buffer the request carried by this step writes.
int skip_cycle(kg_solver_s &h, int &slot)
{
    const int p = (int)((h.steps / kg::SKIP_PERIOD) % kg::SKIP_AGE);
    h.live.n = -1;
    if (h.live_pending[p]) {
        KG_HIP(hipEventSynchronize(h.live_ev[p]));
        h.live_order.clear();
        if (!kg::live_chunk_order(h.nx, h.ny, h.live_host[p], h.live_tag[p], h.live_order))
            ++h.live_stale;
        else if (h.live_order.size() <= (size_t)kg::LIVE_MAX) {
            std::copy(h.live_order.begin(), h.live_order.end(), h.live.c);
            h.live.n = (int)h.live_order.size();
            h.live.dev = nullptr;
        } else if (h.live_dev[0]) {
            // longer than the kernel arguments hold: through device memory.
            // Launches still reading this buffer were queued two periods
            // ago, before the copy; its staging buffer is free once the
            // copy queued then has run.
            const int b = h.live_dev_next;
            h.live_dev_next ^= 1;
            KG_HIP(hipEventSynchronize(h.live_copy_ev[b]));
            std::copy(h.live_order.begin(), h.live_order.end(), h.live_stage[b]);
            KG_HIP(hipMemcpyAsync(h.live_dev[b], h.live_stage[b],
                                  h.live_order.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                                  h.stream));
            KG_HIP(hipEventRecord(h.live_copy_ev[b], h.stream));
            h.live.n = (int)h.live_order.size();
            h.live.dev = reinterpret_cast<const uint32_t *>(h.live_dev[b]);
        }
        h.live_pending[p] = false;
    }
    h.next_tag = h.next_tag % kg::LIVE_TAGS + 1;
    h.live_tag[p] = h.next_tag;
    slot = p;
    return KG_OK;
}

int ensure_temps(kg_solver_s &h)
{
    if (h.temps)
        return KG_OK;
    hipError_t e = hipMalloc(&h.temps, 6 * h.field_bytes());
    if (e != hipSuccess) {
        h.temps = nullptr;
        return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of two-pass temporaries failed");
    }
    return KG_OK;
}

template <typename R>
int step_impl(kg_solver_s &h, int nsteps)
{
    if (h.kernel != KG_KERNEL_FUSED)
        invalidate_status(h);  // the step writes the buffers without keeping the status
    else if (int rc = ensure_status(h))
        return rc;
    const bool skipping = h.kernel == KG_KERNEL_FUSED && h.stat_ok && h.live_host[0] != nullptr;
    for (int s = 0; s < nsteps; ++s) {
        int slot = -1;  // >= 0: this launch carries a liveness request
        if (skipping && h.steps % kg::SKIP_PERIOD == 0)
            if (int rc = skip_cycle(h, slot))
                return rc;
        kg::StepArgs<R> a = make_args<R>(h);
        a.step = (uint32_t)h.steps;
        hipError_t e;
        if (h.kernel == KG_KERNEL_TWO_PASS) {
            R *base = static_cast<R *>(h.temps);
            const size_t n = h.cells();
            kg::PassTemps<R> t{base, base + n, base + 2 * n, base + 3 * n, base + 4 * n,
                               base + 5 * n};
            e = kg::launch_step_two_pass<R>(a, t, h.stream);
        } else {
            const kg::LiveReq req{slot >= 0 ? h.live_host[slot] : nullptr,
                                  slot >= 0 ? h.live_tag[slot] : 0u, 0};
            e = kg::launch_step_fused<R>(a, h.stream, skipping ? &h.live : nullptr,
                                         slot >= 0 ? &req : nullptr);
            if (skipping && h.live.n >= 0 && h.live.n < h.nchunks)
                ++h.listed_launches;
        }
        if (e != hipSuccess)
            return hip_fail(e, "step launch");
        if (slot >= 0) {
            KG_HIP(hipEventRecord(h.live_ev[slot], h.stream));
            h.live_pending[slot] = true;
        }
        h.cur ^= 1;
        h.steps += 1;
    }
    return KG_OK;
}

int init_fields(kg_solver_s &h, int kind, int x, int y, const std::vector<int> &disks, double r2)
{
    KG_HIP(hipSetDevice(h.device));
    const int nd = (int)disks.size() / 2;
    if (nd > h.disks_cap) {
        if (h.disks)
            KG_HIP(hipFree(h.disks));
        h.disks = nullptr;
        KG_HIP(hipMalloc(&h.disks, disks.size() * sizeof(int)));
        h.disks_cap = nd;
    }
    if (nd > 0)
        KG_HIP(hipMemcpyAsync(h.disks, disks.data(), disks.size() * sizeof(int),
                              hipMemcpyHostToDevice, h.stream));
    hipError_t e;
    h.cur = 0;
    h.theta_given = 0.0;
    invalidate_status(h);
    if (h.precision == KG_F64)
        e = kg::launch_init<double>(static_cast<double *>(h.phi[0]), static_cast<double *>(h.T[0]),
                                    static_cast<double *>(h.theta), h.nx, h.ny, kind, x, y, h.disks,
                                    nd, r2, h.stream);
    else
        e = kg::launch_init<float>(static_cast<float *>(h.phi[0]), static_cast<float *>(h.T[0]),
                                   static_cast<float *>(h.theta), h.nx, h.ny, kind, x, y, h.disks,
                                   nd, r2, h.stream);
    if (e != hipSuccess)
        return hip_fail(e, "init launch");
    KG_HIP(hipStreamSynchronize(h.stream));
    h.steps = 0;
    return KG_OK;
}

void free_all(kg_solver_s *h)
{
    if (!h)
        return;
    (void)hipSetDevice(h->device);
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    kg::free_adjoint(h->adj);
    h->adj = nullptr;
    void *bufs[] = {h->phi[0], h->phi[1], h->T[0], h->T[1], h->theta, h->temps, h->partials,
                    h->red, h->disks, h->stat};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    for (int p = 0; p < kg::SKIP_AGE; ++p) {
        if (h->live_host[p])
            (void)hipHostFree(h->live_host[p]);
        if (h->live_ev[p])
            (void)hipEventDestroy(h->live_ev[p]);
    }
    for (int b = 0; b < 2; ++b) {
        if (h->live_dev[b])
            (void)hipFree(h->live_dev[b]);
        if (h->live_stage[b])
            (void)hipHostFree(h->live_stage[b]);
        if (h->live_copy_ev[b])
            (void)hipEventDestroy(h->live_copy_ev[b]);
    }
    if (h->ev0)
        (void)hipEventDestroy(h->ev0);
    if (h->ev1)
        (void)hipEventDestroy(h->ev1);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
}

int copy_field_out(kg_solver_s *h, const void *src, void *dst, size_t bytes)
{
    if (!dst)
        return fail(KG_ERR_INVALID_ARGUMENT, "null destination");
    if (bytes != h->field_bytes())
        return fail(KG_ERR_SIZE, "buffer size " + std::to_string(bytes) + " != field size " +
                                     std::to_string(h->field_bytes()));
    KG_HIP(hipSetDevice(h->device));
    KG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    KG_HIP(hipStreamSynchronize(h->stream));
    return KG_OK;
}

}  // namespace

extern "C" {

int kg_abi_version(void) { return KG_ABI_VERSION; }

const char *kg_last_error(void) { return g_last_error.c_str(); }

int kg_device_count(int *count)
{
    if (!count)
        return fail(KG_ERR_INVALID_ARGUMENT, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        n = 0;
    *count = n;
    return KG_OK;
}

int kg_default_params(kg_params *out)
{
    if (!out)
        return fail(KG_ERR_INVALID_ARGUMENT, "null params");
    std::memset(out, 0, sizeof(*out));
    out->dx = (double)0.03f;
    out->dy = (double)0.03f;
    out->dt = (double)0.0001f;
    out->tau = (double)0.0003f;
    out->eps_bar = (double)0.010f;
    out->mu = (double)1.0f;
    out->K = (double)1.6f;
    out->delta = (double)0.05f;
    out->aniso = (double)6.0f;
    out->alpha = (double)0.9f;
    out->gamma = (double)10.0f;
    out->teq = (double)1.0f;
    out->theta0 = 0.0;
    out->noise_a = 0.0;
    out->boundary = KG_PERIODIC;
    out->instance = 0;
    out->seed = 0;
    return KG_OK;
}

int kg_validate(int nx, int ny, const kg_params *p)
{
    if (!p)
        return fail(KG_ERR_INVALID_ARGUMENT, "null params");
    if (nx < 3 || ny < 3)
        return fail(KG_ERR_INVALID_ARGUMENT, "grid must be at least 3x3");
    if ((long long)nx * (long long)ny > kMaxCells)
        return fail(KG_ERR_INVALID_ARGUMENT, "grid larger than 2^32 cells");
    if (!finite_pos(p->dx) || !finite_pos(p->dy) || !finite_pos(p->dt) || !finite_pos(p->tau) ||
        !finite_pos(p->eps_bar))
        return fail(KG_ERR_INVALID_ARGUMENT, "dx, dy, dt, tau and eps_bar must be finite and > 0");
    const double vals[] = {p->mu, p->K, p->delta, p->aniso, p->alpha, p->gamma,
                           p->teq, p->theta0, p->noise_a};
    for (double v : vals)
        if (!std::isfinite(v))
            return fail(KG_ERR_INVALID_ARGUMENT, "non-finite parameter");
    if (p->boundary != KG_PERIODIC && p->boundary != KG_ZERO_FLUX)
        return fail(KG_ERR_INVALID_ARGUMENT, "boundary must be KG_PERIODIC or KG_ZERO_FLUX");
    // forward Euler with the 9-point stencil: the most negative symbol is
    // -16/(3 dx^2) (checkerboard), stable while dt * 16/(3 dx^2) <= 2
    const double h2 = std::fmin(p->dx, p->dy) * std::fmin(p->dx, p->dy);
    if (p->dt / h2 > 0.375)
        return fail(KG_ERR_UNSTABLE, "dt/dx^2 exceeds 3/8 (explicit-Euler bound of the heat equation)");
    const double d = 1.0 + std::fabs(p->delta);
    if (p->dt * p->eps_bar * p->eps_bar * d * d / (p->tau * h2) > 0.375)
        return fail(KG_ERR_UNSTABLE, "dt eps_bar^2 (1+|delta|)^2 / (tau dx^2) exceeds 3/8");
    return KG_OK;
}

int kg_reference_range_mask(const kg_params *p, uint32_t *mask)
{
    if (!p || !mask)
        return fail(KG_ERR_INVALID_ARGUMENT, "null argument");
    // slider {min, max} of src/Kobayashi.cpp:17-58
    const double v[9] = {p->tau, p->eps_bar, p->mu, p->K, p->delta,
                         p->aniso, p->alpha, p->gamma, p->teq};
    const double lo[9] = {0.0001f, 0.006f, 0.5f, 1.0f, 0.01f, 2.0f, 0.7f, 10.0f, 0.5f};
    const double hi[9] = {0.0009f, 0.015f, 1.4f, 1.9f, 0.09f, 8.0f, 1.2f, 20.0f, 1.5f};
    uint32_t m = 0;
    for (int k = 0; k < 9; ++k)
        if (!(v[k] >= lo[k] - 1e-9 && v[k] <= hi[k] + 1e-9))
            m |= 1u << k;
    *mask = m;
    return KG_OK;
}

int kg_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    if (!ctr || !key || !out)
        return fail(KG_ERR_INVALID_ARGUMENT, "null argument");
    const kg::u32x4 r = kg::philox4x32_10(kg::u32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
    return KG_OK;
}

int kg_create(int nx, int ny, int precision, const kg_params *p, kg_handle *out)
{
    if (!out)
        return fail(KG_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (precision != KG_F32 && precision != KG_F64)
        return fail(KG_ERR_INVALID_ARGUMENT, "precision must be KG_F32 or KG_F64");
    int rc = kg_validate(nx, ny, p);
    if (rc != KG_OK)
        return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(KG_ERR_DEVICE, "no GPU visible to the HIP runtime");
    kg_solver_s *h = new kg_solver_s();
    h->nx = nx;
    h->ny = ny;
    h->precision = precision;
    h->params = *p;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&h->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&h->ev1);
    if (e != hipSuccess) {
        free_all(h);
        return hip_fail(e, "stream/event creation");
    }
    void **bufs[] = {&h->phi[0], &h->phi[1], &h->T[0], &h->T[1], &h->theta};
    for (void **b : bufs) {
        if (hipMalloc(b, h->field_bytes()) != hipSuccess) {
            *b = nullptr;
            free_all(h);
            return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the fields failed");
        }
    }
    if (hipMalloc(&h->partials, kg::reduce_partials_size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&h->red, sizeof(kg::ReduceOut)) != hipSuccess) {
        free_all(h);
        return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of reduction scratch failed");
    }
    {
        const int nt = kg::fused_tiles(nx, ny);
        h->stat_tiles = nt;
        h->tiles_x = (nx + kg::fused_tile_x() - 1) / kg::fused_tile_x();
        h->tiles_y = (ny + kg::fused_tile_y() - 1) / kg::fused_tile_y();
        h->stat_ok = kg::fused_status_ok(nx, ny, p->boundary);
        if (hipMalloc(&h->stat, 2 * (size_t)nt * sizeof(uint32_t)) != hipSuccess) {
            h->stat = nullptr;
            free_all(h);
            return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the tile status failed");
        }
        // chunk lists: chunk indices must fit 16 bits; smaller grids do not
        // fill the GPU and gain less from skipping than the host's waits
        // cost when handles are stepped in turn
        const int nc = kg::fused_chunks_x(nx) * kg::fused_chunks_y(ny);
        if (nc >= kg::SKIP_MIN_CHUNKS && nc <= 65535) {
            h->nchunks = nc;
            bool ok = true;
            for (int q = 0; q < kg::SKIP_AGE && ok; ++q) {
                ok = hipHostMalloc(reinterpret_cast<void **>(&h->live_host[q]), (size_t)nc,
                                   hipHostMallocCoherent) == hipSuccess;
                if (!ok)
                    h->live_host[q] = nullptr;
                else
                    std::memset(h->live_host[q], 0, (size_t)nc);  // tag 0: never written
                ok = ok && hipEventCreateWithFlags(&h->live_ev[q],
                                                   hipEventDisableTiming |
                                                       hipEventDisableSystemFence) == hipSuccess;
            }
            // lists longer than the kernel arguments hold go through device
            // memory (two buffers used in turn, 4-byte padded)
            const size_t lb = ((size_t)nc + 1) / 2 * 4;
            for (int b = 0; b < 2 && ok && nc > kg::LIVE_MAX; ++b) {
                ok = hipMalloc(reinterpret_cast<void **>(&h->live_dev[b]), lb) == hipSuccess;
                if (!ok)
                    h->live_dev[b] = nullptr;
                ok = ok && hipHostMalloc(reinterpret_cast<void **>(&h->live_stage[b]), lb,
                                         hipHostMallocDefault) == hipSuccess;
                if (!ok)
                    h->live_stage[b] = nullptr;
                ok = ok && hipEventCreateWithFlags(&h->live_copy_ev[b], hipEventDisableTiming) ==
                               hipSuccess;
            }
            if (!ok) {
                free_all(h);
                return fail(KG_ERR_OUT_OF_MEMORY, "allocation of the chunk skip state failed");
            }
        }
    }
    rc = kg_reset(h);
    if (rc != KG_OK) {
        free_all(h);
        return rc;
    }
    *out = h;
    return KG_OK;
}

int kg_destroy(kg_handle h)
{
    if (!h)
        return fail(KG_ERR_INVALID_ARGUMENT, "null handle");
    free_all(h);
    return KG_OK;
}

int kg_reset(kg_handle h)
{
    if (int rc = check_handle(h))
        return rc;
    return init_fields(*h, 0, h->nx / 2, h->ny / 2, {}, 0.0);
}

int kg_reset_disk(kg_handle h, int cx, int cy, double r)
{
    if (int rc = check_handle(h))
        return rc;
    if (!(r >= 0.0) || !std::isfinite(r))
        return fail(KG_ERR_INVALID_ARGUMENT, "radius must be finite and >= 0");
    return init_fields(*h, 1, 0, 0, {cx, cy}, r * r);
}

int kg_reset_seeds(kg_handle h, int count, double r, int margin, uint64_t seed, int *cxy)
{
    if (int rc = check_handle(h))
        return rc;
    if (count < 0 || count > 1 << 16)
        return fail(KG_ERR_INVALID_ARGUMENT, "count must be in [0, 65536]");
    if (!(r >= 0.0) || !std::isfinite(r))
        return fail(KG_ERR_INVALID_ARGUMENT, "radius must be finite and >= 0");
    if (margin < 0 || 2 * margin >= h->nx || 2 * margin >= h->ny)
        return fail(KG_ERR_INVALID_ARGUMENT, "margin leaves no room for centres");
    std::vector<int> d(2 * (size_t)count);
    kg::seed_centres(h->nx, h->ny, count, margin, seed, d.data());
    if (cxy)
        std::memcpy(cxy, d.data(), d.size() * sizeof(int));
    return init_fields(*h, 1, 0, 0, d, r * r);
}

int kg_set_params(kg_handle h, const kg_params *p)
{
    if (int rc = check_handle(h))
        return rc;
    if (int rc = kg_validate(h->nx, h->ny, p))
        return rc;
    h->params = *p;
    h->stat_ok = kg::fused_status_ok(h->nx, h->ny, p->boundary);
    invalidate_status(*h);
    return KG_OK;
}

int kg_get_params(kg_handle h, kg_params *p)
{
    if (int rc = check_handle(h))
        return rc;
    if (!p)
        return fail(KG_ERR_INVALID_ARGUMENT, "null params");
    *p = h->params;
    return KG_OK;
}

int kg_get_size(kg_handle h, int *nx, int *ny, int *precision)
{
    if (int rc = check_handle(h))
        return rc;
    if (nx)
        *nx = h->nx;
    if (ny)
        *ny = h->ny;
    if (precision)
        *precision = h->precision;
    return KG_OK;
}

int kg_step(kg_handle h, int nsteps)
{
    if (int rc = check_handle(h))
        return rc;
    if (nsteps < 0)
        return fail(KG_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
    KG_HIP(hipSetDevice(h->device));
    // a captured launch would replay with the step count (noise counter) and
    // buffer parity it was captured with: refuse rather than record it
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    KG_HIP(hipStreamIsCapturing(h->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(KG_ERR_INVALID_ARGUMENT,
                    "kg_step cannot be captured into a graph (each launch carries its step "
                    "count and buffer parity); a batch in KG_BATCH_GRAPH mode replays correctly");
    if (h->kernel == KG_KERNEL_TWO_PASS)
        if (int rc = ensure_temps(*h))
            return rc;
    return h->precision == KG_F64 ? step_impl<double>(*h, nsteps) : step_impl<float>(*h, nsteps);
}

int kg_step_timed(kg_handle h, int nsteps, double *ms)
{
    if (int rc = check_handle(h))
        return rc;
    if (!ms)
        return fail(KG_ERR_INVALID_ARGUMENT, "null ms");
    KG_HIP(hipSetDevice(h->device));
    KG_HIP(hipEventRecord(h->ev0, h->stream));
    if (int rc = kg_step(h, nsteps))
        return rc;
    KG_HIP(hipEventRecord(h->ev1, h->stream));
    // ev1 is the stream's last command: waiting for the stream completes it
    // (and returns sooner after the last launch than waiting on the event)
    KG_HIP(hipStreamSynchronize(h->stream));
    float t = 0.0f;
    KG_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = (double)t;
    return KG_OK;
}

int kg_synchronize(kg_handle h)
{
    if (int rc = check_handle(h))
        return rc;
    KG_HIP(hipSetDevice(h->device));
    KG_HIP(hipStreamSynchronize(h->stream));
    return KG_OK;
}

int kg_get_step_count(kg_handle h, uint64_t *steps)
{
    if (int rc = check_handle(h))
        return rc;
    if (!steps)
        return fail(KG_ERR_INVALID_ARGUMENT, "null steps");
    *steps = h->steps;
    return KG_OK;
}

int kg_get_skip_stats(kg_handle h, uint64_t *listed_launches, uint64_t *stale_answers)
{
    if (int rc = check_handle(h))
        return rc;
    if (!listed_launches || !stale_answers)
        return fail(KG_ERR_INVALID_ARGUMENT, "null output");
    *listed_launches = h->listed_launches;
    *stale_answers = h->live_stale;
    return KG_OK;
}

int kg_set_step_count(kg_handle h, uint64_t steps)
{
    if (int rc = check_handle(h))
        return rc;
    h->steps = steps;
    // the liveness cycle runs on the step count: a list or answer in hand
    // could now be used beyond the steps its margin covers
    h->live.n = -1;
    for (bool &q : h->live_pending)
        q = false;
    return KG_OK;
}

int kg_set_kernel(kg_handle h, int kernel)
{
    if (int rc = check_handle(h))
        return rc;
    if (kernel != KG_KERNEL_FUSED && kernel != KG_KERNEL_TWO_PASS &&
        kernel != KG_KERNEL_FUSED_DENSE)
        return fail(KG_ERR_INVALID_ARGUMENT, "unknown kernel");
    h->kernel = kernel;
    invalidate_status(*h);
    return KG_OK;
}

int kg_get_phase(kg_handle h, void *dst, size_t bytes)
{
    if (int rc = check_handle(h))
        return rc;
    return copy_field_out(h, h->phi[h->cur], dst, bytes);
}

int kg_get_temperature(kg_handle h, void *dst, size_t bytes)
{
    if (int rc = check_handle(h))
        return rc;
    return copy_field_out(h, h->T[h->cur], dst, bytes);
}

int kg_get_angle(kg_handle h, void *dst, size_t bytes)
{
    if (int rc = check_handle(h))
        return rc;
    return copy_field_out(h, h->theta, dst, bytes);
}

int kg_set_fields(kg_handle h, const void *phi, const void *T, const void *theta, size_t bytes)
{
    if (int rc = check_handle(h))
        return rc;
    if (!phi || !T)
        return fail(KG_ERR_INVALID_ARGUMENT, "null phi or T");
    if (bytes != h->field_bytes())
        return fail(KG_ERR_SIZE, "buffer size does not match the field size");
    KG_HIP(hipSetDevice(h->device));
    invalidate_status(*h);
    KG_HIP(hipMemcpyAsync(h->phi[h->cur], phi, bytes, hipMemcpyHostToDevice, h->stream));
    KG_HIP(hipMemcpyAsync(h->T[h->cur], T, bytes, hipMemcpyHostToDevice, h->stream));
    h->theta_given = theta ? kg::max_abs(theta, h->cells(), h->precision) : 0.0;
    if (theta)
        KG_HIP(hipMemcpyAsync(h->theta, theta, bytes, hipMemcpyHostToDevice, h->stream));
    else
        KG_HIP(hipMemsetAsync(h->theta, 0, bytes, h->stream));
    KG_HIP(hipStreamSynchronize(h->stream));
    return KG_OK;
}

int kg_device_fields(kg_handle h, void **phi, void **T, void **theta)
{
    if (int rc = check_handle(h))
        return rc;
    invalidate_status(*h);  // the caller may write through these addresses
    h->theta_given = HUGE_VAL;
    if (phi)
        *phi = h->phi[h->cur];
    if (T)
        *T = h->T[h->cur];
    if (theta)
        *theta = h->theta;
    return KG_OK;
}

int kg_get_stream(kg_handle h, void **stream)
{
    if (int rc = check_handle(h))
        return rc;
    if (!stream)
        return fail(KG_ERR_INVALID_ARGUMENT, "null stream");
    *stream = (void *)h->stream;
    return KG_OK;
}

int kg_reduce(kg_handle h, kg_stats *out)
{
    if (int rc = check_handle(h))
        return rc;
    if (!out)
        return fail(KG_ERR_INVALID_ARGUMENT, "null stats");
    KG_HIP(hipSetDevice(h->device));
    hipError_t e;
    if (h->precision == KG_F64)
        e = kg::launch_reduce<double>(static_cast<const double *>(h->phi[h->cur]),
                                      static_cast<const double *>(h->T[h->cur]), h->cells(),
                                      h->params.K, h->partials, h->red, h->stream);
    else
        e = kg::launch_reduce<float>(static_cast<const float *>(h->phi[h->cur]),
                                     static_cast<const float *>(h->T[h->cur]), h->cells(),
                                     (double)(float)h->params.K, h->partials, h->red, h->stream);
    if (e != hipSuccess)
        return hip_fail(e, "reduce launch");
    kg::ReduceOut r{};
    KG_HIP(hipMemcpyAsync(&r, h->red, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    KG_HIP(hipStreamSynchronize(h->stream));
    out->sum_phi = r.sum_phi;
    out->sum_T = r.sum_T;
    out->sum_T_minus_K_phi = r.sum_T_minus_K_phi;
    out->min_phi = r.min_phi;
    out->max_phi = r.max_phi;
    out->solid_fraction = r.above_half / (double)h->cells();
    out->nonfinite = (int64_t)r.nonfinite;
    out->cells = (int64_t)h->cells();
    return KG_OK;
}

}  // extern "C"

namespace {

// kg_colormap / kg_colormap_rgba8: the ramp on the device into a scratch
// buffer, then one copy to the caller's buffer
int colormap_impl(kg_handle h, void *dst, size_t bytes, int order, bool rgba8)
{
    if (int rc = check_handle(h))
        return rc;
    if (!dst)
        return fail(KG_ERR_INVALID_ARGUMENT, "null rgba");
    if (order != 0 && order != 1)
        return fail(KG_ERR_INVALID_ARGUMENT, "order must be 0 or 1");
    if (order == 1 && h->nx != h->ny)
        return fail(KG_ERR_INVALID_ARGUMENT, "the reference's buffer order needs a square grid");
    const size_t per_cell = rgba8 ? 4 : 4 * sizeof(float);
    if (bytes != h->cells() * per_cell)
        return fail(KG_ERR_SIZE, rgba8 ? "rgba buffer must hold 4 bytes per cell"
                                       : "rgba buffer must hold 4 floats per cell");
    KG_HIP(hipSetDevice(h->device));
    void *d = nullptr;
    KG_HIP(hipMalloc(&d, bytes));
    hipError_t e = h->precision == KG_F64
                       ? kg::launch_colormap<double>(static_cast<const double *>(h->phi[h->cur]),
                                                     h->nx, h->ny, order, d, rgba8, h->stream)
                       : kg::launch_colormap<float>(static_cast<const float *>(h->phi[h->cur]),
                                                    h->nx, h->ny, order, d, rgba8, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return hip_fail(e, "colormap");
    return KG_OK;
}

}  // namespace

extern "C" {

int kg_colormap(kg_handle h, float *rgba, size_t bytes, int order)
{
    return colormap_impl(h, rgba, bytes, order, false);
}

int kg_colormap_rgba8(kg_handle h, uint8_t *rgba, size_t bytes, int order)
{
    return colormap_impl(h, rgba, bytes, order, true);
}

}  // extern "C"

// ---- kg_colormap_rgba8_noalloc --------------------------------------------------
namespace {

// Staging of kg_colormap_rgba8_noalloc: the RGBA8 colours of up to
// KG_COLORMAP_NOALLOC_MAX_CELLS cells (16 MiB).  Every GPU the library runs
// on holds its own copy until the process ends.  In host code `scratch` names
// a host-side stand-in, not a device address: the address of each GPU's copy
// is taken once per device with hipGetSymbolAddress and kept in
// g_scratch_ptr, and only that address is ever handed to a kernel or a copy.
static __device__ __align__(256) uint8_t scratch[(size_t)KG_COLORMAP_NOALLOC_MAX_CELLS * 4];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (device `dev` is current).
int scratch_address(int dev, void **out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return fail(KG_ERR_DEVICE, "device index beyond the library's scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return hip_fail(e, "hipGetSymbolAddress of the colour-ramp scratch");
        if (!p)
            return fail(KG_ERR_DEVICE, "hipGetSymbolAddress gave no address for the scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return KG_OK;
}

}  // namespace

extern "C" {

int kg_colormap_rgba8_noalloc(kg_handle h, uint8_t *rgba, size_t bytes, int order)
{
    if (int rc = check_handle(h))
        return rc;
    if (!rgba)
        return fail(KG_ERR_INVALID_ARGUMENT, "null rgba");
    if (order != 0 && order != 1)
        return fail(KG_ERR_INVALID_ARGUMENT, "order must be 0 or 1");
    if (order == 1 && h->nx != h->ny)
        return fail(KG_ERR_INVALID_ARGUMENT, "the reference's buffer order needs a square grid");
    if (bytes != h->cells() * 4)
        return fail(KG_ERR_SIZE, "rgba buffer must hold 4 bytes per cell");
    if (h->cells() > (size_t)KG_COLORMAP_NOALLOC_MAX_CELLS)
        return fail(KG_ERR_SIZE, "grid of " + std::to_string(h->cells()) +
                                     " cells exceeds the no-allocation colour ramp's maximum of " +
                                     std::to_string((long long)KG_COLORMAP_NOALLOC_MAX_CELLS) +
                                     " cells (use kg_colormap_rgba8)");
    KG_HIP(hipSetDevice(h->device));
    void *d = nullptr;
    if (int rc = scratch_address(h->device, &d))
        return rc;
    // every cell's 4 bytes are written by the ramp, so nothing is cleared
    hipError_t e = h->precision == KG_F64
                       ? kg::launch_colormap<double>(static_cast<const double *>(h->phi[h->cur]),
                                                     h->nx, h->ny, order, d, true, h->stream)
                       : kg::launch_colormap<float>(static_cast<const float *>(h->phi[h->cur]),
                                                    h->nx, h->ny, order, d, true, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(rgba, d, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess)
        return hip_fail(e, "colormap (no allocation)");
    return KG_OK;
}

int kg_front_profile(kg_handle h, double cx, double cy, int n, double level, double *radii)
{
    if (int rc = check_handle(h))
        return rc;
    if (!radii || n < 1 || n > (1 << 24))
        return fail(KG_ERR_INVALID_ARGUMENT, "need 1 <= n <= 2^24 and a radii buffer");
    if (!std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(level))
        return fail(KG_ERR_INVALID_ARGUMENT, "non-finite centre or level");
    KG_HIP(hipSetDevice(h->device));
    double *d = nullptr;
    KG_HIP(hipMalloc(&d, (size_t)n * sizeof(double)));
    hipError_t e = h->precision == KG_F64
                       ? kg::launch_front_profile<double>(static_cast<const double *>(h->phi[h->cur]),
                                                          h->nx, h->ny, cx, cy, n, level, d,
                                                          h->stream)
                       : kg::launch_front_profile<float>(static_cast<const float *>(h->phi[h->cur]),
                                                         h->nx, h->ny, cx, cy, n, level, d,
                                                         h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(radii, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return hip_fail(e, "front profile");
    return KG_OK;
}

}  // extern "C"
