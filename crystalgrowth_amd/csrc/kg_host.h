// kg_host.h -- host-side helpers shared by the C ABI sources: error state,
// HIP status checks, and the StepArgs a kernel launch needs.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <cmath>
#include <vector>

#include "../../include/kobayashi_gpu.h"
#include "kg_internal.h"

namespace kg {
struct AdjointWork;
void free_adjoint(AdjointWork *w);
}  // namespace kg

// A solver handle (kg_handle): one grid's device fields and parameters.
struct kg_solver_s {
    int nx = 0, ny = 0, precision = KG_F32;
    kg_params params{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *phi[2] = {nullptr, nullptr};
    void *T[2] = {nullptr, nullptr};
    void *theta = nullptr;
    void *temps = nullptr;  // 6 fields, two-pass kernel only
    double *partials = nullptr;
    kg::ReduceOut *red = nullptr;
    int *disks = nullptr;
    int disks_cap = 0;
    int cur = 0;
    uint64_t steps = 0;
    int kernel = KG_KERNEL_FUSED;
    kg::AdjointWork *adj = nullptr;  // adjoint buffers, allocated on first use
    // tile status of the fused kernel: 2 x tiles words (one set per buffer
    // index); stat_valid false = unknown, cleared before the next fused step
    uint32_t *stat = nullptr;
    int stat_tiles = 0, tiles_x = 0, tiles_y = 0;
    bool stat_ok = false, stat_valid = false;
    // chunks the fused step works on (kg_api.hip, skip_cycle): the list in
    // use (n < 0: all), and SKIP_AGE pinned host buffers of liveness flags,
    // each with the event of the kernel that writes it, the tag that request
    // writes and whether it is still to be read
    kg::LiveList live{-1, nullptr, {}};
    std::vector<uint16_t> live_order;
    int nchunks = 0;
    uint8_t *live_host[kg::SKIP_AGE] = {};
    hipEvent_t live_ev[kg::SKIP_AGE] = {};
    uint32_t live_tag[kg::SKIP_AGE] = {};
    bool live_pending[kg::SKIP_AGE] = {};
    uint32_t next_tag = 0;
    // lists longer than LIVE_MAX: two device copies used in turn, each
    // filled from a pinned staging buffer by an asynchronous copy whose
    // event guards the staging buffer's next use
    uint16_t *live_dev[2] = {};
    uint16_t *live_stage[2] = {};
    hipEvent_t live_copy_ev[2] = {};
    int live_dev_next = 0;
    uint64_t live_stale = 0;       // answers found not yet written (full launches instead)
    uint64_t listed_launches = 0;  // step launches that covered only listed chunks
    // bound on |theta| of angles not written by the step (resets write 0,
    // kg_set_fields what it is given; infinite after kg_device_fields)
    double theta_given = 0.0;

    size_t cells() const { return (size_t)nx * (size_t)ny; }
    size_t esize() const { return precision == KG_F64 ? sizeof(double) : sizeof(float); }
    size_t field_bytes() const { return cells() * esize(); }
};

// A batch handle (kg_batch): every member's fields in one allocation per field.
struct kg_batch_s {
    int count = 0, precision = KG_F32;
    std::vector<int> nx, ny;
    std::vector<kg_params> params;
    std::vector<size_t> off;   // element offset of each member in the field buffers
    size_t total = 0;          // elements per field buffer
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *phi[2] = {nullptr, nullptr};
    void *T[2] = {nullptr, nullptr};
    void *theta = nullptr;
    void *members = nullptr;   // StepArgs<R>[count]
    int4 *tiles = nullptr;
    int ntiles = 0;
    uint32_t *step_base = nullptr;
    double *partials = nullptr;
    kg::ReduceOut *red = nullptr;
    int *disks = nullptr;
    int disks_cap = 0;
    uint64_t steps = 0;        // steps since the last reset of all members
    uint64_t base_host = 0;    // value of *step_base
    uint32_t base_tmp = 0;     // host source of step_base uploads
    int mode = KG_BATCH_LAUNCH, graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    kg::AdjointWork *adj = nullptr;  // batched adjoint buffers, allocated on first use
    // tile status of the step kernel: two sets (buffer 0, buffer 1) of one
    // word per tile of every member; stat_valid false = unknown
    uint32_t *stat = nullptr;
    int stat_tiles = 0;
    bool stat_valid = false;
    std::vector<double> theta_given;  // per member, as kg_solver_s::theta_given
    // chunk lists (launch mode, batches of at least SKIP_MIN_CHUNKS chunks;
    // kg_batch.hip, batch_skip_cycle): the chunks in the order the tile
    // table deals them (member, chunk column, chunk row), the pinned flags
    // of the request in flight, and the tile table of the listed chunks on
    // the device with its pinned staging copy; ntiles_listed < 0: the full
    // table
    int cw = 0, ch = 0;        // chunk shape (tiles) of the tile table and the lists
    std::vector<int4> chunk_list;
    int4 *chunks_dev = nullptr;
    int nchunks = 0;
    uint8_t *live_flags = nullptr;
    hipEvent_t live_ev = nullptr;
    bool live_pending = false;
    uint32_t live_tag = 0, next_tag = 0;
    int4 *tiles_listed = nullptr, *tiles_stage = nullptr;
    size_t listed_cap = 0;
    int ntiles_listed = -1;
    // the same lists kept on the device in KG_BATCH_GRAPH mode (kg_step.hip,
    // k_batch_compact): two flag buffers, the listed length, lists made
    uint8_t *dflags2 = nullptr;
    int *dcount = nullptr;
    uint32_t *dperiods = nullptr;
    uint64_t listed_launches = 0, live_stale = 0;

    size_t esize() const { return precision == KG_F64 ? sizeof(double) : sizeof(float); }
    size_t cells(int m) const { return (size_t)nx[m] * (size_t)ny[m]; }
    int cur() const { return (int)(steps & 1u); }
    char *field(void *base, int m) const { return static_cast<char *>(base) + off[m] * esize(); }
};

namespace kg {

int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define KG_HIP(call)                               \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess)                      \
            return ::kg::hip_fail(e_, #call);      \
    } while (0)

// Scalar parameters combined in R, in the reference's association order
// (see StepArgs); pointers and step counter filled in by the caller.
template <typename R>
StepArgs<R> make_args(const kg_params &p, int nx, int ny)
{
    StepArgs<R> a{};
    a.nx = nx;
    a.ny = ny;
    a.boundary = p.boundary;
    const R dx = (R)p.dx, dy = (R)p.dy, dt = (R)p.dt, tau = (R)p.tau;
    a.inv_dx = (R)1 / dx;
    a.inv_dy = (R)1 / dy;
    a.inv_3dx2 = (R)1 / ((R)3 * dx * dx);
    a.dt = dt;
    a.dt_tau = dt / tau;
    a.eps_bar = (R)p.eps_bar;
    a.delta = (R)p.delta;
    a.aniso = (R)p.aniso;
    a.theta0 = (R)p.theta0;
    a.neg_ebar_j_delta = -(R)p.eps_bar * (R)p.aniso * (R)p.delta;
    a.alpha_pi = (R)p.alpha / (sizeof(R) == 4 ? (R)3.14159274f : (R)3.14159265358979323846);
    a.gamma = (R)p.gamma;
    a.teq = (R)p.teq;
    a.K = (R)p.K;
    a.noise_a = (R)p.noise_a;
    a.instance = p.instance;
    a.key0 = (uint32_t)p.seed;
    a.key1 = (uint32_t)(p.seed >> 32);
    return a;
}

// StepArgs::stream_out for a launch over `cells` cells of `esize` bytes:
// whether its per-step working set -- phi and T read and written, theta --
// exceeds MI355X's 256 MiB last-level (Infinity) cache, so that its outputs
// cannot be served from cache to the next step anyway.
#ifndef KG_STREAM_OUT_BYTES
#define KG_STREAM_OUT_BYTES (256ull << 20)
#endif
inline int stream_outputs(size_t cells, size_t esize)
{
    return (unsigned long long)cells * esize * 5ull > (unsigned long long)KG_STREAM_OUT_BYTES ? 1 : 0;
}

// |theta| the step itself writes stays below 2 pi + pi/2 < 8.
constexpr double STEP_THETA_BOUND = 8.0;

// StepArgs::wide_angles for parameters p and a bound on the stored |theta|:
// whether |aniso (theta - theta0)| may leave the reduced sin/cos range (with
// a margin for float rounding; NaN or infinite bounds count as wide).
inline int wide_angles(const kg_params &p, double theta_given)
{
    const double b = theta_given > STEP_THETA_BOUND ? theta_given : STEP_THETA_BOUND;
    const double m = std::fabs(p.aniso) * (b + std::fabs(p.theta0));
    return m <= 0.99e4 ? 0 : 1;
}

// largest |v| of n host values of the given precision (NaN -> infinity)
inline double max_abs(const void *v, size_t n, int precision)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double x = precision == KG_F64 ? static_cast<const double *>(v)[i]
                                             : (double)static_cast<const float *>(v)[i];
        const double a = std::fabs(x);
        if (!(a <= m))
            m = a == a ? a : HUGE_VAL;
    }
    return m;
}

// Centres of `count` Philox-drawn nuclei (kg_reset_seeds).
void seed_centres(int nx, int ny, int count, int margin, uint64_t seed, int *cxy);

}  // namespace kg
