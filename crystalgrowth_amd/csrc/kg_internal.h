// kg_internal.h -- arguments of the step kernels and the host-side launchers
// shared by kg_step.hip (device code) and kg_api.hip (the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace kg {

// Everything one explicit-Euler step needs, with the reference's scalar
// parameters pre-combined in the field's precision R the way the reference
// combines them (src/Kobayashi.cpp:139-215):
//   gx = (phiE - phiW) / dx            -> (phiE - phiW) * inv_dx
//   lap = (...) / (3 dx dx)            -> (...) * inv_3dx2
//   eps' = ((-eps_bar) j) delta sin()  -> neg_ebar_j_delta * sin()
//   m = alpha / pi * atan(...)         -> alpha_pi * atan(...)
//   phi' = phi + (...) dt / tau        -> phi + (...) * dt_tau
template <typename R>
struct StepArgs {
    const R *phi_in;
    const R *T_in;
    R *phi_out;
    R *T_out;
    R *theta;        // persistent angle, updated in place where a branch fires
    int nx, ny;
    int boundary;    // 0 periodic, 1 zero-flux
    int ntiles;      // tiles of the grid
    R inv_dx, inv_dy, inv_3dx2, dt, dt_tau;
    R eps_bar, delta, aniso, theta0, neg_ebar_j_delta;
    R alpha_pi, gamma, teq, K, noise_a;
    uint32_t step, instance, key0, key1;
    // 1 when |aniso (theta - theta0)| may exceed the reduced sin/cos range
    // for some stored angle (the host tracks a bound on |theta|)
    int wide_angles;
    // 1: the step's output stores carry the non-temporal hint (set for
    // grids whose per-step working set exceeds the last-level cache: the
    // outputs are not read again within the launch, and without the hint
    // they evict the input lines neighbouring tiles still read as halos)
    int stream_out;
    // Tile status words of the fused kernel (one per 64 x 30 tile, bits in
    // kg_step.hip): stat_in describes the input buffers, stat_out the output
    // buffers; the word of the tile at (x0, y0) sits at tile_base +
    // (y0 / 30) * tiles_x + x0 / 64.  stat_in == nullptr: no status (every
    // tile loads and stores all of its fields).
    const uint32_t *stat_in;
    uint32_t *stat_out;
    int tile_base, tiles_x, tiles_y;
};

// Two-pass variant: pass-1 temporaries in global memory, as the reference's
// member arrays (src/Kobayashi.h:109-114).
template <typename R>
struct PassTemps {
    R *gx, *gy, *lap, *lapT, *eps, *epsd;
};

struct ReduceOut {
    double sum_phi, sum_T, sum_T_minus_K_phi, min_phi, max_phi, above_half;
    double nonfinite;
};

// The chunks of 2 x 4 tiles that one launch of the fused step works on,
// when not all: n chunk indices (row-major chunk order, c = cy * chunks_x +
// cx) in the order the launch deals them to the XCDs.  Every chunk left out
// holds phi = +0 and T = +0 on both buffers and keeps them over the launch,
// and its status words already say so (the host derives the list from
// a liveness request; kg_api.hip states the conditions).  Up to LIVE_MAX
// entries travel by value in the kernel arguments; a longer list is copied
// to device memory first and `dev` points at it (the same 16-bit entries).
// Either way the launch has only the listed chunks' blocks.
constexpr int LIVE_MAX = 1020;
struct LiveList {
    int n;
    const uint32_t *dev;              // n > LIVE_MAX: the entries, in device memory
    alignas(4) uint16_t c[LIVE_MAX];  // read by the kernel as 32-bit words
};
static_assert(LIVE_MAX % 2 == 0, "LiveList entries are read in pairs");
// A liveness request carried by a step launch: leading blocks of the launch
// write one flag byte per chunk to `flags` (pinned host memory, one byte per
// chunk, each written with a system-scope store):
//   tag << 5 | na << 1 | lv,
// lv = 1 if some tile within SKIP_MARGIN_X tile columns and SKIP_MARGIN_Y
// tile rows of the chunk (mapped by the boundary) has an input status word
// other than exactly "phi = +0, T = +0"; na = tiles of the chunk whose last
// step took the full path; tag in 1..LIVE_TAGS.  `blocks` is set by
// launch_step_fused.
struct LiveReq {
    uint8_t *flags;
    uint32_t tag;
    int blocks;
};
constexpr uint32_t LIVE_TAGS = 7;
// live == nullptr or live->n < 0: every chunk; req != nullptr (and the tile
// status on): the launch also answers a liveness request
template <typename R>
hipError_t launch_step_fused(const StepArgs<R> &a, hipStream_t s, const LiveList *live = nullptr,
                             const LiveReq *req = nullptr);
// Append to out the chunks flagged live in flags[c] (one byte per chunk, as
// a LiveReq writes it), in the order a LiveList must hold them: by
// their count of active tiles, most first (a launch then ends on cheap
// tiles), the full launch's chunk order within one count.  False, with out
// incomplete, if some byte does not carry `tag` (not yet written by the
// request that tag names).
bool live_chunk_order(int nx, int ny, const uint8_t *flags, uint32_t tag,
                      std::vector<uint16_t> &out);
// Liveness request over the chunks of a batch: flags[c] as a LiveReq writes
// them, for chunk c = (member, chunk column, chunk row) of chunks, from the
// status set stat_set (the input set of the next step).
// Chunks are cw x ch tiles, a shape of batch_chunk_shapes (else
// hipErrorInvalidValue).
template <typename R>
hipError_t launch_batch_live(const StepArgs<R> *members, const int4 *chunks, int nchunks,
                             const uint32_t *stat_set, uint32_t tag, uint8_t *flags, int cw,
                             int ch, hipStream_t s);
// Chunk grid of the fused step for an nx x ny grid (chunks_x * chunks_y chunks).
int fused_chunks_x(int nx);
int fused_chunks_y(int ny);
// Tile rows/columns of margin around a chunk that a liveness request checks,
// and the period (steps) of the host's requests.
constexpr int SKIP_MARGIN_X = 1, SKIP_MARGIN_Y = 1;  // tile columns, tile rows
constexpr int SKIP_PERIOD = 7;
// periods between a liveness request and the first launch that uses it
constexpr int SKIP_AGE = 1;
// grids with fewer chunks launch every chunk: a chunk list saves them at
// most about a microsecond per step, while its waits keep the host within
// two periods of the GPU, so handles stepped in turn would stop overlapping
// on the GPU (256 chunks of 2 x 4 tiles: about 2000 x 2000 cells)
constexpr int SKIP_MIN_CHUNKS = 256;
template <typename R>
hipError_t launch_step_two_pass(const StepArgs<R> &a, const PassTemps<R> &t, hipStream_t s);

// Initial state: all fields zero, then phi = 1 on the reference's plus at
// (x, y) (kind 0) or on `ndisks` disks (kind 1; disks holds cx, cy pairs in
// device memory, r2 the squared radius).
template <typename R>
hipError_t launch_init(R *phi, R *T, R *theta, int nx, int ny, int kind, int x, int y,
                       const int *disks, int ndisks, double r2, hipStream_t s);

// Deterministic FP64 reduction; `partials` holds reduce_partials_size()
// doubles, `out` one ReduceOut (both device memory).
size_t reduce_partials_size();
template <typename R>
hipError_t launch_reduce(const R *phi, const R *T, size_t n, double K, double *partials,
                         ReduceOut *out, hipStream_t s);

// phi -> colour ramp: 4 floats per cell (rgba8 false) or 4 bytes (true)
template <typename R>
hipError_t launch_colormap(const R *phi, int nx, int ny, int order, void *rgba, bool rgba8,
                           hipStream_t s);

// Front distances along n rays from (cx, cy) to phi = level (FP64 results,
// `radii` device memory of n doubles).
template <typename R>
hipError_t launch_front_profile(const R *phi, int nx, int ny, double cx, double cy, int n,
                                double level, double *radii, hipStream_t s);

// Number of 64 x 30 tiles of the fused kernel on an nx x ny grid.
int fused_tiles(int nx, int ny);
// Tile status can be kept for an nx x ny grid with this boundary (every halo
// of a tile lies in its 3 x 3 tile neighbourhood); fill `n` words with v.
bool fused_status_ok(int nx, int ny, int boundary);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
// Output tile size of the fused kernel (members of a batch are cut into it).
int fused_tile_x();
int fused_tile_y();
// Chunk shape of the fused step's chunk lists (tiles) and the rotation of
// successive chunk rows; a batch cuts its members into chunks of one of the
// batch_chunk_shapes (writes up to cap (width, height) pairs, returns how
// many there are), batch_chunk_w x batch_chunk_h unless it chooses another,
// and its liveness requests check the same chunks.
int fused_chunk_w();
int fused_chunk_h();
int fused_chunk_skew();
int batch_chunk_w();
int batch_chunk_h();
int batch_chunk_shapes(int (*out)[2], int cap);

// A batch's chunk lists kept on the device: its chunks (member, chunk
// column, chunk row) of cw x ch tiles, two flag buffers of nchunks bytes,
// the listed table (listed_cap entries), its length (-1: the full table),
// a count of the lists made, and the tile status sets (stat_tiles words
// each).
struct DeviceLists {
    const int4 *chunks;
    int nchunks, cw, ch;
    uint8_t *flags2;
    int4 *listed;
    int listed_cap;
    int *count;
    uint32_t *periods;
    const uint32_t *stat;
    int stat_tiles;
};
// Batched step over a tile table (member, x0, y0); the step counter is
// *step_base + step_add and its parity picks the ping-pong direction.  bcm:
// the boundary every member has (0, 1), or 2 when they differ.  With dl,
// the launch is preceded by the device lists' list and request kernels
// (both return at once except at multiples of SKIP_PERIOD) and uses the
// listed table when there is one.
template <typename R>
hipError_t launch_step_batch(const StepArgs<R> *members, const int4 *tiles, int ntiles,
                             const uint32_t *step_base, uint32_t step_add, int bcm, hipStream_t s,
                             const DeviceLists *dl = nullptr);
// Batched step with the members' pointers as given and an absolute step.
template <typename R>
hipError_t launch_step_batch_direct(const StepArgs<R> *members, const int4 *tiles, int ntiles,
                                    uint32_t step, hipStream_t s);
hipError_t launch_advance(uint32_t *step_base, uint32_t n, hipStream_t s);

}  // namespace kg
