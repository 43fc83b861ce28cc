// kg_batch.hip -- batched ensemble of independent Kobayashi grids (C ABI).
//
// All members live in one allocation per field (member offsets rounded up to
// 256 bytes, so interior tiles keep their 16-byte vector loads).  A device
// table holds each member's StepArgs; a tile table lists (member, x0, y0) for
// every 64 x 30 tile of every member, in the block order of batch_tile_table
// (chunks of tiles dealt round-robin to the XCDs).  One launch steps every
// tile of every member; KG_BATCH_GRAPH replays a captured hipGraph of an even
// number of launches followed by a counter kernel, with the noise counter
// kept on the device so replays draw fresh noise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kobayashi_gpu.h"
#include "kg_host.h"
#include "kg_internal.h"

using kg::fail;
using kg::hip_fail;


namespace {

int check_batch(kg_batch b, int member, bool allow_all)
{
    if (!b)
        return fail(KG_ERR_INVALID_ARGUMENT, "null batch");
    if (member >= b->count || (member < 0 && !allow_all))
        return fail(KG_ERR_INVALID_ARGUMENT, "member index out of range");
    return KG_OK;
}

void drop_graph(kg_batch_s *b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_batch(kg_batch_s *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    kg::free_adjoint(b->adj);
    b->adj = nullptr;
    void *bufs[] = {b->phi[0], b->phi[1], b->T[0], b->T[1], b->theta, b->members,
                    b->tiles, b->step_base, b->partials, b->red, b->disks, b->stat,
                    b->chunks_dev, b->tiles_listed, b->dflags2, b->dcount, b->dperiods};
    for (void *p : bufs)
        if (p)
            (void)hipFree(p);
    if (b->live_flags)
        (void)hipHostFree(b->live_flags);
    if (b->tiles_stage)
        (void)hipHostFree(b->tiles_stage);
    if (b->live_ev)
        (void)hipEventDestroy(b->live_ev);
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

// The tile table, in block order: block b steps the tile of entry b.  Every
// member is cut into chunks of cw x ch tiles (2 x 2 unless the batch chooses
// another of kg_step.hip's batch shapes, whose liveness requests check the
// same chunks; chunk rows rotated by 3, as the single-grid kernel orders its
// chunks); the chunks, member after member, are dealt round-robin to the 8
// XCDs (block b runs on XCD b % 8), each XCD walks its chunks one after the
// other, and its sequence is padded to the longest one with empty entries
// (member -1).  So every XCD steps tiles of many members at any time, and
// tiles that share halos meet in one L2.
constexpr int BNX = 8;
const int BSKEW = kg::fused_chunk_skew();

// the chunks of every member in that order: (member, chunk column, chunk row)
std::vector<int4> batch_chunks(const int *nx, const int *ny, int count, int cw, int ch)
{
    const int TX = kg::fused_tile_x(), TY = kg::fused_tile_y();
    std::vector<int4> out;
    for (int m = 0; m < count; ++m) {
        const int tx = (nx[m] + TX - 1) / TX, ty = (ny[m] + TY - 1) / TY;
        const int ncx = (tx + cw - 1) / cw, ncy = (ty + ch - 1) / ch;
        for (int cy = 0; cy < ncy; ++cy)
            for (int c = 0; c < ncx; ++c)
                out.push_back(make_int4(m, (c + BSKEW * cy) % ncx, cy, 0));
    }
    return out;
}

// The tile table of chunks[order[0]], chunks[order[1]], ... (cw x ch tiles
// each): the j-th goes to XCD j % 8.  *real = number of tiles (entries that
// are not padding).
std::vector<int4> deal_chunks(const std::vector<int4> &chunks, const std::vector<int> &order,
                              const int *nx, const int *ny, int cw, int ch, size_t *real)
{
    const int TX = kg::fused_tile_x(), TY = kg::fused_tile_y();
    std::vector<std::vector<int4>> seq(BNX);
    size_t n = 0;
    for (size_t j = 0; j < order.size(); ++j) {
        const int4 c = chunks[(size_t)order[j]];
        const int m = c.x;
        const int tx = (nx[m] + TX - 1) / TX, ty = (ny[m] + TY - 1) / TY;
        for (int k = 0; k < cw * ch; ++k) {
            const int x = c.y * cw + k % cw, y = c.z * ch + k / cw;
            if (x < tx && y < ty) {
                seq[j % BNX].push_back(make_int4(m, x * TX, y * TY, 0));
                ++n;
            }
        }
    }
    size_t len = 0;
    for (const auto &q : seq)
        len = std::max(len, q.size());
    std::vector<int4> table(len * BNX, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < BNX; ++x)
        for (size_t i = 0; i < seq[x].size(); ++i)
            table[i * BNX + x] = seq[x][i];
    *real = n;
    return table;
}

std::vector<int4> batch_tile_table(const int *nx, const int *ny, int count, int cw, int ch,
                                   size_t *real)
{
    const std::vector<int4> chunks = batch_chunks(nx, ny, count, cw, ch);
    std::vector<int> order(chunks.size());
    for (size_t j = 0; j < order.size(); ++j)
        order[j] = (int)j;
    return deal_chunks(chunks, order, nx, ny, cw, ch, real);
}

// What depends on the chunk shape: the tile table, and the chunk lists'
// chunks, flags, event and listed tables (batches of at least
// SKIP_MIN_CHUNKS chunks).  Built beside the batch's own and exchanged with
// them, so a failed build leaves the batch as it was.
struct Tables {
    int cw = 0, ch = 0;
    int4 *tiles = nullptr;
    int ntiles = 0;
    std::vector<int4> chunk_list;
    int4 *chunks_dev = nullptr;
    int nchunks = 0;
    uint8_t *live_flags = nullptr;
    hipEvent_t live_ev = nullptr;
    int4 *tiles_listed = nullptr, *tiles_stage = nullptr;
    size_t listed_cap = 0;
    uint8_t *dflags2 = nullptr;
    int *dcount = nullptr;
    uint32_t *dperiods = nullptr;
};

// bytes of the two device flag buffers, rounded up to whole words
size_t dflags_bytes(int nchunks) { return (2 * (size_t)nchunks + 3) / 4 * 4; }

void free_tables(Tables &t)
{
    for (void *p : {static_cast<void *>(t.tiles), static_cast<void *>(t.chunks_dev),
                    static_cast<void *>(t.tiles_listed), static_cast<void *>(t.dflags2),
                    static_cast<void *>(t.dcount), static_cast<void *>(t.dperiods)})
        if (p)
            (void)hipFree(p);
    if (t.tiles_stage)
        (void)hipHostFree(t.tiles_stage);
    if (t.live_flags)
        (void)hipHostFree(t.live_flags);
    if (t.live_ev)
        (void)hipEventDestroy(t.live_ev);
    t = Tables();
}

void swap_tables(kg_batch_s &b, Tables &t)
{
    std::swap(b.cw, t.cw);
    std::swap(b.ch, t.ch);
    std::swap(b.tiles, t.tiles);
    std::swap(b.ntiles, t.ntiles);
    std::swap(b.chunk_list, t.chunk_list);
    std::swap(b.chunks_dev, t.chunks_dev);
    std::swap(b.nchunks, t.nchunks);
    std::swap(b.live_flags, t.live_flags);
    std::swap(b.live_ev, t.live_ev);
    std::swap(b.tiles_listed, t.tiles_listed);
    std::swap(b.tiles_stage, t.tiles_stage);
    std::swap(b.listed_cap, t.listed_cap);
    std::swap(b.dflags2, t.dflags2);
    std::swap(b.dcount, t.dcount);
    std::swap(b.dperiods, t.dperiods);
}

// Builds the tables of chunk shape cw x ch into t (free_tables(t) after a
// failure).  Uses the null stream: the batch's stream must be idle.
int make_tables(const kg_batch_s &b, int cw, int ch, Tables &t)
{
    t.cw = cw;
    t.ch = ch;
    size_t real = 0;
    const std::vector<int4> tiles =
        batch_tile_table(b.nx.data(), b.ny.data(), b.count, cw, ch, &real);
    if (tiles.size() > (size_t)(1 << 30))
        return fail(KG_ERR_INVALID_ARGUMENT, "too many tiles");
    t.ntiles = (int)tiles.size();
    if (hipMalloc(&t.tiles, tiles.size() * sizeof(int4)) != hipSuccess) {
        t.tiles = nullptr;
        return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the batch tile table failed");
    }
    KG_HIP(hipMemcpy(t.tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
    // chunk lists (up to cw * ch tiles per chunk, every XCD padded to
    // ceil(chunks / 8) chunks)
    t.chunk_list = batch_chunks(b.nx.data(), b.ny.data(), b.count, cw, ch);
    t.nchunks = (int)t.chunk_list.size();
    if (t.nchunks < kg::SKIP_MIN_CHUNKS)
        return KG_OK;
    t.listed_cap = ((size_t)t.nchunks + BNX - 1) / BNX * BNX * cw * ch;
    bool ok = hipMalloc(&t.chunks_dev, t.chunk_list.size() * sizeof(int4)) == hipSuccess;
    if (!ok)
        t.chunks_dev = nullptr;
    ok = ok && hipMalloc(&t.tiles_listed, t.listed_cap * sizeof(int4)) == hipSuccess;
    if (!ok)
        t.tiles_listed = nullptr;
    ok = ok && hipHostMalloc(reinterpret_cast<void **>(&t.tiles_stage),
                             t.listed_cap * sizeof(int4), hipHostMallocDefault) == hipSuccess;
    if (!ok)
        t.tiles_stage = nullptr;
    ok = ok && hipHostMalloc(reinterpret_cast<void **>(&t.live_flags), (size_t)t.nchunks,
                             hipHostMallocCoherent) == hipSuccess;
    if (!ok)
        t.live_flags = nullptr;
    ok = ok && hipEventCreateWithFlags(&t.live_ev, hipEventDisableTiming |
                                                       hipEventDisableSystemFence) == hipSuccess;
    if (!ok)
        t.live_ev = nullptr;
    ok = ok && hipMalloc(&t.dflags2, dflags_bytes(t.nchunks)) == hipSuccess;
    if (!ok)
        t.dflags2 = nullptr;
    ok = ok && hipMalloc(&t.dcount, sizeof(int)) == hipSuccess;
    if (!ok)
        t.dcount = nullptr;
    ok = ok && hipMalloc(&t.dperiods, sizeof(uint32_t)) == hipSuccess;
    if (!ok) {
        t.dperiods = nullptr;
        return fail(KG_ERR_OUT_OF_MEMORY, "allocation of the batch chunk lists failed");
    }
    std::memset(t.live_flags, 0, (size_t)t.nchunks);  // tag 0: never written
    KG_HIP(hipMemcpy(t.chunks_dev, t.chunk_list.data(), t.chunk_list.size() * sizeof(int4),
                     hipMemcpyHostToDevice));
    KG_HIP(hipMemset(t.dflags2, 0, dflags_bytes(t.nchunks)));  // no answer
    KG_HIP(hipMemset(t.dcount, 0xff, sizeof(int)));            // -1: the full table
    KG_HIP(hipMemset(t.dperiods, 0, sizeof(uint32_t)));
    return KG_OK;
}

// The boundary every member has (0 periodic, 1 zero-flux), or 2 when they
// differ: a uniform batch launches the step kernel with one tile body.
int boundary_mode(const kg_batch_s &b)
{
    const int first = b.params[0].boundary;
    for (int m = 1; m < b.count; ++m)
        if (b.params[m].boundary != first)
            return 2;
    return first == 1 ? 1 : 0;
}

// The device lists' arguments (graph mode), or null without them.
bool dev_lists(const kg_batch_s &b, kg::DeviceLists &dl)
{
    if (b.dflags2 == nullptr)
        return false;
    dl = kg::DeviceLists{b.chunks_dev, b.nchunks, b.cw, b.ch, b.dflags2, b.tiles_listed,
                         (int)b.listed_cap, b.dcount, b.dperiods, b.stat, b.stat_tiles};
    return true;
}

// Answers and lists kept on the device are void: flags to tag 0 (never a
// request's), length to -1 (the full table).  Fill kernels, queued on the
// batch's stream.
int clear_dev_lists(kg_batch_s &b)
{
    if (b.dflags2 == nullptr)
        return KG_OK;
    KG_HIP(kg::launch_fill_u32(reinterpret_cast<uint32_t *>(b.dflags2),
                               dflags_bytes(b.nchunks) / 4, 0u, b.stream));
    KG_HIP(kg::launch_fill_u32(reinterpret_cast<uint32_t *>(b.dcount), 1, 0xffffffffu, b.stream));
    return KG_OK;
}

template <typename R>
std::vector<kg::StepArgs<R>> member_args(const kg_batch_s &b)
{
    std::vector<kg::StepArgs<R>> v(b.count);
    const int TX = kg::fused_tile_x(), TY = kg::fused_tile_y();
    int tile_base = 0;
    for (int m = 0; m < b.count; ++m) {
        kg::StepArgs<R> a = kg::make_args<R>(b.params[m], b.nx[m], b.ny[m]);
        a.wide_angles = kg::wide_angles(b.params[m], b.theta_given[m]);
        // status launches re-read much of what they wrote from the cache
        a.stream_out = 0;
        a.phi_in = reinterpret_cast<const R *>(b.field(b.phi[0], m));
        a.T_in = reinterpret_cast<const R *>(b.field(b.T[0], m));
        a.phi_out = reinterpret_cast<R *>(b.field(b.phi[1], m));
        a.T_out = reinterpret_cast<R *>(b.field(b.T[1], m));
        a.theta = reinterpret_cast<R *>(b.field(b.theta, m));
        // status words: set 0 describes buffer 0 (the kernel swaps the sets
        // with the buffers on odd steps); members' tiles are numbered in
        // the order of the tile table
        const int tx = (b.nx[m] + TX - 1) / TX, ty = (b.ny[m] + TY - 1) / TY;
        if (kg::fused_status_ok(b.nx[m], b.ny[m], b.params[m].boundary)) {
            a.stat_in = b.stat;
            a.stat_out = b.stat + b.stat_tiles;
            a.tile_base = tile_base;
            a.tiles_x = tx;
            a.tiles_y = ty;
        }
        tile_base += tx * ty;
        v[m] = a;
    }
    return v;
}

// The tile status no longer describes the buffers: cleared before the next step.
void invalidate_status(kg_batch_s &b) { b.stat_valid = false; }

// Answers and lists about other fields or steps are void.
void drop_lists(kg_batch_s &b)
{
    b.ntiles_listed = -1;
    b.live_pending = false;
}

int ensure_status(kg_batch_s &b)
{
    if (b.stat_valid)
        return KG_OK;
    KG_HIP(kg::launch_fill_u32(b.stat, 2 * (size_t)b.stat_tiles, 0u, b.stream));
    b.stat_valid = true;
    drop_lists(b);
    return clear_dev_lists(b);
}

// The chunk lists of the plain solver (kg_api.hip, skip_cycle) over the
// chunks of every member: every SKIP_PERIOD steps the host takes the answer
// it asked for one period ago as the listed chunks' tile table (live chunks
// by active-tile count, most first, dealt to the XCDs as the full table;
// copied to the device behind the launches that use the previous one) and
// asks again with k_batch_live on the input status set of the next step.
// Each member's chunk keeps its margin inside the member (its boundary map),
// so the plain solver's exactness argument holds member by member.
int batch_skip_cycle(kg_batch_s &b)
{
    b.ntiles_listed = -1;
    if (b.live_pending) {
        KG_HIP(hipEventSynchronize(b.live_ev));
        b.live_pending = false;
        bool ok = true;
        for (int c = 0; c < b.nchunks && ok; ++c)
            ok = (uint32_t)(b.live_flags[c] >> 5) == b.live_tag;
        if (!ok) {
            ++b.live_stale;
        } else {
            std::vector<int> order;
            order.reserve((size_t)b.nchunks);
            for (int na = b.cw * b.ch; na >= 0; --na) {
                const uint32_t want = (uint32_t)na << 1 | 1u;
                for (int c = 0; c < b.nchunks; ++c)
                    if ((b.live_flags[c] & 31u) == want)
                        order.push_back(c);
            }
            if ((int)order.size() < b.nchunks) {
                size_t real = 0;
                const std::vector<int4> t =
                    deal_chunks(b.chunk_list, order, b.nx.data(), b.ny.data(), b.cw, b.ch, &real);
                if (t.size() <= b.listed_cap) {
                    std::copy(t.begin(), t.end(), b.tiles_stage);
                    if (!t.empty())
                        KG_HIP(hipMemcpyAsync(b.tiles_listed, b.tiles_stage,
                                              t.size() * sizeof(int4), hipMemcpyHostToDevice,
                                              b.stream));
                    b.ntiles_listed = (int)t.size();
                }
            }
        }
    }
    b.next_tag = b.next_tag % kg::LIVE_TAGS + 1;
    b.live_tag = b.next_tag;
    const uint32_t *set = b.stat + (size_t)(b.steps & 1u) * b.stat_tiles;
    hipError_t e = b.precision == KG_F64
                       ? kg::launch_batch_live<double>(
                             static_cast<const kg::StepArgs<double> *>(b.members), b.chunks_dev,
                             b.nchunks, set, b.live_tag, b.live_flags, b.cw, b.ch, b.stream)
                       : kg::launch_batch_live<float>(
                             static_cast<const kg::StepArgs<float> *>(b.members), b.chunks_dev,
                             b.nchunks, set, b.live_tag, b.live_flags, b.cw, b.ch, b.stream);
    if (e != hipSuccess)
        return hip_fail(e, "batch liveness launch");
    KG_HIP(hipEventRecord(b.live_ev, b.stream));
    b.live_pending = true;
    return KG_OK;
}

int launch_steps(kg_batch_s &b, int n)
{
    if (n > 0)
        if (int rc = ensure_status(b))
            return rc;
    const bool lists = b.mode == KG_BATCH_LAUNCH && b.live_flags != nullptr;
    if (!lists)
        drop_lists(b);
    // graph modes: the steps between whole replays run the graph's own
    // kernels, so the device lists see every step
    kg::DeviceLists dl;
    const bool dev = b.mode == KG_BATCH_GRAPH_LISTS && dev_lists(b, dl);
    for (int s = 0; s < n; ++s) {
        if (lists && b.steps % kg::SKIP_PERIOD == 0)
            if (int rc = batch_skip_cycle(b))
                return rc;
        const bool listed = b.ntiles_listed >= 0;
        const int4 *tiles = listed ? b.tiles_listed : b.tiles;
        const int nt = listed ? b.ntiles_listed : b.ntiles;
        const uint32_t add = (uint32_t)(b.steps - b.base_host);
        hipError_t e = hipSuccess;
        if (nt > 0 || dev)
            e = b.precision == KG_F64
                    ? kg::launch_step_batch<double>(
                          static_cast<const kg::StepArgs<double> *>(b.members), tiles, nt,
                          b.step_base, add, boundary_mode(b), b.stream, dev ? &dl : nullptr)
                    : kg::launch_step_batch<float>(
                          static_cast<const kg::StepArgs<float> *>(b.members), tiles, nt,
                          b.step_base, add, boundary_mode(b), b.stream, dev ? &dl : nullptr);
        if (e != hipSuccess)
            return hip_fail(e, "batch step launch");
        if (listed)
            ++b.listed_launches;
        b.steps += 1;
    }
    return KG_OK;
}

int capture_graph(kg_batch_s &b)
{
    drop_graph(&b);
    kg::DeviceLists dl;
    const kg::DeviceLists *pdl =
        b.mode == KG_BATCH_GRAPH_LISTS && dev_lists(b, dl) ? &dl : nullptr;
    KG_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int s = 0; s < b.graph_steps && e == hipSuccess; ++s)
        e = b.precision == KG_F64
                ? kg::launch_step_batch<double>(static_cast<const kg::StepArgs<double> *>(b.members),
                                                b.tiles, b.ntiles, b.step_base, (uint32_t)s,
                                                boundary_mode(b), b.stream, pdl)
                : kg::launch_step_batch<float>(static_cast<const kg::StepArgs<float> *>(b.members),
                                               b.tiles, b.ntiles, b.step_base, (uint32_t)s,
                                               boundary_mode(b), b.stream, pdl);
    if (e == hipSuccess)
        e = kg::launch_advance(b.step_base, (uint32_t)b.graph_steps, b.stream);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b.stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        return hip_fail(e != hipSuccess ? e : e2, "graph capture");
    }
    b.graph = g;
    KG_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    return KG_OK;
}

int step_impl(kg_batch_s &b, int nsteps)
{
    if (b.mode != KG_BATCH_LAUNCH && nsteps >= b.graph_steps) {
        if (!b.exec)
            if (int rc = capture_graph(b))
                return rc;
        if (b.base_host != b.steps) {
            b.base_tmp = (uint32_t)b.steps;
            KG_HIP(hipMemcpyAsync(b.step_base, &b.base_tmp, sizeof(uint32_t),
                                  hipMemcpyHostToDevice, b.stream));
            KG_HIP(hipStreamSynchronize(b.stream));
            b.base_host = b.steps;
        }
        if (int rc = ensure_status(b))
            return rc;
        drop_lists(b);  // the host's lists: launch mode only
        while (nsteps >= b.graph_steps) {
            KG_HIP(hipGraphLaunch(b.exec, b.stream));
            b.steps += (uint64_t)b.graph_steps;
            b.base_host += (uint64_t)b.graph_steps;
            nsteps -= b.graph_steps;
        }
    }
    return launch_steps(b, nsteps);
}

int init_member(kg_batch_s &b, int m, int kind, int x, int y, const std::vector<int> &disks,
                double r2)
{
    const int nd = (int)disks.size() / 2;
    if (nd > b.disks_cap) {
        if (b.disks)
            KG_HIP(hipFree(b.disks));
        b.disks = nullptr;
        KG_HIP(hipMalloc(&b.disks, disks.size() * sizeof(int)));
        b.disks_cap = nd;
    }
    if (nd > 0) {
        KG_HIP(hipMemcpyAsync(b.disks, disks.data(), disks.size() * sizeof(int),
                              hipMemcpyHostToDevice, b.stream));
    }
    const int c = b.cur();
    invalidate_status(b);
    hipError_t e;
    if (b.precision == KG_F64)
        e = kg::launch_init<double>(reinterpret_cast<double *>(b.field(b.phi[c], m)),
                                    reinterpret_cast<double *>(b.field(b.T[c], m)),
                                    reinterpret_cast<double *>(b.field(b.theta, m)), b.nx[m],
                                    b.ny[m], kind, x, y, b.disks, nd, r2, b.stream);
    else
        e = kg::launch_init<float>(reinterpret_cast<float *>(b.field(b.phi[c], m)),
                                   reinterpret_cast<float *>(b.field(b.T[c], m)),
                                   reinterpret_cast<float *>(b.field(b.theta, m)), b.nx[m],
                                   b.ny[m], kind, x, y, b.disks, nd, r2, b.stream);
    if (e != hipSuccess)
        return hip_fail(e, "batch init launch");
    // the disk list is reused by the next member: wait before overwriting it
    KG_HIP(hipStreamSynchronize(b.stream));
    return KG_OK;
}

int copy_member_out(kg_batch_s *b, int m, void *base, void *dst, size_t bytes)
{
    if (!dst)
        return fail(KG_ERR_INVALID_ARGUMENT, "null destination");
    if (bytes != b->cells(m) * b->esize())
        return fail(KG_ERR_SIZE, "buffer size does not match the member's field size");
    KG_HIP(hipSetDevice(b->device));
    KG_HIP(hipMemcpyAsync(dst, b->field(base, m), bytes, hipMemcpyDeviceToHost, b->stream));
    KG_HIP(hipStreamSynchronize(b->stream));
    return KG_OK;
}

}  // namespace

extern "C" {

int kg_batch_create(int count, const int *nx, const int *ny, int precision,
                    const kg_params *params, kg_batch *out)
{
    if (!out)
        return fail(KG_ERR_INVALID_ARGUMENT, "null output batch");
    *out = nullptr;
    if (count < 1 || count > (1 << 20) || !nx || !ny || !params)
        return fail(KG_ERR_INVALID_ARGUMENT, "count must be in [1, 2^20] with nx, ny, params");
    if (precision != KG_F32 && precision != KG_F64)
        return fail(KG_ERR_INVALID_ARGUMENT, "precision must be KG_F32 or KG_F64");
    for (int m = 0; m < count; ++m)
        if (int rc = kg_validate(nx[m], ny[m], &params[m]))
            return fail(rc, "member " + std::to_string(m) + ": " + kg_last_error());
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(KG_ERR_DEVICE, "no GPU visible to the HIP runtime");

    kg_batch_s *b = new kg_batch_s();
    b->count = count;
    b->precision = precision;
    b->nx.assign(nx, nx + count);
    b->ny.assign(ny, ny + count);
    b->params.assign(params, params + count);
    b->theta_given.assign(count, 0.0);
    const size_t align = 256 / b->esize();
    size_t total = 0;
    for (int m = 0; m < count; ++m) {
        b->off.push_back(total);
        total += (b->cells(m) + align - 1) / align * align;
    }
    b->total = total;
    size_t real_tiles = 0;
    for (int m = 0; m < count; ++m)
        real_tiles += (size_t)((nx[m] + kg::fused_tile_x() - 1) / kg::fused_tile_x()) *
                      (size_t)((ny[m] + kg::fused_tile_y() - 1) / kg::fused_tile_y());
    if (real_tiles > (size_t)(1 << 30)) {
        delete b;
        return fail(KG_ERR_INVALID_ARGUMENT, "too many tiles");
    }

    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&b->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        free_batch(b);
        return hip_fail(e, "stream/event creation");
    }
    const size_t fb = total * b->esize();
    void **bufs[] = {&b->phi[0], &b->phi[1], &b->T[0], &b->T[1], &b->theta};
    for (void **p : bufs)
        if (hipMalloc(p, fb) != hipSuccess) {
            *p = nullptr;
            free_batch(b);
            return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the batch fields failed");
        }
    const size_t args_bytes = (size_t)count * (precision == KG_F64 ? sizeof(kg::StepArgs<double>)
                                                                    : sizeof(kg::StepArgs<float>));
    b->stat_tiles = (int)real_tiles;
    if (hipMalloc(&b->members, args_bytes) != hipSuccess ||
        hipMalloc(&b->stat, 2 * real_tiles * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&b->step_base, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&b->partials, kg::reduce_partials_size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->red, sizeof(kg::ReduceOut)) != hipSuccess) {
        free_batch(b);
        return fail(KG_ERR_OUT_OF_MEMORY, "hipMalloc of the batch tables failed");
    }
    if (precision == KG_F64) {
        auto v = member_args<double>(*b);
        e = hipMemcpy(b->members, v.data(), args_bytes, hipMemcpyHostToDevice);
    } else {
        auto v = member_args<float>(*b);
        e = hipMemcpy(b->members, v.data(), args_bytes, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemset(b->step_base, 0, sizeof(uint32_t));
    if (e != hipSuccess) {
        free_batch(b);
        return hip_fail(e, "batch table upload");
    }
    // the tile table and the chunk lists' state, in the default chunk shape
    Tables t;
    if (int rc = make_tables(*b, kg::batch_chunk_w(), kg::batch_chunk_h(), t)) {
        free_tables(t);
        free_batch(b);
        return rc;
    }
    swap_tables(*b, t);
    for (int m = 0; m < count; ++m) {
        int rc = init_member(*b, m, 0, nx[m] / 2, ny[m] / 2, {}, 0.0);
        if (rc != KG_OK) {
            free_batch(b);
            return rc;
        }
    }
    *out = b;
    return KG_OK;
}

int kg_batch_destroy(kg_batch b)
{
    if (!b)
        return fail(KG_ERR_INVALID_ARGUMENT, "null batch");
    free_batch(b);
    return KG_OK;
}

int kg_batch_set_params(kg_batch b, int member, const kg_params *p)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    if (int rc = kg_validate(b->nx[member], b->ny[member], p))
        return rc;
    KG_HIP(hipSetDevice(b->device));
    KG_HIP(hipStreamSynchronize(b->stream));
    const int bcm = boundary_mode(*b);
    b->params[member] = *p;
    if (boundary_mode(*b) != bcm)
        drop_graph(b);  // it launches the step kernel compiled for the old boundary mix
    invalidate_status(*b);
    hipError_t e;
    if (b->precision == KG_F64) {
        const auto v = member_args<double>(*b);
        e = hipMemcpy(static_cast<kg::StepArgs<double> *>(b->members) + member, &v[member],
                      sizeof(v[member]), hipMemcpyHostToDevice);
    } else {
        const auto v = member_args<float>(*b);
        e = hipMemcpy(static_cast<kg::StepArgs<float> *>(b->members) + member, &v[member],
                      sizeof(v[member]), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess)
        return hip_fail(e, "member parameter upload");
    return KG_OK;
}

int kg_batch_get_params(kg_batch b, int member, kg_params *p)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    if (!p)
        return fail(KG_ERR_INVALID_ARGUMENT, "null params");
    *p = b->params[member];
    return KG_OK;
}

int kg_batch_count(kg_batch b, int *count)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (!count)
        return fail(KG_ERR_INVALID_ARGUMENT, "null count");
    *count = b->count;
    return KG_OK;
}

int kg_batch_reset_disk(kg_batch b, int member, int cx, int cy, double r)
{
    if (int rc = check_batch(b, member, true))
        return rc;
    if (!(r >= 0.0) || !std::isfinite(r))
        return fail(KG_ERR_INVALID_ARGUMENT, "radius must be finite and >= 0");
    KG_HIP(hipSetDevice(b->device));
    if (member < 0) {
        b->steps = 0;  // every member restarts: counter 0, buffer 0
        for (int m = 0; m < b->count; ++m) {
            const int x = cx < 0 ? b->nx[m] / 2 : cx, y = cx < 0 ? b->ny[m] / 2 : cy;
            if (int rc = init_member(*b, m, 1, 0, 0, {x, y}, r * r))
                return rc;
        }
        return KG_OK;
    }
    return init_member(*b, member, 1, 0, 0,
                       {cx < 0 ? b->nx[member] / 2 : cx, cx < 0 ? b->ny[member] / 2 : cy}, r * r);
}

int kg_batch_reset_seeds(kg_batch b, int member, int count, double r, int margin, uint64_t seed)
{
    if (int rc = check_batch(b, member, true))
        return rc;
    if (count < 0 || count > 1 << 16 || !(r >= 0.0) || !std::isfinite(r) || margin < 0)
        return fail(KG_ERR_INVALID_ARGUMENT, "bad seed count, radius or margin");
    KG_HIP(hipSetDevice(b->device));
    const int lo = member < 0 ? 0 : member, hi = member < 0 ? b->count : member + 1;
    for (int m = lo; m < hi; ++m)
        if (2 * margin >= b->nx[m] || 2 * margin >= b->ny[m])
            return fail(KG_ERR_INVALID_ARGUMENT, "margin leaves no room for centres");
    if (member < 0)
        b->steps = 0;
    for (int m = lo; m < hi; ++m) {
        std::vector<int> d(2 * (size_t)count);
        kg::seed_centres(b->nx[m], b->ny[m], count, margin, seed, d.data());
        if (int rc = init_member(*b, m, 1, 0, 0, d, r * r))
            return rc;
    }
    return KG_OK;
}

int kg_batch_set_fields(kg_batch b, int member, const void *phi, const void *T,
                        const void *theta, size_t bytes)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    if (!phi || !T)
        return fail(KG_ERR_INVALID_ARGUMENT, "null phi or T");
    if (bytes != b->cells(member) * b->esize())
        return fail(KG_ERR_SIZE, "buffer size does not match the member's field size");
    KG_HIP(hipSetDevice(b->device));
    const int c = b->cur();
    invalidate_status(*b);
    KG_HIP(hipMemcpyAsync(b->field(b->phi[c], member), phi, bytes, hipMemcpyHostToDevice,
                          b->stream));
    KG_HIP(hipMemcpyAsync(b->field(b->T[c], member), T, bytes, hipMemcpyHostToDevice, b->stream));
    if (theta)
        KG_HIP(hipMemcpyAsync(b->field(b->theta, member), theta, bytes, hipMemcpyHostToDevice,
                              b->stream));
    else
        KG_HIP(hipMemsetAsync(b->field(b->theta, member), 0, bytes, b->stream));
    KG_HIP(hipStreamSynchronize(b->stream));
    // the member's angle bound decides its sin/cos path: refresh its table entry
    b->theta_given[member] = theta ? kg::max_abs(theta, b->cells(member), b->precision) : 0.0;
    hipError_t e;
    if (b->precision == KG_F64) {
        const auto v = member_args<double>(*b);
        e = hipMemcpy(static_cast<kg::StepArgs<double> *>(b->members) + member, &v[member],
                      sizeof(v[member]), hipMemcpyHostToDevice);
    } else {
        const auto v = member_args<float>(*b);
        e = hipMemcpy(static_cast<kg::StepArgs<float> *>(b->members) + member, &v[member],
                      sizeof(v[member]), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess)
        return hip_fail(e, "member table upload");
    return KG_OK;
}

int kg_batch_get_phase(kg_batch b, int member, void *dst, size_t bytes)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    return copy_member_out(b, member, b->phi[b->cur()], dst, bytes);
}

int kg_batch_get_temperature(kg_batch b, int member, void *dst, size_t bytes)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    return copy_member_out(b, member, b->T[b->cur()], dst, bytes);
}

int kg_batch_get_angle(kg_batch b, int member, void *dst, size_t bytes)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    return copy_member_out(b, member, b->theta, dst, bytes);
}

int kg_batch_front_profile(kg_batch b, int member, double cx, double cy, int n, double level,
                           double *radii)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    if (!radii || n < 1 || n > (1 << 24))
        return fail(KG_ERR_INVALID_ARGUMENT, "need 1 <= n <= 2^24 and a radii buffer");
    if (!std::isfinite(cx) || !std::isfinite(cy) || !std::isfinite(level))
        return fail(KG_ERR_INVALID_ARGUMENT, "non-finite centre or level");
    KG_HIP(hipSetDevice(b->device));
    const int c = b->cur();
    double *d = nullptr;
    KG_HIP(hipMalloc(&d, (size_t)n * sizeof(double)));
    hipError_t e =
        b->precision == KG_F64
            ? kg::launch_front_profile<double>(
                  reinterpret_cast<const double *>(b->field(b->phi[c], member)), b->nx[member],
                  b->ny[member], cx, cy, n, level, d, b->stream)
            : kg::launch_front_profile<float>(
                  reinterpret_cast<const float *>(b->field(b->phi[c], member)), b->nx[member],
                  b->ny[member], cx, cy, n, level, d, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(radii, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return hip_fail(e, "batch front profile");
    return KG_OK;
}

int kg_batch_reduce(kg_batch b, int member, kg_stats *out)
{
    if (int rc = check_batch(b, member, false))
        return rc;
    if (!out)
        return fail(KG_ERR_INVALID_ARGUMENT, "null stats");
    KG_HIP(hipSetDevice(b->device));
    const int c = b->cur();
    const size_t n = b->cells(member);
    hipError_t e;
    if (b->precision == KG_F64)
        e = kg::launch_reduce<double>(reinterpret_cast<const double *>(b->field(b->phi[c], member)),
                                      reinterpret_cast<const double *>(b->field(b->T[c], member)), n,
                                      b->params[member].K, b->partials, b->red, b->stream);
    else
        e = kg::launch_reduce<float>(reinterpret_cast<const float *>(b->field(b->phi[c], member)),
                                     reinterpret_cast<const float *>(b->field(b->T[c], member)), n,
                                     (double)(float)b->params[member].K, b->partials, b->red,
                                     b->stream);
    if (e != hipSuccess)
        return hip_fail(e, "batch reduce launch");
    kg::ReduceOut r{};
    KG_HIP(hipMemcpyAsync(&r, b->red, sizeof(r), hipMemcpyDeviceToHost, b->stream));
    KG_HIP(hipStreamSynchronize(b->stream));
    out->sum_phi = r.sum_phi;
    out->sum_T = r.sum_T;
    out->sum_T_minus_K_phi = r.sum_T_minus_K_phi;
    out->min_phi = r.min_phi;
    out->max_phi = r.max_phi;
    out->solid_fraction = r.above_half / (double)n;
    out->nonfinite = (int64_t)r.nonfinite;
    out->cells = (int64_t)n;
    return KG_OK;
}

int kg_batch_set_mode(kg_batch b, int mode, int graph_steps)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (mode != KG_BATCH_LAUNCH && mode != KG_BATCH_GRAPH && mode != KG_BATCH_GRAPH_LISTS)
        return fail(KG_ERR_INVALID_ARGUMENT,
                    "mode must be KG_BATCH_LAUNCH, KG_BATCH_GRAPH or KG_BATCH_GRAPH_LISTS");
    if (mode != KG_BATCH_LAUNCH && (graph_steps < 2 || (graph_steps & 1) || graph_steps > 4096))
        return fail(KG_ERR_INVALID_ARGUMENT, "graph_steps must be even and in [2, 4096]");
    KG_HIP(hipSetDevice(b->device));
    KG_HIP(hipStreamSynchronize(b->stream));
    if (mode != b->mode || graph_steps != b->graph_steps)
        drop_graph(b);
    if (mode != b->mode)  // steps of the other mode did not keep the device lists
        if (int rc = clear_dev_lists(*b))
            return rc;
    b->mode = mode;
    b->graph_steps = mode != KG_BATCH_LAUNCH ? graph_steps : 0;
    return KG_OK;
}

int kg_batch_get_mode(kg_batch b, int *mode, int *graph_steps)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (mode)
        *mode = b->mode;
    if (graph_steps)
        *graph_steps = b->graph_steps;
    return KG_OK;
}

int kg_batch_set_chunk_shape(kg_batch b, int chunk_w, int chunk_h)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    int shapes[16][2];
    const int n = kg::batch_chunk_shapes(shapes, 16);
    bool known = false;
    for (int i = 0; i < n && i < 16; ++i)
        known = known || (shapes[i][0] == chunk_w && shapes[i][1] == chunk_h);
    if (!known) {
        std::string list;
        for (int i = 0; i < n && i < 16; ++i)
            list += (i ? ", " : "") + std::to_string(shapes[i][0]) + "x" +
                    std::to_string(shapes[i][1]);
        return fail(KG_ERR_INVALID_ARGUMENT, "chunk shape must be one of " + list);
    }
    if (chunk_w == b->cw && chunk_h == b->ch)
        return KG_OK;
    KG_HIP(hipSetDevice(b->device));
    KG_HIP(hipStreamSynchronize(b->stream));
    Tables t;
    if (int rc = make_tables(*b, chunk_w, chunk_h, t)) {
        free_tables(t);
        return rc;
    }
    swap_tables(*b, t);
    free_tables(t);  // the old shape's
    drop_graph(b);   // it launched the old tile table
    drop_lists(*b);
    return KG_OK;
}

int kg_batch_get_chunk_shape(kg_batch b, int *chunk_w, int *chunk_h)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (!chunk_w || !chunk_h)
        return fail(KG_ERR_INVALID_ARGUMENT, "null output");
    *chunk_w = b->cw;
    *chunk_h = b->ch;
    return KG_OK;
}

int kg_batch_step(kg_batch b, int nsteps)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (nsteps < 0)
        return fail(KG_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
    KG_HIP(hipSetDevice(b->device));
    return step_impl(*b, nsteps);
}

int kg_batch_step_timed(kg_batch b, int nsteps, double *ms)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (!ms || nsteps < 0)
        return fail(KG_ERR_INVALID_ARGUMENT, "null ms or negative nsteps");
    KG_HIP(hipSetDevice(b->device));
    // a graph replay may first upload the counter synchronously; do it
    // outside the timed region
    if (b->mode != KG_BATCH_LAUNCH && nsteps >= b->graph_steps && b->base_host != b->steps) {
        b->base_tmp = (uint32_t)b->steps;
        KG_HIP(hipMemcpyAsync(b->step_base, &b->base_tmp, sizeof(uint32_t),
                              hipMemcpyHostToDevice, b->stream));
        KG_HIP(hipStreamSynchronize(b->stream));
        b->base_host = b->steps;
    }
    KG_HIP(hipEventRecord(b->ev0, b->stream));
    if (int rc = step_impl(*b, nsteps))
        return rc;
    KG_HIP(hipEventRecord(b->ev1, b->stream));
    // ev1 is the stream's last command: waiting for the stream completes it
    // (and returns sooner after the last launch than waiting on the event)
    KG_HIP(hipStreamSynchronize(b->stream));
    float t = 0.0f;
    KG_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = (double)t;
    return KG_OK;
}

int kg_batch_synchronize(kg_batch b)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    KG_HIP(hipSetDevice(b->device));
    KG_HIP(hipStreamSynchronize(b->stream));
    return KG_OK;
}

int kg_batch_get_skip_stats(kg_batch b, uint64_t *listed_launches, uint64_t *stale_answers)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (!listed_launches || !stale_answers)
        return fail(KG_ERR_INVALID_ARGUMENT, "null output");
    uint32_t periods = 0;
    if (b->dperiods) {
        KG_HIP(hipSetDevice(b->device));
        KG_HIP(hipStreamSynchronize(b->stream));
        KG_HIP(hipMemcpy(&periods, b->dperiods, sizeof(periods), hipMemcpyDeviceToHost));
    }
    // a list made on the device serves the SKIP_PERIOD launches after it
    *listed_launches = b->listed_launches + (uint64_t)periods * kg::SKIP_PERIOD;
    *stale_answers = b->live_stale;
    return KG_OK;
}

int kg_batch_get_step_count(kg_batch b, uint64_t *steps)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (!steps)
        return fail(KG_ERR_INVALID_ARGUMENT, "null steps");
    *steps = b->steps;
    return KG_OK;
}

int kg_batch_autotune(kg_batch b, int trial_steps, int *mode, int *graph_steps)
{
    if (int rc = check_batch(b, 0, false))
        return rc;
    if (trial_steps < 2)
        return fail(KG_ERR_INVALID_ARGUMENT, "trial_steps must be >= 2");
    KG_HIP(hipSetDevice(b->device));
    KG_HIP(hipStreamSynchronize(b->stream));
    // save the state: the current buffers, theta and the counter
    const size_t fb = b->total * b->esize();
    const int c = b->cur();
    void *save = nullptr;
    if (hipMalloc(&save, 3 * fb) != hipSuccess)
        return fail(KG_ERR_OUT_OF_MEMORY, "autotune needs room for a copy of the state");
    char *s = static_cast<char *>(save);
    hipError_t e = hipMemcpyAsync(s, b->phi[c], fb, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s + fb, b->T[c], fb, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s + 2 * fb, b->theta, fb, hipMemcpyDeviceToDevice, b->stream);
    const uint64_t steps0 = b->steps;
    // every launch form in every chunk shape
    struct Cand { int mode, g, cw, ch; };
    const Cand forms[] = {{KG_BATCH_LAUNCH, 0, 0, 0},      {KG_BATCH_GRAPH, 2, 0, 0},
                          {KG_BATCH_GRAPH, 8, 0, 0},       {KG_BATCH_GRAPH, 32, 0, 0},
                          {KG_BATCH_GRAPH_LISTS, 8, 0, 0}, {KG_BATCH_GRAPH_LISTS, 32, 0, 0}};
    int shapes[16][2];
    const int nshapes = std::min(kg::batch_chunk_shapes(shapes, 16), 16);
    std::vector<Cand> cands;
    for (int k = 0; k < nshapes; ++k)
        for (const Cand &f : forms)
            if (f.mode == KG_BATCH_LAUNCH || f.g <= trial_steps)
                cands.push_back({f.mode, f.g, shapes[k][0], shapes[k][1]});
    double best = 1e300;
    Cand pick = {b->mode, b->graph_steps, b->cw, b->ch};
    for (const Cand &cd : cands) {
        if (e != hipSuccess)
            break;
        int rc = kg_batch_set_chunk_shape(b, cd.cw, cd.ch);
        if (rc == KG_OK)
            rc = kg_batch_set_mode(b, cd.mode, cd.g);
        double ms = 0.0;
        if (rc == KG_OK)
            rc = kg_batch_step(b, cd.g > 0 ? cd.g : 2);  // warm up (capture, caches)
        if (rc == KG_OK)
            rc = kg_batch_step_timed(b, trial_steps, &ms);
        // restore the state before the next candidate
        b->steps = steps0;
        invalidate_status(*b);
        const int cc = b->cur();
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->phi[cc], s, fb, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->T[cc], s + fb, fb, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->theta, s + 2 * fb, fb, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(b->stream);
        if (rc != KG_OK) {
            (void)hipFree(save);
            return rc;
        }
        if (ms < best) {
            best = ms;
            pick = cd;
        }
    }
    (void)hipFree(save);
    if (e != hipSuccess)
        return hip_fail(e, "autotune state copy");
    if (int rc = kg_batch_set_chunk_shape(b, pick.cw, pick.ch))
        return rc;
    if (int rc = kg_batch_set_mode(b, pick.mode, pick.g))
        return rc;
    if (mode)
        *mode = pick.mode;
    if (graph_steps)
        *graph_steps = pick.g;
    return KG_OK;
}

}  // extern "C"
