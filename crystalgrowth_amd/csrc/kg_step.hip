// kg_step.hip -- gfx950 kernels of the Kobayashi phase-field step.
//
// One explicit-Euler step = the reference's pass 1 (_computeGradientLaplacian,
// src/Kobayashi.cpp:125-175) and pass 2 (_evolution, :177-221).  The fused
// kernel does both in one launch per step:
//   * a 512-thread block owns a 64 x 30 output tile; phi (halo 2) and T
//     (halo 1) are staged in LDS as rows x0-4 .. x0+67, loaded with aligned
//     16-byte vector loads whenever nx % 4 == 0 (halo segments beyond an
//     edge wrap onto aligned segments or clamp to one edge value);
//   * the pass-1 quantities that pass 2 reads from neighbours -- eps^2,
//     eps eps' gx, eps eps' gy -- are computed once per cell of the tile plus
//     a 1-cell halo (66 x 32 cells) and kept in LDS; atan and sin/cos run
//     only where a branch of :154-167 fires or the kept angle is not +0
//     (cells never fired on take eps(+0), computed once per tile);
//   * each thread then updates a row of 4 consecutive cells with 16-byte
//     LDS reads, one Philox block for their 4 noise samples, and one 16-byte
//     store per field.
// phi and T ping-pong between two buffers; theta is one buffer, written only
// where a branch of :154-167 fires and read only where none does, so blocks
// never race on it.
//
// Memory-bound by design: per cell-update the algorithmic traffic is one
// read and one write of phi and T (16 B in FP32).  Tiles whose phi is exactly
// 0 or exactly 1 everywhere (halo included) reduce to the T diffusion update,
// bit for bit the result of the full arithmetic (every cross term is a signed
// zero); where phi (1 - phi) == 0 the reaction (m(T), noise) is skipped for
// the same reason.
#include <hip/hip_runtime.h>
#include <float.h>

#include <algorithm>
#include <vector>
#include <math.h>

#include "kg_internal.h"
#include "kg_math.h"
#include "kg_philox.h"

namespace kg {

namespace {

#ifndef KG_TILE_Y
#define KG_TILE_Y 30
#endif
constexpr int TX = 64;        // output tile width (x, contiguous)
constexpr int TY = KG_TILE_Y; // output tile height
constexpr int NG = TX / 4;    // 4-cell groups per row
constexpr int NT = NG * (TY + 2);  // threads per block: one pass-1 group each (512: 8 wave64)
constexpr int LW = TX + 8;    // LDS row: columns x0-4 .. x0+67 (16-byte aligned)
constexpr int PH = TY + 4;    // phi rows y0-2 .. y0+TY+1
constexpr int QH = TY + 2;    // T, theta and pass-1 rows y0-1 .. y0+TY
constexpr int NXCD = 8;       // MI355X: workgroups go round-robin over 8 XCDs
static_assert(NG * QH == NT, "pass 1: one 4-cell group per thread");
static_assert(NG * TY <= NT, "pass 2: one 4-cell group per thread");
static_assert(2 * QH <= NT && NT % 64 == 0 && NT <= 1024, "pass 1 halo columns: one thread each");

template <typename R> struct Consts;
template <> struct Consts<float> {
    // PI_F of the reference's viewer library, taken as float(pi)
    static __device__ __forceinline__ float pi() { return 3.14159274f; }
};
template <> struct Consts<double> {
    static __device__ __forceinline__ double pi() { return 3.14159265358979323846; }
};

__device__ __forceinline__ bool is_pos_zero(float v) { return __float_as_uint(v) == 0u; }
__device__ __forceinline__ bool is_pos_zero(double v) { return __double_as_longlong(v) == 0ll; }

// a * b rounded to the type and never fused into a neighbouring add: the
// empty asm takes the product as an opaque register value (the compiler's
// contraction of a * b - c into an fma would otherwise be free to differ
// between unrolled copies of the same expression, and with it the bits)
template <typename R>
__device__ __forceinline__ R mul_rn(R a, R b)
{
    R p = a * b;
    asm volatile("" : "+v"(p));
    return p;
}

// 4 consecutive elements, 16-byte aligned (float) or 2 x 16 bytes (double)
__device__ __forceinline__ void load4(const float *p, float v[4])
{
    const float4 q = *reinterpret_cast<const float4 *>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void load4(const double *p, double v[4])
{
    const double2 a = reinterpret_cast<const double2 *>(p)[0];
    const double2 b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void store4(float *p, const float v[4])
{
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
// Stores of step outputs, which this launch never reads back: with nt
// (StepArgs::stream_out, uniform over the launch) they carry the
// non-temporal hint, so they do not evict input lines that neighbouring
// tiles still read as halos.
typedef float kg_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store4_out(float *p, const float v[4], int nt)
{
    if (nt) {
        const kg_f4 q = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(q, reinterpret_cast<kg_f4 *>(p));
    } else {
        store4(p, v);
    }
}
__device__ __forceinline__ void store4(double *p, const double v[4])
{
    reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store4_out(double *p, const double v[4], int nt)
{
    if (nt) {
        __builtin_nontemporal_store(v[0], p);
        __builtin_nontemporal_store(v[1], p + 1);
        __builtin_nontemporal_store(v[2], p + 2);
        __builtin_nontemporal_store(v[3], p + 3);
    } else {
        store4(p, v);
    }
}

// all four exactly +0 / all four exactly 1 (bit patterns, so -0 is not flat)
__device__ __forceinline__ void flat4(const float v[4], int &all0, int &all1)
{
    const uint32_t a = __float_as_uint(v[0]), b = __float_as_uint(v[1]);
    const uint32_t c = __float_as_uint(v[2]), d = __float_as_uint(v[3]);
    const uint32_t o = a | b | c | d, n = a & b & c & d;
    all0 &= (o == 0u);
    all1 &= (o == 0x3f800000u) & (n == 0x3f800000u);
}
__device__ __forceinline__ void flat4(const double v[4], int &all0, int &all1)
{
    for (int e = 0; e < 4; ++e) {
        all0 &= (__double_as_longlong(v[e]) == 0ll);
        all1 &= (v[e] == 1.0);
    }
}
template <typename R>
__device__ __forceinline__ int finite4(const R v[4])
{
    return (int)isfinite(v[0]) & (int)isfinite(v[1]) & (int)isfinite(v[2]) & (int)isfinite(v[3]);
}

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

// The angle update of src/Kobayashi.cpp:154-167 applied to th (the stored
// angle on entry).  Returns true when a branch fires and rewrote th; false
// leaves th, the stored angle, in force.  NaN
// gradients fire nothing, as in the reference.  atan(gy / gx) is evaluated
// only where a branch that uses it fires, so a wave whose cells all keep
// their angle (flat gradients) or take the axis values skips it.
template <typename R>
__device__ __forceinline__ bool theta_branch(R gx, R gy, R &th)
{
    const R E = (R)1.1920928955078125e-7;  // FLT_EPSILON in both precisions
    const R PI = Consts<R>::pi();
    const bool yneg = gy < -E, ypos = gy > E;
    const bool xneg = gx < -E;
    const bool with_atan = (gx > E && (yneg || ypos)) || xneg;      // :160-167
    const bool on_axis = gx <= E && gx >= -E && (yneg || ypos);     // :154-158
    if (with_atan) {
        const R t = kg_atan_ratio(gy, gx);  // atan(gy / gx)
        th = xneg ? PI + t : (yneg ? (R)2.0 * PI + t : t);
    }
    th = on_axis ? (yneg ? (R)-0.5 * PI : (R)0.5 * PI) : th;  // the two never both hold
    return with_atan || on_axis;
}

// eps (:170) and the factor eps eps' of the cross terms (:171 times eps)
// for angle th
template <typename R>
__device__ __forceinline__ void eps_of_angle(const StepArgs<R> &a, R th, R &eps, R &f)
{
    R sn, cs;
    // wide_angles (uniform): some |j (theta - theta0)| may leave the reduced
    // range, so every call checks; otherwise no per-lane branch
    if (a.wide_angles)
        kg_sincos(a.aniso * (th - a.theta0), &sn, &cs);
    else
        kg_sincos_reduced(a.aniso * (th - a.theta0), &sn, &cs);
    eps = a.eps_bar * ((R)1.0 + a.delta * cs);
    f = eps * (a.neg_ebar_j_delta * sn);
}

// (2 (E + W + N + S) + NE + SW + NW + SE - 12 C), the order of :142-151,
// for centre index c of three rows (m = j-1, z = j, p = j+1)
template <typename R>
__device__ __forceinline__ R lap9(const R *m, const R *z, const R *p, int c)
{
    return (R)2.0 * (z[c + 1] + z[c - 1] + p[c] + m[c]) + p[c + 1] + m[c - 1] + p[c - 1] +
           m[c + 1] - (R)12.0 * z[c];
}

// End of a pass-2 phase: the four values (the phase's results) are taken to
// exist in registers here and LDS reads may not move across, so the
// compiler keeps a phase's arithmetic ahead of the next phase's window reads
// and a phase's windows die before the next are read (no spills at the
// register budget of 6 waves per SIMD).
template <typename R>
__device__ __forceinline__ void phase_end(R (&v)[4])
{
#pragma unroll
    for (int e = 0; e < 4; ++e)
        asm volatile("" : "+v"(v[e]));
}
__device__ __forceinline__ void phase_fence()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// Pass-1 and pass-2 threads map one tile row to 16 consecutive lanes (one
// DPP row).  A thread's left and right neighbour columns are then held by
// the neighbouring lanes: row_shr:1 hands lane l the value of lane l-1,
// row_shl:1 that of lane l+1; the first / last lane of each row keeps `edge`.
// Every lane of the row must be active.
__device__ __forceinline__ float from_left_lane(float edge, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v),
                                                      0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right_lane(float edge, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v),
                                                      0x101, 0xf, 0xf, false));
}

// 6-wide window of the LDS row `rowp` around the 4 cells starting at column
// kc = 4 g + 4 (g = lane within its 16-lane row): w[0] = column kc-1,
// w[1..4] = the cells, w[5] = column kc+4.  FP32 takes the two edge values
// from the neighbouring lanes; the row's first and last lanes use columns 3
// and TX+4, read by every lane of the row at one address (a broadcast), so
// no read conflicts on LDS banks.  FP64 reads them directly.
__device__ __forceinline__ void win6(const float *rowp, int kc, float w[6])
{
    load4(rowp + kc, w + 1);
    const float le = rowp[3], re = rowp[TX + 4];
    w[0] = from_left_lane(le, w[4]);
    w[5] = from_right_lane(re, w[1]);
}
__device__ __forceinline__ void win6(const double *rowp, int kc, double w[6])
{
    w[0] = rowp[kc - 1];
    load4(rowp + kc, w + 1);
    w[5] = rowp[kc + 4];
}

// Order of the tiles over the blocks of the single-grid kernel.  Block b
// runs on XCD b % 8 (workgroups go round-robin over the XCDs) and takes the
// (b / 8)-th tile of that XCD's sequence.  The grid is cut into chunks of
// 2 x 4 tiles; chunks go round-robin over the XCDs in row-major chunk order,
// each chunk row rotated by 3 chunks, and an XCD walks its chunks one after
// the other, the 8 tiles of a chunk in row-major order.  So at any time
// every XCD works on tiles spread over the grid (clustered dendrites load
// the XCDs evenly, and cheap and expensive tiles run side by side), while
// tiles that share halos are processed together on one XCD and meet in its
// L2.  Measured on 20000 steps of a 4096 x 4096 grid growing nine seeds
// (mean of 10 checkpoint steps), all chunks launched: 55.1 us row-major,
// 55.8 us one band per XCD, 46.0 us 2 x 8 chunks unrotated, 44.3 us rotated
// (2 x 4 the same); permuting the chunk order (49.7 - 51.5 us) loses the
// reuse between neighbouring chunks.  With launches of the live chunks only
// (below) the smaller chunk skips more: 40.7 us at 2 x 8, 40.1 us at 2 x 4.
#ifndef KG_CHUNK_W
#define KG_CHUNK_W 2
#endif
#ifndef KG_CHUNK_H
#define KG_CHUNK_H 4
#endif
constexpr int CHUNK_W = KG_CHUNK_W, CHUNK_H = KG_CHUNK_H, CHUNK_SKEW = 3;
// The dense launch (no status, so no chunk lists) deals wider chunks: a
// tile row reads a 16-byte segment on each side of its 64 columns, i.e. a
// 128-byte line of each neighbour, so a chunk's left and right halo lines
// add 2 lines per chunk row to its own 2 per tile; wider chunks share more
// of those lines in one XCD's L2.  One tile row of 8 tiles (512 x 30 cells)
// measured best at 4096^2 (means of two runs, us per launch at steps 2000 /
// 10000 / 19000): 8 x 1 48.3 / 52.0 / 59.6, 8 x 2 48.5 / 52.2 / 59.6,
// 8 x 3 48.9 / 52.3 / 60.2, 8 x 4 48.7 / 53.0 / 60.3, 4 x 2 .. 4 x 4 as
// 8 x 4, 16 x 1 49.1 / 55.5 / 65.7.
#ifndef KG_DENSE_CHUNK_W
#define KG_DENSE_CHUNK_W 8
#endif
#ifndef KG_DENSE_CHUNK_H
#define KG_DENSE_CHUNK_H 1
#endif
constexpr int DENSE_CHUNK_W = KG_DENSE_CHUNK_W, DENSE_CHUNK_H = KG_DENSE_CHUNK_H;

// Tile (tx, ty) of tile block b and its chunk's row-major index c (chunks
// of CW x CH tiles); false for blocks past the grid.
template <int CW = CHUNK_W, int CH = CHUNK_H>
__device__ __forceinline__ bool chunk_tile(int b, int tiles_x, int tiles_y, int &tx, int &ty,
                                           int &c)
{
    const int i = b / NXCD;
    const int chunk = (i / (CW * CH)) * NXCD + b % NXCD;
    const int k = i % (CW * CH);
    const int ncx = (tiles_x + CW - 1) / CW;
    const int cy = chunk / ncx;
    const int cx = (chunk - cy * ncx + CHUNK_SKEW * cy) % ncx;  // a rotation of the row
    tx = cx * CW + k % CW;
    ty = cy * CH + k / CW;
    c = cy * ncx + cx;
    return tx < tiles_x && ty < tiles_y;
}

// blocks of the single-grid launch: whole chunk rounds over the 8 XCDs
int chunk_blocks(int tiles_x, int tiles_y, int cw = CHUNK_W, int ch = CHUNK_H)
{
    const int ncx = (tiles_x + cw - 1) / cw, ncy = (tiles_y + ch - 1) / ch;
    return (ncx * ncy + NXCD - 1) / NXCD * NXCD * cw * ch;
}

// Tile status: one word per tile and buffer, written by the block that
// wrote the tile; a bit is set only when it holds on every cell of the tile.
constexpr uint32_t S_PHI0 = 1u;    // phi is +0 (bit pattern)
constexpr uint32_t S_PHI1 = 2u;    // phi is 1
constexpr uint32_t S_T0 = 4u;      // T is +0 (bit pattern)
constexpr uint32_t S_ACTIVE = 8u;  // the writing step took the full (non-flat) path

// tile coordinate t of a neighbour, mapped like the cells (clamp or wrap)
template <int BC>
__device__ __forceinline__ int map_tile(int t, int n)
{
    if (BC == 1)
        return t < 0 ? 0 : (t >= n ? n - 1 : t);
    return t < 0 ? t + n : (t >= n ? t - n : t);
}

// LDS of one tile: phi (halo 2), T (halo 1) and the pass-1 values that pass 2
// reads from neighbours -- eps^2 and the factor F = eps eps' of the cross
// terms -- each row covering columns x0-4 .. x0+67.  The cross terms need F
// times a neighbour's gx or gy; pass 2 recomputes those gradients from the
// staged phi with the operands and operations pass 1 used, so they round
// alike and only F is kept (37.5 KB per FP32 tile: 4 blocks of 8 waves fit a
// CU's 160 KB).  The old theta of non-flat tiles is staged in F: every pass-1
// cell reads its own theta slot before it writes its own F there.
template <typename R>
struct TileSmem {
    R P[PH * LW];
    R T[QH * LW];
    R E2[QH * LW];
    R F[QH * LW];
    R Z[2];           // eps, eps eps' at theta = +0
    int W[NT / 64];   // per-wave flatness bits
    int W2[NT / 64];  // per-wave status bits of the outputs
};

// LDS rows of width LW all set to v
template <typename R>
__device__ __forceinline__ void fill_rows(R *__restrict__ dst, int rows, R v)
{
    const R q[4] = {v, v, v, v};
    for (int k = threadIdx.x; k < rows * (LW / 4); k += NT)
        store4(&dst[4 * k], q);
}

// Stage `rows` grid rows starting at y (mapped by the boundary) of field f,
// columns x0-4 .. x0+67, into LDS rows of width LW, cell by cell through the
// boundary map (widths nx % 4 != 0; other widths use issue_rows and
// commit_rows below).  CHECK 1 accumulates phi's flatness (all +0 / all 1)
// into f0/f1, CHECK 2 T's finiteness into f0.
template <typename R, int BC, int CHECK>
__device__ __forceinline__ void stage_rows(const R *__restrict__ f, R *__restrict__ dst, int nx,
                                           int ny, int x0, int y, int rows, int &f0, int &f1)
{
    for (int k = threadIdx.x; k < rows * LW; k += NT) {
        const int r = k / LW, c = k - r * LW;
        const R v = f[(size_t)map_index<BC>(y + r, ny) * nx + map_index<BC>(x0 - 4 + c, nx)];
        dst[k] = v;
        if (CHECK == 1) {
            f0 &= is_pos_zero(v);
            f1 &= (v == (R)1.0);
        } else if (CHECK == 2) {
            f0 &= isfinite(v) ? 1 : 0;
        }
    }
}

// Split staging for nx % 4 == 0.  Every 4-column segment then lies wholly
// inside or wholly outside the grid: inside it is one aligned 16-byte load;
// outside, periodic wraps it onto another aligned segment and zero-flux
// repeats one edge value.  A field's ROWS x (LW / 4) segments are at most two
// per thread (segment tid and tid + NT).  issue_rows loads a
// thread's segments into registers and commit_rows writes them to LDS, so a
// tile issues the loads of all of its fields before it waits on any: one
// memory round trip for the tile instead of one per loop trip and field.
template <typename R>
struct Seg2 {
    R a[4], b[4];
    int ea, eb;  // zero-flux edge segment: 0 inside, 1 left (repeat a[0]), 2 right (a[3])
};

// Issue the load of segment k (no use of its data here, so the loads of a
// tile's fields stay in flight together); returns its edge code.
// Element offset of segment k in the field and its edge code.
template <int BC>
__device__ __forceinline__ size_t segment_offset(int nx, int ny, int x0, int y, int k, int &edge)
{
    constexpr int V = LW / 4;
    const int r = k / V, v = k - r * V;
    const int gx = x0 - 4 + 4 * v;
    const bool inside = gx >= 0 && gx < nx;
    // one aligned load in every case: periodic wraps the segment, zero-flux
    // loads the edge segment, whose edge value commit_rows repeats
    const int gl = inside ? gx : (BC == 0 ? map_index<0>(gx, nx) : (gx < 0 ? 0 : nx - 4));
    edge = (BC == 1 && !inside) ? (gx < 0 ? 1 : 2) : 0;
    return (size_t)map_index<BC>(y + r, ny) * nx + gl;
}

template <typename R, int BC>
__device__ __forceinline__ int load_segment(const R *__restrict__ f, int nx, int ny, int x0, int y,
                                            int k, R q[4], bool real = true)
{
    int edge;
    const size_t o = segment_offset<BC>(nx, ny, x0, y, k, edge);
    load4(real ? f + o : f, q);  // !real: an aligned dummy load
    return edge;
}

// real false: the same loads from f's first segment (results unused), for a
// caller that must issue them without a branch
template <typename R, int BC, int ROWS>
__device__ __forceinline__ void issue_rows(const R *__restrict__ f, int nx, int ny, int x0, int y,
                                           Seg2<R> &s, bool real = true)
{
    constexpr int N = ROWS * (LW / 4);
    static_assert(N > NT && N <= 2 * NT, "two segments per thread");
    const int tid = threadIdx.x;
    s.ea = load_segment<R, BC>(f, nx, ny, x0, y, tid, s.a, real);
    // the second segment's load is issued by every thread (no branch around
    // a load separates it from the loads before it); only waves that have
    // second segments compute their addresses, the others load f[0..3]
    size_t ob = 0;
    s.eb = 0;
    if ((tid & ~63) + NT < N)
        ob = segment_offset<BC>(nx, ny, x0, y, tid + NT < N ? tid + NT : N - 1, s.eb);
    load4(real ? f + ob : f, s.b);
}

// The same, with the second segment issued only by waves that have one
// (a wave-uniform branch: the other waves skip its address arithmetic);
// for loads whose wait may cover every load issued so far.
template <typename R, int BC, int ROWS>
__device__ __forceinline__ void issue_rows_w(const R *__restrict__ f, int nx, int ny, int x0, int y,
                                             Seg2<R> &s)
{
    constexpr int N = ROWS * (LW / 4);
    static_assert(N > NT && N <= 2 * NT, "two segments per thread");
    const int tid = threadIdx.x;
    s.ea = load_segment<R, BC>(f, nx, ny, x0, y, tid, s.a);
    s.eb = 0;
    if ((tid & ~63) + NT < N)
        s.eb = load_segment<R, BC>(f, nx, ny, x0, y, tid + NT < N ? tid + NT : N - 1, s.b);
}

template <typename R>
__device__ __forceinline__ void fix_edge(R q[4], int e)
{
    const R v = e == 1 ? q[0] : q[3];
    if (e != 0)
        q[0] = q[1] = q[2] = q[3] = v;
}

// CHECK as in stage_rows
template <typename R, int ROWS, int CHECK>
__device__ __forceinline__ void commit_rows(Seg2<R> &s, R *__restrict__ dst, int &f0, int &f1)
{
    constexpr int V = LW / 4, N = ROWS * V;
    const int tid = threadIdx.x;
    fix_edge(s.a, s.ea);
    store4(&dst[tid / V * LW + 4 * (tid % V)], s.a);
    if (CHECK == 1)
        flat4(s.a, f0, f1);
    else if (CHECK == 2)
        f0 &= finite4(s.a);
    if (tid + NT < N) {
        const int k = tid + NT;
        fix_edge(s.b, s.eb);
        store4(&dst[k / V * LW + 4 * (k % V)], s.b);
        if (CHECK == 1)
            flat4(s.b, f0, f1);
        else if (CHECK == 2)
            f0 &= finite4(s.b);
    }
}

// eps and eps eps' at theta = +0, the angle of every cell no branch has
// fired on yet; computed by the same code as any other angle, so cells that
// take them get the bits the full evaluation would give.
template <typename R>
struct ZeroAngle {
    R eps, f;
};

// pass-1 values of one cell from its phi neighbourhood (rows m, z, p of sP
// around column c) and its old theta.  Returns e^2, F = eps eps' and the
// theta the step leaves (fired: true when a branch rewrote it).
template <typename R>
__device__ __forceinline__ void pass1_cell(const StepArgs<R> &a, const ZeroAngle<R> &z0,
                                           const R *m, const R *z, const R *p, int c,
                                           R th_old, R &e2, R &F, R &th, bool &fired)
{
    const R gx = (z[c + 1] - z[c - 1]) * a.inv_dx;
    const R gy = (p[c] - m[c]) * a.inv_dy;
    th = th_old;  // kept unless a branch fires
    fired = theta_branch(gx, gy, th);
    R eps;
    if (is_pos_zero(th)) {
        eps = z0.eps;
        F = z0.f;
    } else {
        eps_of_angle(a, th, eps, F);
    }
    e2 = eps * eps;
}

// One explicit-Euler step of the 64 x 30 output tile at (x0, y0).  TRACK
// false: the caller passes no tile status (the dense launch), and the status
// code is not compiled in.
template <typename R, int BC, bool TRACK = true>
__device__ __forceinline__ void step_tile(const StepArgs<R> &a, int x0, int y0, TileSmem<R> &sm)
{
    R *__restrict__ sP = sm.P;
    R *__restrict__ sT = sm.T;
    R *__restrict__ sE2 = sm.E2;
    R *__restrict__ sF = sm.F;
    const int nx = a.nx, ny = a.ny;
    const int tid = threadIdx.x;

    // ---- tile status ------------------------------------------------------
    // known: bits that hold on the input buffers over the 3 x 3 tiles around
    // this one (so on the tile and all of its halos); has: bits that already
    // hold on this tile of the output buffers; active: the tile took the full
    // path last step, so theta is staged with phi and T.
    const bool track = TRACK && a.stat_in != nullptr;
    uint32_t known = 0, has = 0;
    bool active = false;
    int sidx = 0;
    // With the status, T is loaded right after the status words, before they
    // are read: every tile but a quiescent one stages T, and the two loads
    // then share one memory round trip.
    const bool spec_T = track && (nx & 3) == 0;
    Seg2<R> gt;
    if (track) {
        const int tx = x0 / TX, ty = y0 / TY;
        sidx = a.tile_base + ty * a.tiles_x + tx;
        // lanes 0..8 of every wave load the 3 x 3 neighbourhood's input
        // words (lane 4 = this tile), lane 9 this tile's output word
        // (lanes past 9 load lane 9's word again: one load, no branch, so
        // the wait below is for it alone and not for the T loads after it)
        const int lane = tid & 63;
        const int ln = lane < 9 ? lane : 4;
        const int yy = map_tile<BC>(ty + ln / 3 - 1, a.tiles_y);
        const int xx = map_tile<BC>(tx + ln % 3 - 1, a.tiles_x);
        const uint32_t wv =
            lane < 9 ? a.stat_in[a.tile_base + yy * a.tiles_x + xx] : a.stat_out[sidx];
        issue_rows<R, BC, QH>(a.T_in, nx, ny, x0, y0 - 1, gt, spec_T);
        __builtin_amdgcn_sched_barrier(0);  // keep the T loads ahead of the wait
        const uint64_t nine = 0x1ffull;
        known = ((__ballot(lane < 9 && (wv & S_PHI0)) & nine) == nine ? S_PHI0 : 0u) |
                ((__ballot(lane < 9 && (wv & S_PHI1)) & nine) == nine ? S_PHI1 : 0u) |
                ((__ballot(lane < 9 && (wv & S_T0)) & nine) == nine ? S_T0 : 0u);
        has = (uint32_t)__builtin_amdgcn_readlane((int)wv, 9);
        active = ((uint32_t)__builtin_amdgcn_readlane((int)wv, 4) & S_ACTIVE) != 0;
        if ((known & (S_PHI0 | S_T0)) == (S_PHI0 | S_T0) && isfinite(a.K) && isfinite(a.dt) &&
            isfinite(a.inv_3dx2)) {
            // phi = +0 and T = +0 on the tile and its halos: the step leaves
            // both +0 (every term is a signed zero summing to +0); store only
            // what the output buffers do not already hold.  When the output
            // word already reads exactly this status, nothing changes.
            if (has == (S_PHI0 | S_T0))
                return;
            const bool w_phi = !(has & S_PHI0), w_T = !(has & S_T0);
            if (w_phi || w_T) {
                const int g = tid % NG, row = tid / NG;
                const int i0 = x0 + 4 * g, j = y0 + row;
                if (row < TY && j < ny && i0 < nx) {
                    const R zero4[4] = {(R)0, (R)0, (R)0, (R)0};
                    const size_t o = (size_t)j * nx + i0;
                    if (i0 + 3 < nx && (nx & 3) == 0) {
                        if (w_phi)
                            store4_out(a.phi_out + o, zero4, a.stream_out);
                        if (w_T)
                            store4_out(a.T_out + o, zero4, a.stream_out);
                    } else {
                        for (int e = 0; e < 4 && i0 + e < nx; ++e) {
                            if (w_phi)
                                a.phi_out[o + e] = (R)0;
                            if (w_T)
                                a.T_out[o + e] = (R)0;
                        }
                    }
                }
            }
            __syncthreads();  // every wave has read `has` before it changes
            if (tid == 0)
                a.stat_out[sidx] = S_PHI0 | S_T0;
            return;
        }
    }

    // ---- stage phi (rows y0-2..y0+TY+1) and T (rows y0-1..y0+TY) ---------
    // (a field known to be constant is not loaded; its LDS rows are filled
    // only if the full path needs them)
    const bool phi_known = (known & (S_PHI0 | S_PHI1)) != 0;
    const bool T_known = (known & S_T0) != 0;
    int all0 = (known & S_PHI0) ? 1 : 0, all1 = (known & S_PHI1) ? 1 : 0, finite = 1;
    if (!phi_known)
        all0 = all1 = 1;
    // eps and eps eps' at theta = +0 (see ZeroAngle), computed by thread 0
    // into LDS.  Without the status it is done while the staging loads are
    // in flight, and the flatness barrier publishes it; with the status (whose
    // quiescent tiles return before any of this) only for a non-flat tile,
    // behind a barrier of its own.
    const bool z_early = !track;
    auto zero_angle = [&]() {
        R e0, f0;
        eps_of_angle(a, (R)0, e0, f0);
        sm.Z[0] = e0;
        sm.Z[1] = f0;
    };
    if ((nx & 3) == 0) {
        Seg2<R> gp, gh;
        if (!phi_known)
            issue_rows_w<R, BC, PH>(a.phi_in, nx, ny, x0, y0 - 2, gp);
        if (!spec_T && !T_known)
            issue_rows_w<R, BC, QH>(a.T_in, nx, ny, x0, y0 - 1, gt);
        if (active)
            issue_rows_w<R, BC, QH>(a.theta, nx, ny, x0, y0 - 1, gh);
        if (z_early && tid == 0)
            zero_angle();
        int unused0 = 1, unused1 = 1;
        if (!phi_known)
            commit_rows<R, PH, 1>(gp, sP, all0, all1);
        // a T known to be +0 was loaded as +0 (the status holds bitwise)
        if (spec_T || !T_known)
            commit_rows<R, QH, 2>(gt, sT, finite, finite);
        if (active)
            commit_rows<R, QH, 0>(gh, sF, unused0, unused1);
    } else {
        if (z_early && tid == 0)
            zero_angle();
        if (!phi_known)
            stage_rows<R, BC, 1>(a.phi_in, sP, nx, ny, x0, y0 - 2, PH, all0, all1);
        if (!T_known)
            stage_rows<R, BC, 2>(a.T_in, sT, nx, ny, x0, y0 - 1, QH, finite, finite);
        if (active) {
            int unused0 = 1, unused1 = 1;
            stage_rows<R, BC, 0>(a.theta, sF, nx, ny, x0, y0 - 1, QH, unused0, unused1);
        }
    }
    if (T_known && !spec_T)
        fill_rows(sT, QH, (R)0);
    // The old theta of the pass-1 cells: staged in F with phi and T for a
    // tile active last step; otherwise each pass-1 thread loads its own four
    // angles (the 16-byte segment staging would put in its slots) and each
    // halo-column thread its one, straight into registers, waited on only
    // where pass 1 first needs an old angle.  A wave that has seen a
    // non-flat phi knows the tile is non-flat and issues them before the
    // flatness barrier; the other waves after it, if the tile is non-flat.
    // Widths not a multiple of 4 stage cell by cell into F.
    const bool theta_direct = !active && (nx & 3) == 0;
    R thold[4] = {(R)0, (R)0, (R)0, (R)0}, thalo = (R)0;
    int thedge = 0;  // zero-flux edge code of the segment, applied where it is used
    bool theta_issued = false;
    auto issue_theta = [&]() {
        const size_t o = segment_offset<BC>(nx, ny, x0, y0 - 1,
                                            (tid / NG) * (LW / 4) + tid % NG + 1, thedge);
        load4(a.theta + o, thold);
        if (tid < 2 * QH) {
            const int r = tid >> 1, ux = (tid & 1) ? x0 + TX : x0 - 1;
            thalo = a.theta[(size_t)map_index<BC>(y0 - 1 + r, ny) * nx + map_index<BC>(ux, nx)];
        }
        theta_issued = true;
    };
    // block-wide AND of the two flatness flags with one barrier: each wave
    // posts its ballots, every thread combines the eight
    {
        const bool w0 = __ballot(all0 & finite) == __ballot(1);
        const bool w1 = __ballot(all1 & finite) == __ballot(1);
        if ((tid & 63) == 0)
            sm.W[tid >> 6] = (w0 ? 1 : 0) | (w1 ? 2 : 0);
        if (theta_direct && !w0 && !w1)
            issue_theta();
    }
    __syncthreads();
    int fl = 3;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w)
        fl &= sm.W[w];
    const bool flat0 = (fl & 1) != 0;
    const bool flat = fl != 0;

    if (!flat) {
        if (!active && (nx & 3) != 0) {
            int unused0 = 1, unused1 = 1;
            stage_rows<R, BC, 0>(a.theta, sF, nx, ny, x0, y0 - 1, QH, unused0, unused1);
        }
        if (theta_direct && !theta_issued)
            issue_theta();
        if (phi_known)
            fill_rows(sP, PH, (known & S_PHI0) ? (R)0 : (R)1);
        if (!z_early && tid == 0)
            zero_angle();
        // LDS written since the flatness barrier: the zero angle (status
        // kernels), phi rows of a known constant, theta staged cell by cell
        if (!z_early || phi_known || (!active && (nx & 3) != 0))
            __syncthreads();
        const ZeroAngle<R> z0{sm.Z[0], sm.Z[1]};

        // ---- pass 1: one 4-cell group per thread (columns x0 .. x0+63) ----
        {
            const int g = tid % NG, r = tid / NG;       // pass-1 row r: global y0-1+r
            const int kc = 4 * g + 4;                   // LDS column of the first cell
            const int ux = x0 + 4 * g, uy = y0 - 1 + r;
            R e2[4], F[4], th[4];
            bool fired[4];
            if (!theta_direct)
                load4(&sF[r * LW + kc], thold);
            // zero-flux ghost rows/columns take the values of the cell they
            // clamp to; periodic halos already hold wrapped values
            const bool shifted = BC == 1 && (uy < 0 || uy >= ny || ux + 3 >= nx);
            // every lane reads its window (the lane exchange needs all lanes)
            R m[6], z[6], p[6];
            {
                const R *pm = &sP[r * LW], *pz = &sP[(r + 1) * LW], *pp = &sP[(r + 2) * LW];
                win6(pz, kc, z);
                load4(pm + kc, m + 1);
                load4(pp + kc, p + 1);
            }
            if (theta_direct)
                fix_edge(thold, thedge);
            if (!shifted) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    pass1_cell(a, z0, m, z, p, e + 1, thold[e], e2[e], F[e], th[e], fired[e]);
            } else {
                const int cy = uy < 0 ? 0 : (uy >= ny ? ny - 1 : uy);
                const int lr = r + (cy - uy);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int cx = ux + e >= nx ? nx - 1 : ux + e;
                    const int lc = kc + e + (cx - (ux + e));
                    pass1_cell(a, z0, &sP[lr * LW], &sP[(lr + 1) * LW], &sP[(lr + 2) * LW], lc,
                               thold[e], e2[e], F[e], th[e], fired[e]);
                }
            }
            store4(&sE2[r * LW + kc], e2);
            store4(&sF[r * LW + kc], F);
            // theta of owned cells: rewrite the group (kept values unchanged)
            // when any of its cells fired
            if (r >= 1 && r <= TY && uy < ny && (fired[0] | fired[1] | fired[2] | fired[3])) {
                R *dst = a.theta + (size_t)uy * nx + ux;
                if (ux + 3 < nx && (nx & 3) == 0) {
                    store4_out(dst, th, a.stream_out);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (ux + e < nx)
                            dst[e] = th[e];
                }
            }
        }
        // ---- pass 1, halo columns x0-1 and x0+64 (one wave) ---------------
        if (tid < 2 * QH) {
            const int r = tid >> 1, side = tid & 1;
            const int lcol = side ? TX + 4 : 3;          // LDS column
            const int ux = x0 - 4 + lcol, uy = y0 - 1 + r;
            int lc = lcol, lr = r;
            if (BC == 1) {
                const int cx = ux < 0 ? 0 : (ux >= nx ? nx - 1 : ux);
                const int cy = uy < 0 ? 0 : (uy >= ny ? ny - 1 : uy);
                lc += cx - ux;
                lr += cy - uy;
            }
            R e2, F, th;
            bool fired;
            pass1_cell(a, z0, &sP[lr * LW], &sP[(lr + 1) * LW], &sP[(lr + 2) * LW], lc,
                       theta_direct ? thalo : sF[r * LW + lcol], e2, F, th, fired);
            sE2[r * LW + lcol] = e2;
            sF[r * LW + lcol] = F;
        }
        __syncthreads();
    }

    // ---- pass 2: one row of 4 cells per thread ----------------------------
    // threads past the tile's last row (or its grid edge) compute on a
    // valid row and store nothing: every lane takes part in the lane exchange
    const int g = tid % NG, row = tid / NG;
    const int rw = row < TY ? row : TY - 1;
    const int i0 = x0 + 4 * g, j = y0 + row;
    const bool mine = row < TY && j < ny && i0 < nx;
    uint32_t obits = S_PHI0 | S_PHI1 | S_T0;  // what holds on this thread's outputs
    const int kc = 4 * g + 4;  // LDS column of the first cell
    // 6-wide windows: index 0 = column i0-1, 1..4 = the cells, 5 = i0+4
    R Tm[6], Tz[6], Tp[6];
    win6(&sT[rw * LW], kc, Tm);
    win6(&sT[(rw + 1) * LW], kc, Tz);
    win6(&sT[(rw + 2) * LW], kc, Tp);

    R phin[4], Tn[4];
    if (flat) {
        // phi is the tile's constant (+0 or 1); phi' = phi
        const R pc = flat0 ? (R)0 : (R)1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const R lapT = lap9(Tm, Tz, Tp, e + 1) * a.inv_3dx2;
            phin[e] = pc;
            Tn[e] = Tz[e + 1] + lapT * a.dt + a.K * (phin[e] - pc);
        }
    } else {
        // in phases, so that each set of LDS windows dies before the next is
        // read (register pressure): T, phi, the cross terms, then eps^2
        R lapT[4], Tc[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            lapT[e] = lap9(Tm, Tz, Tp, e + 1) * a.inv_3dx2;
            Tc[e] = Tz[e + 1];
        }
        phase_end(lapT);
        phase_end(Tc);
        phase_fence();
        // phi, its Laplacian and gradient, and the neighbours' gradients the
        // cross terms take: gx at N and S, gy at the window's 6 columns (the
        // pass-1 expressions of those cells on the same staged values)
        R phi[4], lap[4], gx[4], gy[4], gxN[4], gxS[4], gyW, gyE;
        {
            R Pm[6], Pz[6], Pp[6];
            win6(&sP[(rw + 1) * LW], kc, Pm);
            win6(&sP[(rw + 2) * LW], kc, Pz);
            win6(&sP[(rw + 3) * LW], kc, Pp);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = e + 1;
                phi[e] = Pz[c];
                lap[e] = lap9(Pm, Pz, Pp, c) * a.inv_3dx2;
                gx[e] = (Pz[c + 1] - Pz[c - 1]) * a.inv_dx;
                gy[e] = (Pp[c] - Pm[c]) * a.inv_dy;
                gxN[e] = (Pp[c + 1] - Pp[c - 1]) * a.inv_dx;
                gxS[e] = (Pm[c + 1] - Pm[c - 1]) * a.inv_dx;
            }
            gyW = (Pp[0] - Pm[0]) * a.inv_dy;
            gyE = (Pp[5] - Pm[5]) * a.inv_dy;
        }
        phase_end(phi);
        phase_end(lap);
        phase_end(gx);
        phase_end(gy);
        phase_end(gxN);
        phase_end(gxS);
        phase_fence();
        // t1 = (F_N gx_N - F_S gx_S) / dy, t2 = -(F_E gy_E - F_W gy_W) / dx
        // (:197-203), each product rounded as the reference rounds it
        R t1[4], t2[4];
        {
            R Fm[4], Fz[6], Fp[4];
            load4(&sF[rw * LW + kc], Fm);
            win6(&sF[(rw + 1) * LW], kc, Fz);
            load4(&sF[(rw + 2) * LW + kc], Fp);
            const R gy6[6] = {gyW, gy[0], gy[1], gy[2], gy[3], gyE};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                t1[e] = (mul_rn(Fp[e], gxN[e]) - mul_rn(Fm[e], gxS[e])) * a.inv_dy;
                t2[e] = -(mul_rn(Fz[e + 2], gy6[e + 2]) - mul_rn(Fz[e], gy6[e])) * a.inv_dx;
            }
        }
        phase_end(t1);
        phase_end(t2);
        phase_fence();
        R e2c[4], t3[4];
        {
            R E2m[4], E2z[6], E2p[4];
            load4(&sE2[rw * LW + kc], E2m);
            win6(&sE2[(rw + 1) * LW], kc, E2z);
            load4(&sE2[(rw + 2) * LW + kc], E2p);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = e + 1;
                const R gEx = (E2z[c + 1] - E2z[c - 1]) * a.inv_dx;
                const R gEy = (E2p[e] - E2m[e]) * a.inv_dy;
                e2c[e] = E2z[c];
                t3[e] = gEx * gx[e] + gEy * gy[e];
            }
        }
        phase_end(e2c);
        phase_end(t3);
        phase_fence();

        // the reaction matters only where phi (1 - phi) != 0 (or T is not
        // finite, which the reference would turn into NaN)
        bool need[4], any = false;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            need[e] = (phi[e] * ((R)1.0 - phi[e]) != (R)0) || !isfinite(Tc[e]);
            any |= need[e];
        }
        R chi[4] = {(R)0, (R)0, (R)0, (R)0};
        if (any && a.noise_a != (R)0) {
            const u32x4 w = philox4x32_10(
                u32x4{(uint32_t)i0 >> 2, (uint32_t)j, a.step, a.instance}, a.key0, a.key1);
            chi[0] = noise_from_word<R>(w.x);
            chi[1] = noise_from_word<R>(w.y);
            chi[2] = noise_from_word<R>(w.z);
            chi[3] = noise_from_word<R>(w.w);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            R react = (R)0;
            if (need[e]) {
                const R m = a.alpha_pi * kg_atan(a.gamma * (a.teq - Tc[e]));
                R q = phi[e] - (R)0.5 + m;
                if (a.noise_a != (R)0)
                    q = q + a.noise_a * chi[e];
                react = phi[e] * ((R)1.0 - phi[e]) * q;
            }
            const R sum = t1[e] + t2[e] + e2c[e] * lap[e] + t3[e] + react;
            phin[e] = phi[e] + sum * a.dt_tau;
            Tn[e] = Tc[e] + lapT[e] * a.dt + a.K * (phin[e] - phi[e]);
        }
    }

    if (mine) {
    // a flat tile's phi is its constant: skip the store where the output
    // buffer already holds it
    const bool w_phi = !(flat && (has & (flat0 ? S_PHI0 : S_PHI1)));
    const size_t o = (size_t)j * nx + i0;
    if (i0 + 3 < nx && (nx & 3) == 0) {
        if (w_phi)
            store4_out(a.phi_out + o, phin, a.stream_out);
        store4_out(a.T_out + o, Tn, a.stream_out);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (i0 + e < nx) {
                if (w_phi)
                    a.phi_out[o + e] = phin[e];
                a.T_out[o + e] = Tn[e];
            }
        }
    }
    if (track) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (i0 + e < nx) {
                if (!is_pos_zero(phin[e]))
                    obits &= ~S_PHI0;
                if (!(phin[e] == (R)1))
                    obits &= ~S_PHI1;
                if (!is_pos_zero(Tn[e]))
                    obits &= ~S_T0;
            }
        }
    }
    }  // mine

    // ---- status of the output buffers on this tile ------------------------
    if (track) {
        const uint32_t wb = (__ballot((obits & S_PHI0) != 0) == __ballot(1) ? S_PHI0 : 0u) |
                            (__ballot((obits & S_PHI1) != 0) == __ballot(1) ? S_PHI1 : 0u) |
                            (__ballot((obits & S_T0) != 0) == __ballot(1) ? S_T0 : 0u);
        if ((tid & 63) == 0)
            sm.W2[tid >> 6] = (int)wb;
        __syncthreads();
        if (tid == 0) {
            uint32_t st = flat ? 0u : S_ACTIVE;
            uint32_t acc = S_PHI0 | S_PHI1 | S_T0;
            for (int w = 0; w < NT / 64; ++w)
                acc &= (uint32_t)sm.W2[w];
            a.stat_out[sidx] = st | acc;
        }
    }
}

// FP32: 4 blocks of 8 waves per CU fit the LDS (4 x 37.5 KB); the register
// allocator is asked for the matching 8 waves per SIMD (at most 64 VGPRs).
// With fewer waves asked for, the allocator is handed the registers the LDS
// limit leaves unused anyway (3 blocks: 6 waves, up to 80 VGPRs), which is
// why the kernel descriptor then shows more VGPRs than the code needs.
// FP64 needs twice the LDS and registers and runs two blocks per CU.
#ifndef KG_WAVES_PER_EU
#define KG_WAVES_PER_EU 8
#endif
template <typename R> struct Occupancy;
template <> struct Occupancy<float> { static constexpr int waves = KG_WAVES_PER_EU; };
template <> struct Occupancy<double> { static constexpr int waves = 1; };

// Tile of tile block b when the launch covers only the chunks of `live`:
// the j-th listed chunk goes to XCD j % 8 and its 8 tiles to consecutive
// blocks of that XCD, as chunk_tile deals all chunks.
// DEV: the list is in device memory (live.dev), else in the kernel
// arguments (live.c); two kernels, so neither carries the other's select.
template <bool DEV>
__device__ __forceinline__ bool listed_tile(const LiveList &live, int b, int tiles_x, int tiles_y,
                                            int &tx, int &ty)
{
    const int i = b / NXCD;
    const int j = (i / (CHUNK_W * CHUNK_H)) * NXCD + b % NXCD;
    if (j >= live.n)
        return false;
    // the 32-bit word holding entry j (a scalar load; there is no 16-bit one)
    const uint32_t w =
        DEV ? live.dev[j >> 1] : reinterpret_cast<const uint32_t *>(live.c)[j >> 1];
    const int c = (int)((j & 1) ? w >> 16 : w & 0xffffu);
    const int k = i % (CHUNK_W * CHUNK_H);
    const int ncx = (tiles_x + CHUNK_W - 1) / CHUNK_W;
    tx = (c % ncx) * CHUNK_W + k % CHUNK_W;
    ty = (c / ncx) * CHUNK_H + k / CHUNK_W;
    return tx < tiles_x && ty < tiles_y;
}

// Liveness request folded into a step launch (every SKIP_PERIOD steps): the
// launch's first q.blocks blocks (a multiple of 8, so that the tile blocks
// keep their XCDs) take NT chunks each, one thread per chunk.  A thread
// reads the input status words of its chunk's tiles and of a margin of
// SKIP_MARGIN_X tile columns and SKIP_MARGIN_Y tile rows (mapped by the
// boundary), all loads in flight together, and stores the flag byte
// tag << 5 | na << 1 | lv (lv: some word is not exactly "phi = +0, T = +0";
// na: tiles of the chunk whose last step took the full path) with a
// system-scope store, straight into the host buffer; the host checks the
// tag, so no fence has to write back the L2.  The launch writes only the
// other status set, so the words read are those of the state it starts from.
// Cells a margin must span (see the static_assert below the kernels).  A
// periodic grid's last tile column or row can be narrower: where a margin
// tile is that partial tile, the window takes the tile beyond it as well
// (a full tile spans enough); elsewhere that extra slot repeats the margin
// tile, adding no condition.
constexpr int LIVE_NEED = 2 * (SKIP_AGE + 1) * SKIP_PERIOD + 2;
static_assert(SKIP_MARGIN_X == 1 && SKIP_MARGIN_Y == 1 && LIVE_NEED <= TX && LIVE_NEED <= TY,
              "one full margin tile spans a margin; the window logic assumes one");
static_assert(CHUNK_W * CHUNK_H < 16, "active-tile count fits 4 bits");
// A batch deals its own chunks, by default smaller than the single grid's:
// its members are small (a 128^2 member has 6 tiles) and mostly busy, so
// finer chunks leave fewer idle tiles in a listed launch.  Measured on the
// 512-member ensemble over 5000 steps: 2 x 2 300 Gcell-updates/s, 2 x 4 293,
// 4 x 2 288, 4 x 1 292, 8 x 1 280.  The shape is the batch's choice (set
// directly or by its autotuner) among these; the liveness request is
// compiled for each.
constexpr int BATCH_CHUNK_W = 2, BATCH_CHUNK_H = 2;
constexpr int BATCH_SHAPES[][2] = {{2, 2}, {2, 4}, {4, 2}, {4, 1}, {8, 1}};
__device__ __forceinline__ int live_wrap(int v, int n, int boundary)
{
    if (boundary == 1)
        return v < 0 ? 0 : (v >= n ? n - 1 : v);
    return ((v % n) + n) % n;
}
// Flag byte of chunk (cx, cy) of a grid of tiles_x x tiles_y tiles (nx x ny
// cells) from the input status words sa of its tiles and margin.
template <int CW = CHUNK_W, int CH = CHUNK_H>
__device__ __forceinline__ uint8_t chunk_flag(const uint32_t *__restrict__ sa, int tiles_x,
                                              int tiles_y, int nx, int ny, int boundary, int cx,
                                              int cy, uint32_t tag)
{
    constexpr uint32_t Q = S_PHI0 | S_T0;
    constexpr int LIVE_W = CW + 2 * SKIP_MARGIN_X + 2, LIVE_H = CH + 2 * SKIP_MARGIN_Y + 2;
    // Window columns cx * CW - SKIP_MARGIN_X - 1 + i, i < LIVE_W, and rows
    // likewise, mapped by the boundary; the outermost column (row) on each
    // side reaches one tile further only where the tile next to it is a
    // narrow periodic edge tile, and otherwise repeats its neighbour.
    // Indices are computed as used (no index arrays: they cost the step
    // kernels that carry this request register spills).
    const int wx = nx - (tiles_x - 1) * TX, wy = ny - (tiles_y - 1) * TY;
    const bool narrow_x = boundary == 0 && wx < LIVE_NEED;
    const bool narrow_y = boundary == 0 && wy < LIVE_NEED;
    const int x_1 = cx * CW - SKIP_MARGIN_X - 1, y_1 = cy * CH - SKIP_MARGIN_Y - 1;
    const int c1 = live_wrap(x_1 + 1, tiles_x, boundary);
    const int cl = live_wrap(x_1 + LIVE_W - 2, tiles_x, boundary);
    const int r1 = live_wrap(y_1 + 1, tiles_y, boundary);
    const int rl = live_wrap(y_1 + LIVE_H - 2, tiles_y, boundary);
    const int col0 = narrow_x && c1 == tiles_x - 1 ? live_wrap(x_1, tiles_x, boundary) : c1;
    const int coll =
        narrow_x && cl == tiles_x - 1 ? live_wrap(x_1 + LIVE_W - 1, tiles_x, boundary) : cl;
    const int row0 = narrow_y && r1 == tiles_y - 1 ? live_wrap(y_1, tiles_y, boundary) : r1;
    const int rowl =
        narrow_y && rl == tiles_y - 1 ? live_wrap(y_1 + LIVE_H - 1, tiles_y, boundary) : rl;
    uint32_t any = 0, act = 0;  // act: tiles of the chunk that took the full path
    for (int j = 0; j < LIVE_H; ++j) {
        const int rr = j == 0 ? row0
                              : (j == LIVE_H - 1 ? rowl : live_wrap(y_1 + j, tiles_y, boundary));
        for (int i = 0; i < LIVE_W; ++i) {
            const int cc = i == 0 ? col0
                                  : (i == LIVE_W - 1 ? coll : live_wrap(x_1 + i, tiles_x, boundary));
            const uint32_t w = sa[rr * tiles_x + cc];
            any |= w ^ Q;
            if (j > SKIP_MARGIN_Y && j <= SKIP_MARGIN_Y + CH && i > SKIP_MARGIN_X &&
                i <= SKIP_MARGIN_X + CW)
                act += (w & S_ACTIVE) ? 1u : 0u;
        }
    }
    return (uint8_t)(tag << 5 | act << 1 | (any ? 1u : 0u));
}

template <typename R>
__device__ __forceinline__ void live_request(const StepArgs<R> &a, const LiveReq &q, int tiles_x,
                                             int tiles_y)
{
    const int ncx = (tiles_x + CHUNK_W - 1) / CHUNK_W, ncy = (tiles_y + CHUNK_H - 1) / CHUNK_H;
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= ncx * ncy)
        return;
    const uint8_t f = chunk_flag(a.stat_in + a.tile_base, tiles_x, tiles_y, a.nx, a.ny, a.boundary,
                                 c % ncx, c / ncx, q.tag);
    __hip_atomic_store(q.flags + c, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The same request over the chunks of a batch (chunks[c] = member, chunk
// column, chunk row): the member's input status set is stat_set + its
// tile_base.  A member without status has every chunk live and heavy.
template <typename R, int CW, int CH>
__global__ __launch_bounds__(256) void k_batch_live(const StepArgs<R> *__restrict__ members,
                                                    const int4 *__restrict__ chunks, int nchunks,
                                                    const uint32_t *__restrict__ stat_set,
                                                    uint32_t tag, uint8_t *__restrict__ flags)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nchunks)
        return;
    const int4 d = chunks[c];
    const StepArgs<R> &a = members[d.x];
    static_assert(CW * CH < 16, "active-tile count fits 4 bits");
    uint8_t f = (uint8_t)(tag << 5 | (uint32_t)(CW * CH) << 1 | 1u);
    if (a.stat_in != nullptr)
        f = chunk_flag<CW, CH>(stat_set + a.tile_base, a.tiles_x, a.tiles_y, a.nx, a.ny,
                               a.boundary, d.y, d.z, tag);
    __hip_atomic_store(flags + c, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Chunk lists of a batch kept on the device (KG_BATCH_GRAPH: no host wait,
// so a captured graph keeps skipping).  The same request and answer cycle as
// the host's (kg_batch.hip, batch_skip_cycle), driven by the device step
// counter t = *step_base + step_add: at every t = q * SKIP_PERIOD,
// k_batch_compact turns the answer of request q - 1 into the listed table
// for steps t .. t + SKIP_PERIOD - 1, and k_batch_live_dev asks request q.
// Request q writes flag buffer q & 1 with tag q % LIVE_TAGS + 1; an answer
// whose flags do not all carry the expected tag (cleared buffers: tag 0) is
// no answer, and the full table is used.  At other steps both return at once.
template <typename R, int CW, int CH>
__global__ __launch_bounds__(256) void k_batch_live_dev(const StepArgs<R> *__restrict__ members,
                                                        const int4 *__restrict__ chunks,
                                                        int nchunks,
                                                        const uint32_t *__restrict__ stat,
                                                        int stat_tiles,
                                                        const uint32_t *__restrict__ step_base,
                                                        uint32_t step_add,
                                                        uint8_t *__restrict__ flags2)
{
    static_assert(CW * CH < 16, "active-tile count fits 4 bits");
    const uint32_t t = *step_base + step_add;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (t % SKIP_PERIOD != 0 || c >= nchunks)
        return;
    const uint32_t q = t / SKIP_PERIOD;
    const uint32_t tag = q % LIVE_TAGS + 1;
    const uint32_t *set = stat + (size_t)(t & 1u) * (size_t)stat_tiles;  // input set of step t
    const int4 d = chunks[c];
    const StepArgs<R> &a = members[d.x];
    uint8_t f = (uint8_t)(tag << 5 | (uint32_t)(CW * CH) << 1 | 1u);
    if (a.stat_in != nullptr)
        f = chunk_flag<CW, CH>(set + a.tile_base, a.tiles_x, a.tiles_y, a.nx, a.ny, a.boundary,
                               d.y, d.z, tag);
    flags2[(size_t)(q & 1u) * (size_t)nchunks + c] = f;
}

// Exclusive prefix sum of v over the 1024 threads of a block; *total gets
// the sum.  Every thread must call it (two barriers).
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int s = wsum[i];
        off += i < w ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

// One block of 1024 threads: the listed table from the answer of request
// q - 1, as the host builds it -- live chunks by active-tile count, most
// first (a stable order within a count), the j-th chunk dealt to XCD j % 8
// at position j / 8 of that XCD's sequence, cw * ch entries per chunk
// (tiles outside the member and the last round's missing chunks are padding
// entries, member -1).  *count = entries, or -1: no valid answer, or every
// chunk live (the full table).  *periods counts the lists made.
template <typename R>
__global__ __launch_bounds__(1024) void k_batch_compact(const StepArgs<R> *__restrict__ members,
                                                        const int4 *__restrict__ chunks,
                                                        int nchunks, int cw, int ch,
                                                        const uint32_t *__restrict__ step_base,
                                                        uint32_t step_add,
                                                        const uint8_t *__restrict__ flags2,
                                                        int4 *__restrict__ out,
                                                        int *__restrict__ count,
                                                        uint32_t *__restrict__ periods)
{
    constexpr int FLAGS_LDS = 49152;  // answers of up to this many chunks are read from LDS
    __shared__ int wsum[16];
    __shared__ uint8_t sflags[FLAGS_LDS];
    const uint32_t t = *step_base + step_add;
    if (t % SKIP_PERIOD != 0)
        return;
    const uint32_t q = t / SKIP_PERIOD;
    const int tid = threadIdx.x;
    if (q == 0) {
        if (tid == 0)
            *count = -1;
        return;
    }
    const uint32_t want_tag = (q - 1) % LIVE_TAGS + 1;
    const uint8_t *flags = flags2 + (size_t)((q - 1) & 1u) * (size_t)nchunks;
    if (nchunks <= FLAGS_LDS) {  // one coalesced read; the passes below read LDS
        for (int c = tid; c < nchunks; c += 1024)
            sflags[c] = flags[c];
        __syncthreads();
        flags = sflags;
    }
    const int per = (nchunks + 1023) / 1024;
    const int lo = min(tid * per, nchunks), hi = min(lo + per, nchunks);
    int bad = 0;
    for (int c = lo; c < hi; ++c)
        bad |= (uint32_t)(flags[c] >> 5) != want_tag;
    if (__syncthreads_or(bad)) {
        if (tid == 0)
            *count = -1;
        return;
    }
    const int area = cw * ch;
    int base = 0;  // chunks placed so far (the same in every thread)
    for (int na = area; na >= 0; --na) {
        const uint32_t want = (uint32_t)na << 1 | 1u;
        int n = 0;
        for (int c = lo; c < hi; ++c)
            n += (flags[c] & 31u) == want ? 1 : 0;
        int total = 0;
        int j = base + block_excl_scan(n, wsum, total);
        for (int c = lo; c < hi; ++c) {
            if ((flags[c] & 31u) != want)
                continue;
            const int4 d = chunks[c];
            const StepArgs<R> &a = members[d.x];
            const int tx = (a.nx + TX - 1) / TX, ty = (a.ny + TY - 1) / TY;
            int4 *o = out + (size_t)(j / 8) * area * 8 + j % 8;
            for (int k = 0; k < area; ++k) {
                const int x = d.y * cw + k % cw, y = d.z * ch + k / cw;
                o[(size_t)k * 8] = x < tx && y < ty ? make_int4(d.x, x * TX, y * TY, 0)
                                                    : make_int4(-1, 0, 0, 0);
            }
            ++j;
        }
        base += total;
    }
    // the last round's missing chunks: padding
    const int rounded = (base + 7) / 8 * 8;
    for (int i = tid; i < (rounded - base) * area; i += 1024) {
        const int jj = base + i / area, k = i % area;
        out[((size_t)(jj / 8) * area + k) * 8 + jj % 8] = make_int4(-1, 0, 0, 0);
    }
    if (tid == 0) {
        const bool all = base == nchunks;
        *count = all ? -1 : rounded * area;
        if (!all)
            *periods += 1u;
    }
}

// REQ as in k_step_fused_listed
template <typename R, int BC, bool REQ>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Occupancy<R>::waves))) void k_step_fused(StepArgs<R> a, LiveReq q)
{
    __shared__ __align__(16) TileSmem<R> sm;
    const int tiles_x = (a.nx + TX - 1) / TX, tiles_y = (a.ny + TY - 1) / TY;
    if (REQ && (int)blockIdx.x < q.blocks) {
        live_request(a, q, tiles_x, tiles_y);
        return;
    }
    int tx, ty, c;
    if (!chunk_tile((int)blockIdx.x - (REQ ? q.blocks : 0), tiles_x, tiles_y, tx, ty, c))
        return;
    step_tile<R, BC>(a, tx * TX, ty * TY, sm);
}

// The step without tile status: every tile loads and stores all of its
// fields (the same results as k_step_fused, whose status only proves some of
// that work unneeded).
template <typename R, int BC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Occupancy<R>::waves))) void k_step_dense(StepArgs<R> a)
{
    __shared__ __align__(16) TileSmem<R> sm;
    const int tiles_x = (a.nx + TX - 1) / TX, tiles_y = (a.ny + TY - 1) / TY;
    int tx, ty, c;
    if (!chunk_tile<DENSE_CHUNK_W, DENSE_CHUNK_H>((int)blockIdx.x, tiles_x, tiles_y, tx, ty, c))
        return;
    step_tile<R, BC, false>(a, tx * TX, ty * TY, sm);
}

// The same step over the chunks of `live` only (a separate kernel, so that
// launches of every chunk keep their small kernel arguments).
// REQ: the launch carries a liveness request in its first q.blocks blocks
// (one launch in SKIP_PERIOD); the others use the kernel without that path,
// which spills fewer scalar registers.
template <typename R, int BC, bool DEV, bool REQ>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Occupancy<R>::waves))) void k_step_fused_listed(StepArgs<R> a, LiveList live, LiveReq q)
{
    __shared__ __align__(16) TileSmem<R> sm;
    const int tiles_x = (a.nx + TX - 1) / TX, tiles_y = (a.ny + TY - 1) / TY;
    if (REQ && (int)blockIdx.x < q.blocks) {
        live_request(a, q, tiles_x, tiles_y);
        return;
    }
    int tx, ty;
    if (!listed_tile<DEV>(live, (int)blockIdx.x - (REQ ? q.blocks : 0), tiles_x, tiles_y, tx, ty))
        return;
    step_tile<R, BC>(a, tx * TX, ty * TY, sm);
}

template <typename R, int BC, bool DEV>
void launch_listed(dim3 grid, hipStream_t s, const StepArgs<R> &b, const LiveList &live,
                   const LiveReq &q)
{
    if (q.blocks > 0)
        hipLaunchKernelGGL((k_step_fused_listed<R, BC, DEV, true>), grid, dim3(NT), 0, s, b, live,
                           q);
    else
        hipLaunchKernelGGL((k_step_fused_listed<R, BC, DEV, false>), grid, dim3(NT), 0, s, b, live,
                           q);
}

// A chunk skipped by a launch must stay all +0 for the launch.  The host uses
// a liveness answer for the SKIP_PERIOD steps that start SKIP_AGE periods
// after it asked, so up to (SKIP_AGE + 1) SKIP_PERIOD steps separate the
// state whose status the request read from the last step that uses it.  A
// step moves nonzero values at most 2 cells (pass 2 reads pass-1 values of
// neighbours, which read phi one cell further), and a tile reads 2 cells
// beyond itself; the margin must cover both.  The request reads the input
// status set only: by the first launch that uses the answer, SKIP_AGE >= 1
// periods later, both buffers of the chunk have been rewritten with states
// the margin keeps all +0 there (or left alone by launches that skipped the
// chunk under an earlier answer, which held the same zeros), so a skipped
// launch's output buffer already holds what the step would write.
static_assert(SKIP_AGE >= 1, "the request reads only the input status set");
static_assert(2 * (SKIP_AGE + 1) * SKIP_PERIOD + 2 <= SKIP_MARGIN_Y * TY &&
                  2 * (SKIP_AGE + 1) * SKIP_PERIOD + 2 <= SKIP_MARGIN_X * TX,
              "chunk skip margin too small for the request period");

// Batched step: every tile of every ensemble member in one launch.  tiles[t]
// = (member, x0, y0); members[m] holds the member's fields (buffer 0 as
// phi_in/T_in, buffer 1 as phi_out/T_out) and parameters.  The step counter
// is *step_base + step_add (the noise counter); its parity is the ping-pong
// direction (a batch always restarts at step 0 in buffer 0), so a captured
// graph replays correctly after its k_advance node moves *step_base on.
// DL: the table is chosen on the device -- the listed table `listed` of
// *listed_n entries when *listed_n >= 0 (built by k_batch_compact), else the
// full one; the grid covers the longer of the two, and both entries are
// loaded together with the count so the choice adds no round trip.
// BCM: the boundary of every member (0 periodic, 1 zero-flux), or 2 for a
// batch that mixes them (both tile bodies compiled in, picked per member:
// twice the code, and about 30 more scalar registers spilled to VGPR lanes).
template <typename R, bool DL, int BCM>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Occupancy<R>::waves))) void k_step_batch(const StepArgs<R> *__restrict__ members,
                                                   const int4 *__restrict__ tiles, int ntiles,
                                                   const uint32_t *__restrict__ step_base,
                                                   uint32_t step_add,
                                                   const int4 *__restrict__ listed,
                                                   const int *__restrict__ listed_n,
                                                   int listed_cap)
{
    __shared__ __align__(16) TileSmem<R> sm;
    const int t = blockIdx.x;
    int4 d;
    if (DL) {
        const int4 none = make_int4(-1, 0, 0, 0);
        const int4 dl = t < listed_cap ? listed[t] : none;
        const int4 df = t < ntiles ? tiles[t] : none;
        const int n = *listed_n;
        d = n >= 0 ? (t < n ? dl : none) : df;
    } else {
        if (t >= ntiles)
            return;
        d = tiles[t];
    }
    if (d.x < 0)
        return;  // padding of the table
    StepArgs<R> a = members[d.x];
    a.step = *step_base + step_add;
    if (a.step & 1u) {
        const R *pi = a.phi_in, *ti = a.T_in;
        a.phi_in = a.phi_out;
        a.T_in = a.T_out;
        a.phi_out = const_cast<R *>(pi);
        a.T_out = const_cast<R *>(ti);
        const uint32_t *si = a.stat_in;
        a.stat_in = a.stat_out;
        a.stat_out = const_cast<uint32_t *>(si);
    }
    if (BCM == 1 || (BCM == 2 && a.boundary == 1))
        step_tile<R, 1>(a, d.y, d.z, sm);
    else
        step_tile<R, 0>(a, d.y, d.z, sm);
}

// Batched step with the members' pointers used as given (no ping-pong) and an
// absolute step counter: the adjoint's recompute of a stored segment.
template <typename R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(Occupancy<R>::waves))) void k_step_batch_direct(
    const StepArgs<R> *__restrict__ members, const int4 *__restrict__ tiles, int ntiles,
    uint32_t step)
{
    __shared__ __align__(16) TileSmem<R> sm;
    const int t = blockIdx.x;
    if (t >= ntiles)
        return;
    const int4 d = tiles[t];
    if (d.x < 0)
        return;  // padding of the table
    StepArgs<R> a = members[d.x];
    a.step = step;
    a.stat_in = nullptr;  // the buffers are not the ones the status describes
    if (a.boundary == 1)
        step_tile<R, 1>(a, d.y, d.z, sm);
    else
        step_tile<R, 0>(a, d.y, d.z, sm);
}

__global__ void k_advance(uint32_t *step_base, uint32_t n) { *step_base += n; }

// Reaction + combination of :206-215 for one cell (two-pass kernel).
template <typename R>
__device__ __forceinline__ void update_cell(const StepArgs<R> &a, int i, int j, R phi, R T,
                                            R lap, R lapT, R e2c, R t1, R t2, R t3, R &phin,
                                            R &Tn)
{
    const R m = a.alpha_pi * kg_atan(a.gamma * (a.teq - T));
    R q = phi - (R)0.5 + m;
    if (a.noise_a != (R)0) {
        const u32x4 w = philox4x32_10(u32x4{(uint32_t)i >> 2, (uint32_t)j, a.step, a.instance},
                                      a.key0, a.key1);
        q = q + a.noise_a * noise_from_word<R>(select_word(w, (uint32_t)i & 3u));
    }
    const R sum = t1 + t2 + e2c * lap + t3 + phi * ((R)1.0 - phi) * q;
    phin = phi + sum * a.dt_tau;
    Tn = T + lapT * a.dt + a.K * (phin - phi);
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
    R th = a.theta[c];
    if (theta_branch(gx, gy, th))
        a.theta[c] = th;
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
// Colour of phi: three linear segments with breaks at 0.9, 0.99 and 1.0
// between the reference's four colours (:318-325), in its float arithmetic.
__device__ __forceinline__ float3 ramp(float phi)
{
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
    return make_float3(lo[0] * (1.0f - ratio) + hi[0] * ratio,
                       lo[1] * (1.0f - ratio) + hi[1] * ratio,
                       lo[2] * (1.0f - ratio) + hi[2] * ratio);
}

// float -> 8-bit unsigned normalised: clamp to [0, 1], scale by 255, round
// to the nearest integer (ties to even)
__device__ __forceinline__ uint32_t unorm8(float c)
{
    c = fminf(fmaxf(c, 0.0f), 1.0f);
    return (uint32_t)rintf(c * 255.0f);
}

// order 0: entry c is cell c (x fastest); order 1: the reference's constant
// buffer order, entry c showing _phi[_INDEX(c / n, c % n)] (square grids).
// RGBA8 false: 4 floats per entry (alpha 1); true: 4 bytes R, G, B, A = 255.
template <typename R, bool RGBA8>
__global__ void k_colormap(const R *phi_f, int nx, int ny, int order, void *out)
{
    const size_t n = (size_t)nx * ny;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        size_t src = c;
        if (order == 1) {
            const size_t side = (size_t)nx;
            src = (c / side) + (size_t)nx * (c % side);
        }
        const float3 col = ramp((float)phi_f[src]);
        if (RGBA8) {
            const uint32_t w = unorm8(col.x) | unorm8(col.y) << 8 | unorm8(col.z) << 16 | 255u << 24;
            static_cast<uint32_t *>(out)[c] = w;
        } else {
            static_cast<float4 *>(out)[c] = make_float4(col.x, col.y, col.z, 1.0f);
        }
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

int fused_tiles(int nx, int ny) { return ((nx + TX - 1) / TX) * ((ny + TY - 1) / TY); }

template <typename R>
hipError_t launch_step_fused(const StepArgs<R> &a, hipStream_t s, const LiveList *live,
                             const LiveReq *req)
{
    StepArgs<R> b = a;
    b.ntiles = fused_tiles(a.nx, a.ny);
    const int tiles_x = (a.nx + TX - 1) / TX, tiles_y = (a.ny + TY - 1) / TY;
    const int nchunks = ((tiles_x + CHUNK_W - 1) / CHUNK_W) * ((tiles_y + CHUNK_H - 1) / CHUNK_H);
    LiveReq q{nullptr, 0u, 0};
    if (req && a.stat_in) {
        q = *req;
        q.blocks = (nchunks + NT - 1) / NT;
        q.blocks = (q.blocks + NXCD - 1) / NXCD * NXCD;  // tile blocks keep their XCDs
    }
    if (live && live->n >= 0 && live->n < nchunks) {
        const int nblocks = (live->n + NXCD - 1) / NXCD * NXCD * CHUNK_W * CHUNK_H + q.blocks;
        if (nblocks == 0)
            return hipSuccess;  // no chunk to step, nothing asked
        const bool dev = live->dev != nullptr;
        if (a.boundary == 1 && dev)
            launch_listed<R, 1, true>(dim3(nblocks), s, b, *live, q);
        else if (a.boundary == 1)
            launch_listed<R, 1, false>(dim3(nblocks), s, b, *live, q);
        else if (dev)
            launch_listed<R, 0, true>(dim3(nblocks), s, b, *live, q);
        else
            launch_listed<R, 0, false>(dim3(nblocks), s, b, *live, q);
        return hipGetLastError();
    }
    if (!a.stat_in) {
        const int nblocks = chunk_blocks(tiles_x, tiles_y, DENSE_CHUNK_W, DENSE_CHUNK_H);
        if (a.boundary == 1)
            hipLaunchKernelGGL((k_step_dense<R, 1>), dim3(nblocks), dim3(NT), 0, s, b);
        else
            hipLaunchKernelGGL((k_step_dense<R, 0>), dim3(nblocks), dim3(NT), 0, s, b);
        return hipGetLastError();
    }
    const int nblocks = chunk_blocks(tiles_x, tiles_y) + q.blocks;
    if (a.boundary == 1 && q.blocks > 0)
        hipLaunchKernelGGL((k_step_fused<R, 1, true>), dim3(nblocks), dim3(NT), 0, s, b, q);
    else if (a.boundary == 1)
        hipLaunchKernelGGL((k_step_fused<R, 1, false>), dim3(nblocks), dim3(NT), 0, s, b, q);
    else if (q.blocks > 0)
        hipLaunchKernelGGL((k_step_fused<R, 0, true>), dim3(nblocks), dim3(NT), 0, s, b, q);
    else
        hipLaunchKernelGGL((k_step_fused<R, 0, false>), dim3(nblocks), dim3(NT), 0, s, b, q);
    return hipGetLastError();
}

// The chunk indices of the chunks flagged live, in the order chunk_tile
// deals all chunks (chunk rows rotated by CHUNK_SKEW): appended to out.
bool live_chunk_order(int nx, int ny, const uint8_t *flags, uint32_t tag,
                      std::vector<uint16_t> &out)
{
    const int tiles_x = (nx + TX - 1) / TX, tiles_y = (ny + TY - 1) / TY;
    const int ncx = (tiles_x + CHUNK_W - 1) / CHUNK_W, ncy = (tiles_y + CHUNK_H - 1) / CHUNK_H;
    // live chunks by their number of active tiles, most first, so that the
    // launch ends on cheap tiles
    for (int chunk = 0; chunk < ncx * ncy; ++chunk)
        if ((uint32_t)(flags[chunk] >> 5) != tag)
            return false;
    for (int na = CHUNK_W * CHUNK_H; na >= 0; --na) {
        const uint32_t want = (uint32_t)na << 1 | 1u;
        for (int chunk = 0; chunk < ncx * ncy; ++chunk) {
            const int cy = chunk / ncx;
            const int cx = (chunk - cy * ncx + CHUNK_SKEW * cy) % ncx;
            const int c = cy * ncx + cx;
            if ((flags[c] & 31u) == want)
                out.push_back((uint16_t)c);
        }
    }
    return true;
}

template <typename R, int CW, int CH>
hipError_t launch_batch_live_shape(const StepArgs<R> *members, const int4 *chunks, int nchunks,
                                   const uint32_t *stat_set, uint32_t tag, uint8_t *flags,
                                   hipStream_t s)
{
    hipLaunchKernelGGL((k_batch_live<R, CW, CH>), dim3((nchunks + 255) / 256), dim3(256), 0, s,
                       members, chunks, nchunks, stat_set, tag, flags);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_batch_live(const StepArgs<R> *members, const int4 *chunks, int nchunks,
                             const uint32_t *stat_set, uint32_t tag, uint8_t *flags, int cw,
                             int ch, hipStream_t s)
{
    if (nchunks <= 0)
        return hipSuccess;
#define KG_BATCH_SHAPE(W, H)                                                                  \
    if (cw == W && ch == H)                                                                   \
        return launch_batch_live_shape<R, W, H>(members, chunks, nchunks, stat_set, tag, flags, \
                                                s);
    KG_BATCH_SHAPE(2, 2)
    KG_BATCH_SHAPE(2, 4)
    KG_BATCH_SHAPE(4, 2)
    KG_BATCH_SHAPE(4, 1)
    KG_BATCH_SHAPE(8, 1)
#undef KG_BATCH_SHAPE
    return hipErrorInvalidValue;  // not a shape of BATCH_SHAPES
}

int fused_chunks_x(int nx) { return ((nx + TX - 1) / TX + CHUNK_W - 1) / CHUNK_W; }
int fused_chunks_y(int ny) { return ((ny + TY - 1) / TY + CHUNK_H - 1) / CHUNK_H; }

int fused_tile_x() { return TX; }
int fused_tile_y() { return TY; }
int fused_chunk_w() { return CHUNK_W; }
int fused_chunk_h() { return CHUNK_H; }
int batch_chunk_w() { return BATCH_CHUNK_W; }
int batch_chunk_h() { return BATCH_CHUNK_H; }
int batch_chunk_shapes(int (*out)[2], int cap)
{
    const int n = (int)(sizeof(BATCH_SHAPES) / sizeof(BATCH_SHAPES[0]));
    for (int i = 0; i < n && i < cap; ++i) {
        out[i][0] = BATCH_SHAPES[i][0];
        out[i][1] = BATCH_SHAPES[i][1];
    }
    return n;
}
int fused_chunk_skew() { return CHUNK_SKEW; }

bool fused_status_ok(int nx, int ny, int boundary)
{
    // zero-flux maps halos onto the tile itself or its edge neighbours;
    // periodic wraps a 2-cell halo onto the far tile row/column, which must
    // then hold at least 2 cells
    if (boundary == 1)
        return true;
    return nx % TX != 1 && ny % TY != 1;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

template <typename R, bool DL>
hipError_t launch_step_batch_bc(int bcm, dim3 grid, hipStream_t s, const StepArgs<R> *members,
                                const int4 *tiles, int ntiles, const uint32_t *step_base,
                                uint32_t step_add, const int4 *listed, const int *listed_n,
                                int listed_cap)
{
    if (bcm == 0)
        hipLaunchKernelGGL((k_step_batch<R, DL, 0>), grid, dim3(NT), 0, s, members, tiles, ntiles,
                           step_base, step_add, listed, listed_n, listed_cap);
    else if (bcm == 1)
        hipLaunchKernelGGL((k_step_batch<R, DL, 1>), grid, dim3(NT), 0, s, members, tiles, ntiles,
                           step_base, step_add, listed, listed_n, listed_cap);
    else
        hipLaunchKernelGGL((k_step_batch<R, DL, 2>), grid, dim3(NT), 0, s, members, tiles, ntiles,
                           step_base, step_add, listed, listed_n, listed_cap);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_step_batch(const StepArgs<R> *members, const int4 *tiles, int ntiles,
                             const uint32_t *step_base, uint32_t step_add, int bcm, hipStream_t s,
                             const DeviceLists *dl)
{
    if (dl == nullptr)
        return launch_step_batch_bc<R, false>(bcm, dim3(ntiles), s, members, tiles, ntiles,
                                              step_base, step_add, nullptr, nullptr, 0);
    hipLaunchKernelGGL((k_batch_compact<R>), dim3(1), dim3(1024), 0, s, members, dl->chunks,
                       dl->nchunks, dl->cw, dl->ch, step_base, step_add, dl->flags2, dl->listed,
                       dl->count, dl->periods);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const dim3 lg((dl->nchunks + 255) / 256), lb(256);
#define KG_BATCH_SHAPE(W, H)                                                                   \
    else if (dl->cw == W && dl->ch == H) hipLaunchKernelGGL((k_batch_live_dev<R, W, H>), lg, lb, \
                                                            0, s, members, dl->chunks,          \
                                                            dl->nchunks, dl->stat,              \
                                                            dl->stat_tiles, step_base,          \
                                                            step_add, dl->flags2);
    if (dl->nchunks <= 0)
        return hipErrorInvalidValue;
    KG_BATCH_SHAPE(2, 2)
    KG_BATCH_SHAPE(2, 4)
    KG_BATCH_SHAPE(4, 2)
    KG_BATCH_SHAPE(4, 1)
    KG_BATCH_SHAPE(8, 1)
    else
        return hipErrorInvalidValue;  // not a shape of BATCH_SHAPES
#undef KG_BATCH_SHAPE
    e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const int grid = std::max(ntiles, dl->listed_cap);
    return launch_step_batch_bc<R, true>(bcm, dim3(grid), s, members, tiles, ntiles, step_base,
                                         step_add, dl->listed, dl->count, dl->listed_cap);
}

template <typename R>
hipError_t launch_step_batch_direct(const StepArgs<R> *members, const int4 *tiles, int ntiles,
                                    uint32_t step, hipStream_t s)
{
    hipLaunchKernelGGL((k_step_batch_direct<R>), dim3(ntiles), dim3(NT), 0, s, members, tiles,
                       ntiles, step);
    return hipGetLastError();
}

hipError_t launch_advance(uint32_t *step_base, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s, step_base, n);
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
hipError_t launch_colormap(const R *phi, int nx, int ny, int order, void *rgba, bool rgba8,
                           hipStream_t s)
{
    const size_t n = (size_t)nx * ny;
    if (rgba8)
        hipLaunchKernelGGL((k_colormap<R, true>), dim3(grid_1d(n, 256)), dim3(256), 0, s, phi, nx,
                           ny, order, rgba);
    else
        hipLaunchKernelGGL((k_colormap<R, false>), dim3(grid_1d(n, 256)), dim3(256), 0, s, phi, nx,
                           ny, order, rgba);
    return hipGetLastError();
}

// ---- front profile (morphology) ---------------------------------------------
// Distance from (cx, cy) to the first point where phi drops below `level`
// along n equally spaced rays: march in steps of 1/4 cell, sample phi
// bilinearly (points outside the grid read 0), interpolate linearly between
// the last two samples.  FP64 throughout; one thread per ray.
template <typename R>
__global__ void k_front_profile(const R *__restrict__ phi, int nx, int ny, double cx, double cy,
                                int n, double level, double *__restrict__ radii)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const double ang = 2.0 * 3.14159265358979323846 * (double)k / (double)n;
    const double c = cos(ang), sn = sin(ang);
    const double rmax = sqrt((double)nx * nx + (double)ny * ny);
    const long nsteps = (long)ceil(rmax / 0.25);
    double prev = 0.0;
    double out = (double)(nsteps - 1) * 0.25;  // no crossing: the last step
    for (long i = 0; i < nsteps; ++i) {
        const double st = 0.25 * (double)i;
        const double x = cx + c * st, y = cy + sn * st;
        double v = 0.0;
        if (x >= 0.0 && x <= (double)(nx - 1) && y >= 0.0 && y <= (double)(ny - 1)) {
            int x0 = (int)floor(x), y0 = (int)floor(y);
            x0 = x0 < 0 ? 0 : (x0 > nx - 2 ? nx - 2 : x0);
            y0 = y0 < 0 ? 0 : (y0 > ny - 2 ? ny - 2 : y0);
            double fx = x - x0, fy = y - y0;
            fx = fx < 0.0 ? 0.0 : (fx > 1.0 ? 1.0 : fx);
            fy = fy < 0.0 ? 0.0 : (fy > 1.0 ? 1.0 : fy);
            const size_t o = (size_t)y0 * nx + x0;
            v = (double)phi[o] * (1.0 - fx) * (1.0 - fy) + (double)phi[o + 1] * fx * (1.0 - fy) +
                (double)phi[o + nx] * (1.0 - fx) * fy + (double)phi[o + nx + 1] * fx * fy;
        }
        if (v < level) {
            if (i == 0) {
                out = 0.0;
            } else {
                const double t = prev != v ? (prev - level) / (prev - v) : 0.0;
                out = 0.25 * (double)(i - 1) + t * 0.25;
            }
            break;
        }
        prev = v;
    }
    radii[k] = out;
}

template <typename R>
hipError_t launch_front_profile(const R *phi, int nx, int ny, double cx, double cy, int n,
                                double level, double *radii, hipStream_t s)
{
    hipLaunchKernelGGL((k_front_profile<R>), dim3((n + 63) / 64), dim3(64), 0, s, phi, nx, ny, cx,
                       cy, n, level, radii);
    return hipGetLastError();
}

#define KG_INSTANTIATE(R)                                                                        \
    template hipError_t launch_step_fused<R>(const StepArgs<R> &, hipStream_t, const LiveList *, \
                                             const LiveReq *);                                \
    template hipError_t launch_batch_live<R>(const StepArgs<R> *, const int4 *, int,             \
                                             const uint32_t *, uint32_t, uint8_t *, int, int,     \
                                             hipStream_t);                                        \
    template hipError_t launch_step_batch<R>(const StepArgs<R> *, const int4 *, int,             \
                                             const uint32_t *, uint32_t, int, hipStream_t,       \
                                             const DeviceLists *);                               \
    template hipError_t launch_step_batch_direct<R>(const StepArgs<R> *, const int4 *, int,      \
                                                    uint32_t, hipStream_t);                      \
    template hipError_t launch_step_two_pass<R>(const StepArgs<R> &, const PassTemps<R> &,       \
                                                hipStream_t);                                    \
    template hipError_t launch_init<R>(R *, R *, R *, int, int, int, int, int, const int *, int, \
                                       double, hipStream_t);                                     \
    template hipError_t launch_reduce<R>(const R *, const R *, size_t, double, double *,         \
                                         ReduceOut *, hipStream_t);                              \
    template hipError_t launch_colormap<R>(const R *, int, int, int, void *, bool, hipStream_t);  \
    template hipError_t launch_front_profile<R>(const R *, int, int, double, double, int, double,  \
                                                double *, hipStream_t);
KG_INSTANTIATE(float)
KG_INSTANTIATE(double)

}  // namespace kg
