// rcg_halo.hip -- FP32 gridder/degridder over halo records: whole windows,
// walked in runs of one width with the lane geometry fixed at compile time.
//
// Same mathematics as the reference's scatter gridder and its transpose
// degridder (MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p
// .cpp:146-176 and :368-431), organised for CDNA4 around the record layout
// RCG_LAYOUT_HALO (rcg_tiled.hip, point_records): every record is a whole
// window of w x h cells (w, h <= wmax <= 32; larger windows arrive as
// near-equal sub-windows), binned in the T x T tile that holds its first
// cell.  A tile's records therefore reach at most T + wmax - 1 cells down
// and across, the tile's "region", which the kernels hold in LDS.  Both
// kernels run one wavefront per workgroup, each on its own work item (at
// most MAX_HALO_ITEM records of one tile):
//
//  * gridding: the wavefront zeroes its private FP32 copy of the region and
//    adds every record into it (a cell's read-modify-write needs no atomic:
//    the lanes of one record own distinct cells, and a wavefront's LDS
//    operations complete in order).  At the end of the item the region goes
//    to the item's buffer with plain stores, and k_gather_halo sums the
//    buffers of the overlapping regions into each grid cell in FP64, in a
//    fixed order: the grid is reproducible and needs no FP32 atomic (items
//    past the buffers' budget add their cells with atomics instead);
//  * degridding: the wavefront stages the region of the grid once and forms
//    four partial products per lane for each record, summed across the
//    wavefront four records at a time; a whole window's sum is its
//    visibility, a sub-window's goes to its slot of the partial buffer.
//
// Runs.  The scalar unit issues one instruction per CU per cycle, as many as
// the vector units retire, so the work per record must be a handful of
// scalar and vector instructions besides its FMAs.  The wavefront therefore
// sorts its item's records by width (and squareness) in LDS and walks them
// in runs: per run one dispatch on the width, then a loop compiled for that
// width.  The lanes take K adjacent columns each (two for an even width,
// read as one 16-byte tap load; one for an odd width), C = w / K lanes per
// window row, R = 64 / C rows per 64-lane slab, lanes from R * C on idle; a
// lane's row and column, every slab's LDS offset (an immediate) and the
// last slab's lane mask are constants of the run.  Per record there remain
// the readlanes of its fields, its tap offsets (scalar), one vector address
// and the FMAs.
//
// Latency.  A vector memory instruction costs the CU about 16 cycles whatever
// its width, and an L2 hit takes hundreds: a wavefront keeps the taps of the
// next records in flight while it works on the current one (three for small
// windows, fewer for large ones, whose taps take more registers).  In a run
// of square windows every record issues the same number of tap loads (past
// the run's end at an out-of-range offset, which reads zero at no cost), so
// the compiler's wait on a record's taps is exact.  Non-square sub-windows
// (windows above wmax) load their taps when they are worked on.
#include <type_traits>

#include "rcg_internal.h"
#include "rcg_tiled.h"
#include "rcg_tiledev.h"

namespace rcg {

namespace {

// Run keys: (w - 1) + 32 (h != w), 64 of them.
constexpr int NKEY = 64;

// LDS region of a tile: T + H rows and columns (H = wmax - 1), row stride LS.
template <int T, int H>
struct Region {
  static constexpr int RW = T + H;
  static constexpr int LS = RW + 1;
  static constexpr int NC = RW * LS;
};

// f(std::integral_constant<int, W>()) for W = w, 1 <= w <= WMAX.
template <int WMAX, class F>
__device__ __forceinline__ void with_width(int w, F &&f)
{
#define RCG_W(W)                                                      \
  case W:                                                             \
    if constexpr (W <= WMAX) f(std::integral_constant<int, W>());     \
    break;
  switch (w) {
    RCG_W(1) RCG_W(2) RCG_W(3) RCG_W(4) RCG_W(5) RCG_W(6) RCG_W(7) RCG_W(8)
    RCG_W(9) RCG_W(10) RCG_W(11) RCG_W(12) RCG_W(13) RCG_W(14) RCG_W(15) RCG_W(16)
    RCG_W(17) RCG_W(18) RCG_W(19) RCG_W(20) RCG_W(21) RCG_W(22) RCG_W(23) RCG_W(24)
    RCG_W(25) RCG_W(26) RCG_W(27) RCG_W(28) RCG_W(29) RCG_W(30) RCG_W(31) RCG_W(32)
    default: break;
  }
#undef RCG_W
}

// Lane geometry of a record of width W (see the file comment).  A square
// record (h == W) has NSM = ceil(W / R) slabs, the last with HL rows; a
// non-square one (a sub-window, h <= WMAX) at most NSM = ceil(WMAX / R).
template <int W, bool SQ, int WMAX>
struct Geo {
  static constexpr int K = W % 2 == 0 ? 2 : 1;
  static constexpr int C = W / K;
  static constexpr int R = 64 / C;
  static constexpr int NSM = SQ ? (W + R - 1) / R : (WMAX + R - 1) / R;
  static constexpr int HL = SQ ? W - (NSM - 1) * R : 0;
  using Taps = std::conditional_t<K == 2, float4, float2>;
};

// Square records of a run in flight at once (the one worked on and those
// whose taps are loading): their taps take F x TV vector registers, TV per
// record, kept near 64 so that several wavefronts fit on a SIMD.
template <class G>
constexpr int depth()
{
  constexpr int TV = G::NSM * (G::K == 2 ? 4 : 2);
  return TV >= 32 ? 2 : TV >= 16 ? 3 : 4;
}

template <int K>
__device__ __forceinline__ void load_taps(std::conditional_t<K == 2, float4, float2> &t,
                                          __amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
  if constexpr (K == 2) {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    t = make_float4(__int_as_float(v[0]), __int_as_float(v[1]), __int_as_float(v[2]),
                    __int_as_float(v[3]));
  } else {
    t = tap_load(r, voff, soff);
  }
}

__device__ __forceinline__ float2 tap_k(const float4 &t, int k)
{
  return k ? make_float2(t.z, t.w) : make_float2(t.x, t.y);
}
__device__ __forceinline__ float2 tap_k(const float2 &t, int) { return t; }

// The lane index, opaque to the compiler, so that the lane constants of a
// run are computed at the run rather than hoisted out of the item loop for
// every width at once (which would hold them all in vector registers).
__device__ __forceinline__ int opaque_lane(int lane)
{
  asm volatile("" : "+v"(lane));
  return lane;
}

__device__ __forceinline__ int run_key(int px)
{
  const int w = ((px >> 12) & 63) + 1, h = ((px >> 18) & 63) + 1;
  return (w - 1) + (h != w ? 32 : 0);
}

// The wavefront's next work item (-1 when all are drawn): the draw of
// draw_item by lane 0, broadcast.
__device__ __forceinline__ int next_item(int *q, int total, int &tried, int lane)
{
  int item = -1;
  if (lane == 0) item = draw_item(q, total, blockIdx.x & 7, tried);
  return __builtin_amdgcn_readfirstlane(item);
}

// Run table of a work item: lane `key` gets its key's first record (within
// the item, whose records are sorted by key) and count.
__device__ __forceinline__ void item_runs(const TiledDev &d, int item, int lane, int &first,
                                          int &count)
{
  count = d.runs[(long long)item * NKEY + lane];
  int x = count;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  first = x - count;
}

// The records of a run, 64 at a time (one per lane, the next batch loaded a
// batch ahead) with, for gridding, their visibilities: batch b is held in
// (r0, m0) for even b and (r1, m1) for odd b, so that moving on loads into
// the registers just used up and nothing is copied.  at(k) moves to record
// k's batch (k increasing); field(k) reads record k's fields.
template <bool VIS>
struct RunStream {
  const int4 *rp;
  const float2 *vp;
  int n, lane;
  int4 r0, r1;
  float2 m0, m1;
  // Batch b into (r, m): global loads for even batches, raw buffer loads
  // (out of range past the run: zero) for odd ones.  The two kinds keep the
  // compiler from merging the loads of the two branches of at() into one
  // load through a pointer to either member, which would put the stream in
  // scratch memory.
  __device__ __forceinline__ void load0(int b)
  {
    const int k = b * 64 + lane;
    r0 = make_int4(0, 0, 0, 0);
    m0 = make_float2(0.f, 0.f);
    if (k < n) {
      r0 = rp[k];
      if (VIS) m0 = vp[k];
    }
  }
  __device__ __forceinline__ void load1(int b)
  {
    const unsigned voff = (unsigned)(b * 64 + lane);
    auto rr = __builtin_amdgcn_make_buffer_rsrc(const_cast<int4 *>(rp), (short)0, n * 16,
                                                0x00020000);
    auto v = __builtin_amdgcn_raw_buffer_load_b128(rr, voff * 16u, 0, 0);
    r1 = make_int4(v[0], v[1], v[2], v[3]);
    if (VIS) {
      auto vr = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(vp), (short)0, n * 8,
                                                  0x00020000);
      m1 = tap_load(vr, voff * 8u, 0);
    }
  }
  __device__ __forceinline__ RunStream(const int4 *rp_, const float2 *vp_, int n_, int lane_)
      : rp(rp_), vp(vp_), n(n_), lane(lane_)
  {
    load0(0);
    load1(1);
  }
  // at the first record of batch b >= 1, batch b - 1 is used up: its
  // registers take batch b + 1
  __device__ __forceinline__ bool at(int k)
  {
    if ((k & 63) == 0 && k > 0) {
      if ((k >> 6) & 1) load0((k >> 6) + 1);
      else load1((k >> 6) + 1);
      return true;
    }
    return false;
  }
  __device__ __forceinline__ void field(int k, int &px, int &py, int &pz, float &vx,
                                        float &vy) const
  {
    const int j = k & 63;
    if ((k >> 6) & 1) {
      px = rl(r1.x, j), py = rl(r1.y, j), pz = rl(r1.z, j);
      vx = VIS ? rlf(m1.x, j) : 0.f, vy = VIS ? rlf(m1.y, j) : 0.f;
    } else {
      px = rl(r0.x, j), py = rl(r0.y, j), pz = rl(r0.z, j);
      vx = VIS ? rlf(m0.x, j) : 0.f, vy = VIS ? rlf(m0.y, j) : 0.f;
    }
  }
  // output code of this lane's record in record k's batch (degridding)
  __device__ __forceinline__ int code(int k, int npts) const
  {
    int c;
    if ((k >> 6) & 1) c = out_code(r1, npts);
    else c = out_code(r0, npts);
    return c;
  }
};

// A record's scalar fields: first cell (LDS byte address), tap and row
// stride (bytes), height, and the FMA coefficients: for gridding the
// visibility folded with the conjugations (re += a tx + b ty,
// im += c ty + e tx), for degridding a = the sign of ty.  live = 0 past the
// run's end.
struct Rec {
  int live, lbase, tapb, strb, h;
  float a, b, c, e;
};

template <bool VIS, int LS>
__device__ __forceinline__ Rec fields(RunStream<VIS> &rs, int k, int adj)
{
  Rec f;
  if (k >= rs.n) {
    f.live = 0;
    f.lbase = f.tapb = f.strb = 0;
    f.h = 1;
    f.a = f.b = f.c = f.e = 0.f;
    return f;
  }
  rs.at(k);
  int px, py, pz;
  float vx, vy;
  rs.field(k, px, py, pz, vx, vy);
  f.live = 1;
  f.lbase = ((px & 63) * LS + ((px >> 6) & 63)) * 8;
  f.h = ((px >> 18) & 63) + 1;
  f.tapb = py * 8;  // tap offsets are below 2^28 (check_tap_table)
  f.strb = (pz & 0xffff) * 8;
  const int conj = (px >> 30) & 1;
  if (VIS) {
    // taps are conjugated for mirrored (w < 0) layers and, in the adjoint,
    // once more for all of them: b = conj ? vy : -vy, c = conj ? -vx : vx
    const unsigned flip = conj != adj ? 0u : 0x80000000u;
    f.a = vx;
    f.e = vy;
    f.b = __int_as_float(__float_as_int(vy) ^ flip);
    f.c = __int_as_float(__float_as_int(vx) ^ flip ^ 0x80000000u);
  } else {
    f.a = conj ? -1.f : 1.f;
    f.b = f.c = f.e = 0.f;
  }
  return f;
}

// Lane constants of a run of width W.
template <class G, int LS>
struct Lane {
  int lr, lp, lds;  // row, column group, LDS byte offset from the record's first cell
  bool act;         // the lane has a row in a full slab
  bool last;        // ... and in a square's last slab
  __device__ __forceinline__ explicit Lane(int lane)
  {
    lr = lane / G::C;
    lp = lane - lr * G::C;
    lds = (lr * LS + lp * G::K) * 8;
    act = lr < G::R;
    last = lr < G::HL;
  }
  __device__ __forceinline__ unsigned voff(int strb) const
  {
    return (unsigned)(lr * strb + lp * 8 * G::K);
  }
};

// Taps of a square record (all NSM slabs; the lanes outside a slab, and
// every lane past the run's end, at an out-of-range offset).
template <class G, class L>
__device__ __forceinline__ void issue_sq(typename G::Taps (&t)[G::NSM], const Rec &f, const L &ln,
                                         __amdgpu_buffer_rsrc_t rsrc)
{
  const unsigned v = ln.voff(f.strb);
  const unsigned v0 = f.live && ln.act ? v : TAP_OOB;
  const unsigned vl = f.live && ln.last ? v : TAP_OOB;
#pragma unroll
  for (int s = 0; s < G::NSM; s++)
    load_taps<G::K>(t[s], rsrc, s == G::NSM - 1 ? vl : v0, f.tapb + s * G::R * f.strb);
}

// One square record into the region copy.
template <class G, int LS, class L>
__device__ __forceinline__ void grid_sq(const typename G::Taps (&t)[G::NSM], const Rec &f,
                                        const L &ln, float2 *reg)
{
  if (!f.live || !ln.act) return;
  constexpr int K = G::K, R = G::R, NS = G::NSM;
  const int la = f.lbase + ln.lds;
  float2 x[NS][K];
#pragma unroll
  for (int s = 0; s < NS; s++)
    if (s < NS - 1 || ln.last) {
#pragma unroll
      for (int k = 0; k < K; k++) x[s][k] = lds_at(reg, la + (s * R * LS + k) * 8);
    }
#pragma unroll
  for (int s = 0; s < NS; s++)
    if (s < NS - 1 || ln.last) {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const float2 q = tap_k(t[s], k);
        x[s][k].x = fmaf(f.a, q.x, fmaf(f.b, q.y, x[s][k].x));
        x[s][k].y = fmaf(f.c, q.y, fmaf(f.e, q.x, x[s][k].y));
        lds_at(reg, la + (s * R * LS + k) * 8) = x[s][k];
      }
    }
}

// Slab s of a non-square record of height h is worked on by the lanes with a
// row in it.
template <class G, class L>
__device__ __forceinline__ bool ns_on(int s, int h, const L &ln)
{
  const int ns = (h + G::R - 1) / G::R;
  return s < ns && ln.act && (s < ns - 1 || ln.lr < h - (ns - 1) * G::R);
}

// One non-square record (a sub-window) into the region copy, its taps loaded
// here.
template <class G, int LS, class L>
__device__ __forceinline__ void grid_nsq(const Rec &f, const L &ln, float2 *reg,
                                         __amdgpu_buffer_rsrc_t rsrc)
{
  constexpr int K = G::K, R = G::R, NSM = G::NSM;
  if (!ln.act) return;
  const int la = f.lbase + ln.lds;
  const unsigned v = ln.voff(f.strb);
  typename G::Taps t[NSM];
  float2 x[NSM][K];
#pragma unroll
  for (int s = 0; s < NSM; s++)
    if (ns_on<G>(s, f.h, ln)) load_taps<K>(t[s], rsrc, v, f.tapb + s * R * f.strb);
#pragma unroll
  for (int s = 0; s < NSM; s++)
    if (ns_on<G>(s, f.h, ln)) {
#pragma unroll
      for (int k = 0; k < K; k++) x[s][k] = lds_at(reg, la + (s * R * LS + k) * 8);
    }
#pragma unroll
  for (int s = 0; s < NSM; s++)
    if (ns_on<G>(s, f.h, ln)) {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const float2 q = tap_k(t[s], k);
        x[s][k].x = fmaf(f.a, q.x, fmaf(f.b, q.y, x[s][k].x));
        x[s][k].y = fmaf(f.c, q.y, fmaf(f.e, q.x, x[s][k].y));
        lds_at(reg, la + (s * R * LS + k) * 8) = x[s][k];
      }
    }
}

// A run of n records of width W (square or not) into the region copy.
template <int W, bool SQ, int WMAX, int LS>
__device__ __forceinline__ void grid_run(const TiledDev &d, __amdgpu_buffer_rsrc_t rsrc,
                                         float2 *reg, int first, int n, int lane)
{
  using G = Geo<W, SQ, WMAX>;
  const Lane<G, LS> ln(opaque_lane(lane));
  RunStream<true> rs(d.pairs + first, d.vsort + first, n, lane);
  if constexpr (SQ) {
    // F records in flight: record k + F - 1's taps are issued before record
    // k is worked on, into the slot record k - 1 left
    constexpr int F = depth<G>();
    typename G::Taps t[F][G::NSM];
    Rec f[F];
#pragma unroll
    for (int j = 0; j < F - 1; j++) {
      f[j] = fields<true, LS>(rs, j, d.adj);
      issue_sq<G>(t[j], f[j], ln, rsrc);
    }
    for (int k = 0; k < n; k += F) {
#pragma unroll
      for (int j = 0; j < F; j++) {
        const int nx = (j + F - 1) % F;
        f[nx] = fields<true, LS>(rs, k + j + F - 1, d.adj);
        issue_sq<G>(t[nx], f[nx], ln, rsrc);
        grid_sq<G, LS>(t[j], f[j], ln, reg);
      }
    }
  } else {
    for (int k = 0; k < n; k++) grid_nsq<G, LS>(fields<true, LS>(rs, k, d.adj), ln, reg, rsrc);
  }
}

// Degridding counterparts: this lane's part of
// sum_ij grid(r0 + i, c0 + j) * tap'(i, j), tap' = (tx, sg ty), sg = f.a.
__device__ __forceinline__ float2 pqru(float P, float Q, float Rs, float U, float sg)
{
  // sg = +-1, so each fma rounds as the add or subtract would
  return make_float2(fmaf(-sg, Q, P), fmaf(sg, Rs, U));
}

template <class G, int LS, class L>
__device__ __forceinline__ float2 degrid_sq(const typename G::Taps (&t)[G::NSM], const Rec &f,
                                            const L &ln, const float2 *reg)
{
  constexpr int K = G::K, R = G::R, NS = G::NSM;
  float P = 0.f, Q = 0.f, Rs = 0.f, U = 0.f;
  if (f.live && ln.act) {
    const int la = f.lbase + ln.lds;
    float2 x[NS][K];
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (s < NS - 1 || ln.last) {
#pragma unroll
        for (int k = 0; k < K; k++)
          x[s][k] = lds_at(const_cast<float2 *>(reg), la + (s * R * LS + k) * 8);
      }
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (s < NS - 1 || ln.last) {
#pragma unroll
        for (int k = 0; k < K; k++) {
          const float2 q = tap_k(t[s], k);
          P = fmaf(x[s][k].x, q.x, P);
          Q = fmaf(x[s][k].y, q.y, Q);
          Rs = fmaf(x[s][k].x, q.y, Rs);
          U = fmaf(x[s][k].y, q.x, U);
        }
      }
  }
  return pqru(P, Q, Rs, U, f.a);
}

template <class G, int LS, class L>
__device__ __forceinline__ float2 degrid_nsq(const Rec &f, const L &ln, const float2 *reg,
                                             __amdgpu_buffer_rsrc_t rsrc)
{
  constexpr int K = G::K, R = G::R, NSM = G::NSM;
  float P = 0.f, Q = 0.f, Rs = 0.f, U = 0.f;
  if (ln.act) {
    const int la = f.lbase + ln.lds;
    const unsigned v = ln.voff(f.strb);
    typename G::Taps t[NSM];
    float2 x[NSM][K];
#pragma unroll
    for (int s = 0; s < NSM; s++)
      if (ns_on<G>(s, f.h, ln)) load_taps<K>(t[s], rsrc, v, f.tapb + s * R * f.strb);
#pragma unroll
    for (int s = 0; s < NSM; s++)
      if (ns_on<G>(s, f.h, ln)) {
#pragma unroll
        for (int k = 0; k < K; k++)
          x[s][k] = lds_at(const_cast<float2 *>(reg), la + (s * R * LS + k) * 8);
      }
#pragma unroll
    for (int s = 0; s < NSM; s++)
      if (ns_on<G>(s, f.h, ln)) {
#pragma unroll
        for (int k = 0; k < K; k++) {
          const float2 q = tap_k(t[s], k);
          P = fmaf(x[s][k].x, q.x, P);
          Q = fmaf(x[s][k].y, q.y, Q);
          Rs = fmaf(x[s][k].x, q.y, Rs);
          U = fmaf(x[s][k].y, q.x, U);
        }
      }
  }
  return pqru(P, Q, Rs, U, f.a);
}

// Per-wavefront sums of the degridder: put(k, o) leaves this lane's part of
// the run's record k in row k % 4 of sums; every fourth record, and at the
// end of the run, flush_sums adds each row across the wavefront and stores
// the record's sum to the output its code names (codes[k % 128], written
// by lane k % 64 when record k's batch of 64 becomes current: two batches
// held, the current one and the one whose last sums may be pending).
struct Sums {
  float2 *sums;
  int *codes;
  float2 *vis, *partial;
  int npts, lane;
  __device__ __forceinline__ void batch(int code, int k)
  {
    codes[(k & 64) + lane] = code;
  }
  __device__ __forceinline__ void put(int k, float2 o)
  {
    sums[(k & (SUMS - 1)) * 64 + lane] = o;
    if ((k & (SUMS - 1)) == SUMS - 1) flush_sums(sums, codes + (k & 124), vis, partial, SUMS, lane);
  }
  __device__ __forceinline__ void end(int n)
  {
    if (n & (SUMS - 1))
      flush_sums(sums, codes + ((n - 1) & 124), vis, partial, n & (SUMS - 1), lane);
  }
};

// Fields of record k for degridding; at a new batch, its output codes.
template <int LS>
__device__ __forceinline__ Rec degrid_fields(RunStream<false> &rs, Sums &out, int k)
{
  const Rec f = fields<false, LS>(rs, k, 0);
  if (k < rs.n && (k & 63) == 0 && k > 0) out.batch(rs.code(k, out.npts), k);
  return f;
}

template <int W, bool SQ, int WMAX, int LS>
__device__ __forceinline__ void degrid_run(const TiledDev &d, __amdgpu_buffer_rsrc_t rsrc,
                                           const float2 *reg, int first, int n, int lane,
                                           Sums &out)
{
  using G = Geo<W, SQ, WMAX>;
  const Lane<G, LS> ln(opaque_lane(lane));
  RunStream<false> rs(d.pairs + first, nullptr, n, lane);
  out.batch(rs.code(0, out.npts), 0);
#define fetch(k) degrid_fields<LS>(rs, out, k)
  if constexpr (SQ) {
    // as in grid_run: F records in flight
    constexpr int F = depth<G>();
    typename G::Taps t[F][G::NSM];
    Rec f[F];
#pragma unroll
    for (int j = 0; j < F - 1; j++) {
      f[j] = fetch(j);
      issue_sq<G>(t[j], f[j], ln, rsrc);
    }
    for (int k = 0; k < n; k += F) {
#pragma unroll
      for (int j = 0; j < F; j++) {
        const int nx = (j + F - 1) % F;
        f[nx] = fetch(k + j + F - 1);
        issue_sq<G>(t[nx], f[nx], ln, rsrc);
        const float2 o = degrid_sq<G, LS>(t[j], f[j], ln, reg);
        if (j == 0 || k + j < n) out.put(k + j, o);
      }
    }
  } else {
    for (int k = 0; k < n; k++) out.put(k, degrid_nsq<G, LS>(fetch(k), ln, reg, rsrc));
  }
#undef fetch
  out.end(n);
}

// Walks the item's runs: f(key, first, count) per non-empty key, in key
// order (first and count held by lane `key`).
template <class F>
__device__ __forceinline__ void for_each_run(int first, int count, F &&f)
{
  unsigned long long m = __ballot(count > 0);
  while (m) {
    const int key = __builtin_ctzll(m);
    m &= m - 1;
    f(key, rl(first, key), rl(count, key));
  }
}

template <int WMAX, class F>
__device__ __forceinline__ void with_key(int key, F &&f)
{
  with_width<WMAX>((key & 31) + 1, [&](auto W) {
    if (key < 32) f(W, std::true_type());
    else f(W, std::false_type());
  });
}

// LDS of one gridding wavefront: the region copy.
template <int T, int H>
struct GridLDS {
  float2 reg[Region<T, H>::NC];
};

template <int T, int H>
__global__ void __launch_bounds__(64)
k_grid_halo(TiledDev d, const float2 *__restrict__ vis, const float2 *__restrict__ taps,
            long long n_taps, float2 *__restrict__ grid)
{
  using G = Region<T, H>;
  constexpr int WMAX = H + 1;
  __shared__ GridLDS<T, H> sh;
  float2 *reg = sh.reg;
  const __amdgpu_buffer_rsrc_t rsrc = tap_rsrc(taps, n_taps);
  const int lane = threadIdx.x;
  const int total = d.totals[1];
  int tried = 0;
  for (int item = next_item(d.queue, total, tried, lane); item >= 0;
       item = next_item(d.queue, total, tried, lane)) {
    const ItemRange ir = item_of(d, item);
    for (int q = lane; q < G::NC; q += 64) reg[q] = make_float2(0.f, 0.f);
    int first, count;
    item_runs(d, item, lane, first, count);
    for_each_run(first, count, [&](int key, int s, int c) {
      with_key<WMAX>(key, [&](auto W, auto SQ) {
        grid_run<decltype(W)::value, decltype(SQ)::value, WMAX, G::LS>(d, rsrc, reg,
                                                                       ir.begin + s, c, lane);
      });
    });
    // flush: the region to the item's buffer, summed into the grid by
    // k_gather_halo; items past the buffers add their non-zero cells with
    // atomics (neighbouring regions overlap).  Cells past the grid are never
    // written by an in-bound window.
    if (ir.item < d.ibuf_cap) {
      float2 *ob = d.ibuf + (long long)ir.item * (G::RW * G::RW);
      for (int q = lane; q < G::RW * G::RW; q += 64) {
        const int r = q / G::RW, cc = q - r * G::RW;
        ob[q] = reg[r * G::LS + cc];
      }
      continue;
    }
    const int gsz = d.grid_size;
    float2 *gi = grid + ir.goff;
    for (int q = lane; q < G::RW * G::RW; q += 64) {
      const int r = q / G::RW, cc = q - r * G::RW;
      const int gu = ir.tu0 + r, gv = ir.tv0 + cc;
      if (gu >= gsz || gv >= gsz) continue;
      const float2 v = reg[r * G::LS + cc];
      if (v.x == 0.f && v.y == 0.f) continue;
      float2 *g = gi + (long long)gu * d.pitch + gv;
      if (d.lo) {
        // compensated flush: the exact rounding error of each atomic add
        // (old + x, as the atomic computed it) goes to a second grid
        const float ox = unsafeAtomicAdd(&g->x, v.x);
        const float oy = unsafeAtomicAdd(&g->y, v.y);
        float2 *l = d.lo + ir.goff + (long long)gu * d.pitch + gv;
        const float ex = two_sum_err(ox, v.x), ey = two_sum_err(oy, v.y);
        if (ex != 0.f) unsafeAtomicAdd(&l->x, ex);
        if (ey != 0.f) unsafeAtomicAdd(&l->y, ey);
      } else {
        unsafeAtomicAdd(&g->x, v.x);
        unsafeAtomicAdd(&g->y, v.y);
      }
    }
  }
}

// Sum of the item buffers into the grid: one workgroup per tile (T x T
// threads, one cell each).  Cell (u, v) of the tile lies in the regions of
// the tiles up to NB - 1 rows above and columns left of it; each of their
// items holds it at (u - row0, v - col0) of its buffer.  The sum runs in
// FP64 over the items in order, so it is reproducible, and is rounded once
// and added to the grid's cell (zero, or the atomics of items past the
// buffers).
template <int T, int H>
__global__ void __launch_bounds__(T * T)
k_gather_halo(TiledDev d, float2 *__restrict__ grid)
{
  using G = Region<T, H>;
  constexpr int NB = (G::RW + T - 1) / T;
  const int t = blockIdx.x;
  const int inst = t / d.tpi, lt = t - inst * d.tpi;
  const int ti = lt / d.ntx, tj = lt - ti * d.ntx;
  const int r = threadIdx.x / T, c = threadIdx.x - (threadIdx.x / T) * T;
  const int gu = ti * T + r, gv = tj * T + c;
  if (gu >= d.grid_size || gv >= d.grid_size) return;
  const int n_items = d.totals[1];
  double sx = 0.0, sy = 0.0;
  bool any = false;
  for (int a = 0; a < NB; a++) {
    const int si = ti - a, rr = a * T + r;
    if (si < 0 || rr >= G::RW) continue;
    for (int b = 0; b < NB; b++) {
      const int sj = tj - b, cc = b * T + c;
      if (sj < 0 || cc >= G::RW) continue;
      const int st = inst * d.tpi + si * d.ntx + sj;
      const int i0 = d.item_start[st];
      const int i1 = min(st + 1 < d.n_tiles ? d.item_start[st + 1] : n_items, d.ibuf_cap);
      for (int it = i0; it < i1; it++) {
        const float2 v = d.ibuf[(long long)it * (G::RW * G::RW) + rr * G::RW + cc];
        sx += v.x;
        sy += v.y;
        any = true;
      }
    }
  }
  if (!any) return;
  float2 *g = grid + (long long)inst * d.gstride + (long long)gu * d.pitch + gv;
  float2 o = *g;
  o.x += (float)sx;
  o.y += (float)sy;
  *g = o;
}

// LDS of one degridding wavefront: the staged region, the per-record sums
// and output codes.
template <int T, int H>
struct DegridLDS {
  float2 reg[Region<T, H>::NC];
  float2 sums[SUMS * 64];
  int codes[128];
};

template <int T, int H>
__global__ void __launch_bounds__(64)
k_degrid_halo(TiledDev d, const float2 *__restrict__ grid, const float2 *__restrict__ taps,
              long long n_taps, float2 *__restrict__ vis)
{
  using G = Region<T, H>;
  constexpr int WMAX = H + 1;
  __shared__ DegridLDS<T, H> sh;
  const __amdgpu_buffer_rsrc_t rsrc = tap_rsrc(taps, n_taps);
  const int lane = threadIdx.x;
  Sums out{sh.sums, sh.codes, vis, d.partial, (int)d.npts, lane};
  const int total = d.totals[1];
  int tried = 0;
  for (int item = next_item(d.queue + 8, total, tried, lane); item >= 0;
       item = next_item(d.queue + 8, total, tried, lane)) {
    const ItemRange ir = item_of(d, item);
    const int gsz = d.grid_size;
    const float2 *gi = grid + ir.goff;
    for (int q = lane; q < G::RW * G::RW; q += 64) {
      const int r = q / G::RW, c = q - r * G::RW;
      const int gu = ir.tu0 + r, gv = ir.tv0 + c;
      sh.reg[r * G::LS + c] = (gu < gsz && gv < gsz) ? gi[(long long)gu * d.pitch + gv]
                                                     : make_float2(0.f, 0.f);
    }
    int first, count;
    item_runs(d, item, lane, first, count);
    for_each_run(first, count, [&](int key, int s, int c) {
      with_key<WMAX>(key, [&](auto W, auto SQ) {
        degrid_run<decltype(W)::value, decltype(SQ)::value, WMAX, G::LS>(
            d, rsrc, sh.reg, ir.begin + s, c, lane, out);
      });
    });
  }
}

// Sorts each work item's records by run key, in place, and writes its run
// table: one wavefront per item (items drawn from one counter), the item's
// records staged in LDS.  Within a key the records keep their order (ranks
// of equal keys among 64 records by ballot, in lane order).
__global__ void __launch_bounds__(64)
k_sort_items(TiledDev d)
{
  __shared__ int4 stage[MAX_HALO_ITEM];
  __shared__ int cur[NKEY];
  const int lane = threadIdx.x;
  const int total = d.totals[1];
  while (true) {
    int item = 0;
    if (lane == 0) item = atomicAdd(&d.queue[16], 1);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= total) break;
    const ItemRange ir = item_of(d, item);
    const int n = ir.end - ir.begin;
    int4 *rp = d.pairs + ir.begin;
    // The barriers order the lanes' LDS accesses for the compiler as well:
    // without them it may hand a lane its own earlier store to cur[] in
    // place of what the other lanes' atomics left there.  (The workgroup is
    // one wavefront, so they cost next to nothing.)
    __syncthreads();
    cur[lane] = 0;
    __syncthreads();
    for (int k = lane; k < n; k += 64) {
      const int4 r = rp[k];
      stage[k] = r;
      atomicAdd(&cur[run_key(r.x)], 1);
    }
    __syncthreads();
    const int count = cur[lane];
    int x = count;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    cur[lane] = x - count;
    __syncthreads();
    if (item < d.item_cap) d.runs[(long long)item * NKEY + lane] = (uint16_t)count;
    for (int k0 = 0; k0 < n; k0 += 64) {
      const int k = k0 + lane;
      const int key = k < n ? run_key(stage[k].x) : NKEY;
      // this lane's rank among the batch's lanes of its key, and the batch's
      // count of each key (read by the key's lowest lane)
      int below = 0, same = 0;
      for (int q = 0; q < 64; q++) {
        const int kq = __shfl(key, q);
        below += (kq == key && q < lane) ? 1 : 0;
        same += kq == key ? 1 : 0;
      }
      int pos = 0;
      if (key < NKEY) pos = cur[key] + below;
      __syncthreads();
      if (key < NKEY && below == 0) cur[key] += same;
      __syncthreads();
      if (key < NKEY) rp[pos] = stage[k];
    }
  }
}

// Visibilities in record order (the records' point indices) for the halo
// gridder: vsort[j] = vis[pairs[j].w].
__global__ void __launch_bounds__(256)
k_gather_vis(TiledDev d, const float2 *__restrict__ vis)
{
  const int total = d.totals[0];
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x)
    d.vsort[j] = vis[d.pairs[j].w];
}

template <int T, int H>
hipError_t blocks_of(int *gb, int *db)
{
  int dev = 0, cus = 0, g = 0, dg = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &g, reinterpret_cast<const void *>(&k_grid_halo<T, H>), 64, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &dg, reinterpret_cast<const void *>(&k_degrid_halo<T, H>), 64, 0);
  if (e != hipSuccess) return e;
  // a multiple of 8: one share per XCD
  *gb = std::max(8, (std::max(1, g) * std::max(1, cus)) & ~7);
  *db = std::max(8, (std::max(1, dg) * std::max(1, cus)) & ~7);
  return hipSuccess;
}

template <int T, int H>
void launch_grid(const TiledDev &d, int blocks, const float2 *vis, const float2 *taps,
                 long long n_taps, float2 *grid, hipStream_t st)
{
  hipLaunchKernelGGL(k_gather_vis, dim3(2048), dim3(256), 0, st, d, vis);
  hipLaunchKernelGGL((k_grid_halo<T, H>), dim3(blocks), dim3(64), 0, st, d, vis, taps, n_taps,
                     grid);
  hipLaunchKernelGGL((k_gather_halo<T, H>), dim3(d.n_tiles), dim3(T * T), 0, st, d, grid);
}

template <int T, int H>
void launch_degrid(const TiledDev &d, int blocks, const float2 *grid, const float2 *taps,
                   long long n_taps, float2 *vis, hipStream_t st)
{
  hipLaunchKernelGGL((k_degrid_halo<T, H>), dim3(blocks), dim3(64), 0, st, d, grid, taps,
                     n_taps, vis);
}

}  // namespace

rcg_status halo_sort_items(const TiledDev &d, hipStream_t st)
{
  int dev = 0, cus = 0;
  RCG_HIP_TRY(hipGetDevice(&dev));
  RCG_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  hipLaunchKernelGGL(k_sort_items, dim3(cus * 8), dim3(64), 0, st, d);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

rcg_status halo_blocks(int T, int wmax, int *gb, int *db)
{
  hipError_t e;
  if (T == 16) e = wmax == 16 ? blocks_of<16, 15>(gb, db) : blocks_of<16, 31>(gb, db);
  else e = wmax == 16 ? blocks_of<32, 15>(gb, db) : blocks_of<32, 31>(gb, db);
  RCG_HIP_TRY(e);
  return RCG_OK;
}

rcg_status halo_grid(const TiledDev &d, int blocks, const float2 *vis, const float2 *taps,
                     long long n_taps, float2 *grid, hipStream_t st)
{
  if (d.T == 16) {
    if (d.wmax == 16) launch_grid<16, 15>(d, blocks, vis, taps, n_taps, grid, st);
    else launch_grid<16, 31>(d, blocks, vis, taps, n_taps, grid, st);
  } else {
    if (d.wmax == 16) launch_grid<32, 15>(d, blocks, vis, taps, n_taps, grid, st);
    else launch_grid<32, 31>(d, blocks, vis, taps, n_taps, grid, st);
  }
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

rcg_status halo_degrid(const TiledDev &d, int blocks, const float2 *grid, const float2 *taps,
                       long long n_taps, float2 *vis, hipStream_t st)
{
  if (d.T == 16) {
    if (d.wmax == 16) launch_degrid<16, 15>(d, blocks, grid, taps, n_taps, vis, st);
    else launch_degrid<16, 31>(d, blocks, grid, taps, n_taps, vis, st);
  } else {
    if (d.wmax == 16) launch_degrid<32, 15>(d, blocks, grid, taps, n_taps, vis, st);
    else launch_degrid<32, 31>(d, blocks, grid, taps, n_taps, vis, st);
  }
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

}  // namespace rcg
