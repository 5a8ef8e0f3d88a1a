// rcg_tiled.hip -- FP32 gridder/degridder over uv tiles held in LDS.
//
// Same mathematics as the reference's scatter gridder and its transpose
// degridder (MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p
// .cpp:146-176), organised for CDNA4.  A plan covers one or more instances
// (snapshots, facets): each has its own points and geometry and its own
// grid of the common shape; all share one GCF table.
//
//  1. pregrid (FP64, exact; rcg_pregrid.hip) -> one record per point;
//  2. binning: every grid is cut into T x T tiles; every in-bound point is
//     listed in each tile its S x S window overlaps (a point near a tile
//     edge appears in up to four tiles when S <= T).  A counting sort builds
//     the per-tile lists without host synchronisation: LDS histograms per
//     workgroup (each workgroup holds points of one instance) -> global
//     counts -> three-kernel scan -> LDS-ranked fill;
//  3. work items: each tile's list is cut into chunks of at most `chunk`
//     pairs.  The grid and degrid kernels are persistent: a fixed number of
//     workgroups draw items from eight queues, one per XCD, each holding a
//     contiguous range of items (neighbouring tiles share that XCD's L2),
//     and steal from the other queues when theirs is empty;
//  4. gridding: each wavefront accumulates its pairs' taps into a private
//     FP32 copy of the tile in LDS, lanes covering the rows and columns of
//     the window/tile intersection, GCF rows read coalesced; the copies
//     are summed and stored to HBM with plain stores when the tile has a
//     single item, with FP32 atomics (non-zero cells only) otherwise;
//  5. degridding: the workgroup stages its grid tile in LDS, each wavefront
//     forms per-lane partial sums of one pair over the intersection and
//     reduces them across lanes four pairs at a time; points whose window
//     lies in one tile are stored, split points leave one partial per tile,
//     summed in a fixed order by k_degrid_reduce.
#include <algorithm>
#include <climits>

#include "rcg_internal.h"
#include "rcg_tiled.h"
#include "rcg_tiledev.h"

namespace rcg {

// Pair record: one (point, tile) incidence with the geometry of the
// window/tile intersection precomputed, 16 bytes.
//   x: r0 | c0 << 6 | (w-1) << 12 | (rows-1) << 18 | slot << 24 | conj << 30
//      | whole << 31
//      r0, c0: first tile row/column of the intersection (< 64)
//      w:      intersection width (1..64)
//      rows:   64 / w, the rows one 64-lane slab covers
//      slot:   index of this tile among the point's tiles (row-major)
//      whole:  the window lies entirely in this tile
//   y: tap offset (elements) of the intersection's first cell
//   z: stride | h << 16 | nslabs << 23
//      stride: layer row stride (elements, < 65536)
//      h:      intersection height (1..64)
//      nslabs: 64-lane slabs of `rows` rows that cover it, ceil(h / rows)
//   w: point index
__device__ __forceinline__ int4 make_pair(int u0, int v0, const rcg_pg &r, int point,
                                          int tile_r, int tile_c, int T, int slot)
{
  const int S = r.supp, stride = r.stride;
  const int tu0 = tile_r * T, tv0 = tile_c * T;
  const int ra = max(u0, tu0), rb = min(u0 + S, tu0 + T);
  const int ca = max(v0, tv0), cb = min(v0 + S, tv0 + T);
  const int h = rb - ra, w = cb - ca;
  // 64 / w and ceil(h / rows) through the approximate reciprocal: (x + 0.5)
  // / y lies at least 0.5 / y >= 1/128 from an integer for x < 128, y <= 64,
  // far above the reciprocal's error
  const int rows = (int)(64.5f * __builtin_amdgcn_rcpf((float)w));
  const int nslabs = (int)(((float)(h + rows - 1) + 0.5f) * __builtin_amdgcn_rcpf((float)rows));
  const int whole = (h == S) && (w == S);
  int4 p;
  p.x = (ra - tu0) | ((ca - tv0) << 6) | ((w - 1) << 12) | ((rows - 1) << 18) |
        (slot << 24) | ((r.flags & PG_CONJ) ? (1 << 30) : 0) |
        (whole ? (int)0x80000000 : 0);
  p.y = r.tap0 + (ra - u0) * stride + (ca - v0);
  p.z = stride | (h << 16) | (nslabs << 23);
  p.w = point;
  return p;
}

// ---------------------------------------------------------------------------
// small utility kernels
__global__ void k_fill_i32(int *__restrict__ p, int v, long long n)
{
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_f32(float *__restrict__ p, long long n)
{
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n4 = n >> 2;
  float4 *p4 = reinterpret_cast<float4 *>(p);
  for (long long k = i; k < n4; k += (long long)gridDim.x * blockDim.x)
    p4[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (long long k = (n4 << 2) + i; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = 0.f;
}

// ---------------------------------------------------------------------------
// binning
//
// Each workgroup (BIN_THREADS threads) takes ppt consecutive points per
// thread, all of one instance: instance k owns workgroups
// [blk_off[k], blk_off[k + 1]).  Tile ranges of a point: rows
// [u0 >> lt, (u0 + S - 1) >> lt], columns likewise; plan tile index =
// instance * tpi + row * ntx + column.

constexpr int BIN_THREADS = 1024;

// u0 = cell / pitch without an integer division: (cell + 0.5) / pitch lies
// at least 0.5 / pitch (>= 2^-16) from an integer, and the FP64 product
// (cell < 2^31) is within 2^-21 of it.
__device__ __forceinline__ void point_tiles(const rcg_pg &r, int pitch, double inv_pitch, int lt,
                                            int *u0, int *v0, int *tu0,
                                            int *tu1, int *tv0, int *tv1)
{
  *u0 = (int)(((double)r.cell + 0.5) * inv_pitch);
  *v0 = r.cell - *u0 * pitch;
  const int S = r.supp;
  *tu0 = *u0 >> lt;
  *tu1 = (*u0 + S - 1) >> lt;
  *tv0 = *v0 >> lt;
  *tv1 = (*v0 + S - 1) >> lt;
}

struct BinBlock {
  int inst, p0, p1;  // points [p0, p1) of instance inst
};

__device__ __forceinline__ BinBlock bin_block(const TiledDev &d)
{
  // the largest k with blk_off[k] <= blockIdx.x (instances without points
  // own no workgroup and are skipped)
  int lo = 0, hi = d.n_inst - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.blk_off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  BinBlock r;
  r.inst = lo;
  r.p0 = d.pt_off[lo] + (b - d.blk_off[lo]) * (BIN_THREADS * d.ppt);
  r.p1 = min(d.pt_off[lo + 1], r.p0 + BIN_THREADS * d.ppt);
  return r;
}

// Sub-window k of n along a side of S cells: first cell and length.  The
// lengths are near-equal and even, except that an odd S gives its last
// sub-window the odd cell (the kernels read two columns per lane from
// even widths).  With n = ceil(S / wmax) every length is at most wmax.
__device__ __forceinline__ void sub_range(int S, int n, int k, int *off, int *len)
{
  const int P = S >> 1;  // cell pairs
  const int base = P / n, extra = P - base * n;
  *off = 2 * (k * base + min(k, extra));
  *len = 2 * (base + (k < extra ? 1 : 0)) + ((S & 1) && k == n - 1 ? 1 : 0);
}

// Halo record: one sub-window of a point, binned in the tile that holds its
// first cell, 16 bytes.
//   x: r0 | c0 << 6 | (w-1) << 12 | (h-1) << 18 | slot << 24 | conj << 30
//      | whole << 31
//      r0, c0: the sub-window's first cell, relative to its tile (< T)
//      w, h:   width and height (<= wmax <= 32)
//      slot:   index of the sub-window among the point's (row-major)
//      whole:  the sub-window is the whole window
//   y: tap offset (elements) of the sub-window's first cell
//   z: layer row stride (elements, < 65536)
//   w: point index
// The windows of S <= wmax are whole; larger ones are cut into
// ceil(S / wmax)^2 near-equal sub-windows.
//
// point_records(r, i, d, f) calls f(tile, slot, record) for every record of
// point i in the plan's layout (tile: index within the instance) and returns
// their number, 0 for a point left out, or -1 when there are more than the
// plan's d.mpp.
template <class F>
__device__ __forceinline__ int point_records(const rcg_pg &r, int i, const TiledDev &d, F f)
{
  if (r.cell < 0 || r.supp <= 0) return 0;
  int u0, v0, tu0, tu1, tv0, tv1;
  point_tiles(r, d.pitch, d.inv_pitch, d.lt, &u0, &v0, &tu0, &tu1, &tv0, &tv1);
  if (d.layout == LAYOUT_PAIRS) {
    const int np = (tu1 - tu0 + 1) * (tv1 - tv0 + 1);
    if (np > d.mpp) return -1;
    int j = 0;
    for (int a = tu0; a <= tu1; a++)
      for (int b = tv0; b <= tv1; b++, j++)
        f(a * d.ntx + b, j, make_pair(u0, v0, r, i, a, b, d.T, j));
    return np;
  }
  const int S = r.supp, n = (S + d.wmax - 1) >> d.lw;
  if (n * n > d.mpp) return -1;
  const int conj = (r.flags & PG_CONJ) ? (1 << 30) : 0;
  const int whole = n == 1 ? (int)0x80000000 : 0;
  for (int a = 0; a < n; a++) {
    int ro, h;
    sub_range(S, n, a, &ro, &h);
    const int uu = u0 + ro;
    for (int b = 0; b < n; b++) {
      int co, w;
      sub_range(S, n, b, &co, &w);
      const int vv = v0 + co;
      int4 q;
      q.x = (uu & (d.T - 1)) | ((vv & (d.T - 1)) << 6) | ((w - 1) << 12) | ((h - 1) << 18) |
            ((a * n + b) << 24) | conj | whole;
      q.y = r.tap0 + ro * r.stride + co;
      q.z = r.stride;
      q.w = i;
      f((uu >> d.lt) * d.ntx + (vv >> d.lt), a * n + b, q);
    }
  }
  return n * n;
}

// Counts per tile.  With LDS histograms (tiles per instance <= MAX_BIN_TILES)
// each workgroup adds its histogram with one global atomic per touched tile;
// otherwise every record adds one global atomic.
__global__ void __launch_bounds__(BIN_THREADS)
k_bin_count(const rcg_pg *__restrict__ pg, TiledDev d, int *__restrict__ err)
{
  extern __shared__ int hist[];
  const BinBlock bb = bin_block(d);
  int *tc = d.tile_count + (long long)bb.inst * d.tpi;
  if (!d.gbins) {
    for (int t = threadIdx.x; t < d.tpi; t += BIN_THREADS) hist[t] = 0;
    __syncthreads();
  }
  for (int k = 0; k < d.ppt; k++) {
    const int i = bb.p0 + k * BIN_THREADS + (int)threadIdx.x;
    if (i >= bb.p1) break;
    const rcg_pg r = pg[i];
    const int nr = point_records(r, i, d, [&](int t, int, const int4 &) {
      atomicAdd(d.gbins ? &tc[t] : &hist[t], 1);
    });
    if (nr < 0) atomicOr(err, ERR_CAPACITY);  // support above the plan's max_support
  }
  if (d.gbins) return;
  __syncthreads();
  for (int t = threadIdx.x; t < d.tpi; t += BIN_THREADS)
    if (hist[t]) atomicAdd(&tc[t], hist[t]);
}

// Scan of the tile counts (pairs) and per-tile item counts in three
// kernels over segments of SCAN_SEG tiles: segment sums, one workgroup
// scanning the sums, per-segment exclusive scans.  totals[0] = pairs,
// totals[1] = items; the fill cursors are cleared.
constexpr int SCAN_SEG = 1024;

// Exclusive scan of (x, y) over a 1024-thread workgroup; *tot = the sums.
__device__ __forceinline__ int2 block_scan(int2 v, int2 *sh, int2 *tot)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int2 x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int ax = __shfl_up(x.x, off), ay = __shfl_up(x.y, off);
    if (lane >= off) {
      x.x += ax;
      x.y += ay;
    }
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  if (threadIdx.x < 64) {
    int2 w = lane < 16 ? sh[lane] : make_int2(0, 0);
    for (int off = 1; off < 16; off <<= 1) {
      const int ax = __shfl_up(w.x, off), ay = __shfl_up(w.y, off);
      if (lane >= off) {
        w.x += ax;
        w.y += ay;
      }
    }
    if (lane < 16) sh[lane] = w;
  }
  __syncthreads();
  const int2 base = wave ? sh[wave - 1] : make_int2(0, 0);
  *tot = sh[15];
  return make_int2(base.x + x.x - v.x, base.y + x.y - v.y);
}

__global__ void __launch_bounds__(SCAN_SEG)
k_scan_reduce(TiledDev d)
{
  __shared__ int2 sh[16];
  const int t = blockIdx.x * SCAN_SEG + threadIdx.x;
  const int c = t < d.n_tiles ? d.tile_count[t] : 0;
  int2 tot;
  block_scan(make_int2(c, (c + d.chunk - 1) / d.chunk), sh, &tot);
  if (threadIdx.x == 0) {
    d.seg[2 * blockIdx.x] = tot.x;
    d.seg[2 * blockIdx.x + 1] = tot.y;
  }
}

__global__ void __launch_bounds__(SCAN_SEG)
k_scan_top(TiledDev d, int n_seg)
{
  __shared__ int2 sh[16];
  const int s = threadIdx.x;
  const int2 v = s < n_seg ? make_int2(d.seg[2 * s], d.seg[2 * s + 1]) : make_int2(0, 0);
  int2 tot;
  const int2 e = block_scan(v, sh, &tot);
  if (s < n_seg) {
    d.seg[2 * s] = e.x;
    d.seg[2 * s + 1] = e.y;
  }
  if (s == 0) {
    d.totals[0] = tot.x;
    d.totals[1] = tot.y;
  }
}

__global__ void __launch_bounds__(SCAN_SEG)
k_scan_apply(TiledDev d)
{
  __shared__ int2 sh[16];
  const int t = blockIdx.x * SCAN_SEG + threadIdx.x;
  const int c = t < d.n_tiles ? d.tile_count[t] : 0;
  int2 tot;
  const int2 e = block_scan(make_int2(c, (c + d.chunk - 1) / d.chunk), sh, &tot);
  if (t < d.n_tiles) {
    d.tile_start[t] = d.seg[2 * blockIdx.x] + e.x;
    d.item_start[t] = d.seg[2 * blockIdx.x + 1] + e.y;
    d.tile_cursor[t] = 0;
  }
}

// Fill: LDS counts per tile within the workgroup, one global atomic per
// touched tile for the workgroup's range of slots, then every record takes
// the next slot of its tile's range with an LDS atomic (records of a tile
// land in any order, as with global atomics); with global bins, one global
// atomic per record gives its slot.  Only the histogram lives in LDS, so two
// workgroups share a CU (a per-record rank array, 73 KB at 4 points per
// thread and 9 records per point, had left room for one).
__global__ void __launch_bounds__(BIN_THREADS)
k_bin_fill(const rcg_pg *__restrict__ pg, TiledDev d, int *__restrict__ err)
{
  extern __shared__ int hist[];  // tpi
  const BinBlock bb = bin_block(d);
  const long long t0 = (long long)bb.inst * d.tpi;
  auto put = [&](long long pos, const int4 &q) {
    if (pos < d.pair_cap) d.pairs[pos] = q;
    else atomicOr(err, ERR_CAPACITY);
  };
  if (d.gbins) {
    for (int k = 0; k < d.ppt; k++) {
      const int i = bb.p0 + k * BIN_THREADS + (int)threadIdx.x;
      if (i >= bb.p1) break;
      const rcg_pg r = pg[i];
      point_records(r, i, d, [&](int t, int, const int4 &q) {
        put((long long)d.tile_start[t0 + t] + atomicAdd(&d.tile_cursor[t0 + t], 1), q);
      });
    }
    return;
  }
  for (int t = threadIdx.x; t < d.tpi; t += BIN_THREADS) hist[t] = 0;
  __syncthreads();
  for (int k = 0; k < d.ppt; k++) {
    const int i = bb.p0 + k * BIN_THREADS + (int)threadIdx.x;
    if (i >= bb.p1) break;
    const rcg_pg r = pg[i];
    point_records(r, i, d, [&](int t, int, const int4 &) { atomicAdd(&hist[t], 1); });
  }
  __syncthreads();
  for (int t = threadIdx.x; t < d.tpi; t += BIN_THREADS) {
    const int c = hist[t];
    hist[t] = c ? d.tile_start[t0 + t] + atomicAdd(&d.tile_cursor[t0 + t], c) : 0;
  }
  __syncthreads();
  for (int k = 0; k < d.ppt; k++) {
    const int i = bb.p0 + k * BIN_THREADS + (int)threadIdx.x;
    if (i >= bb.p1) break;
    const rcg_pg r = pg[i];
    point_records(r, i, d, [&](int t, int, const int4 &q) { put(atomicAdd(&hist[t], 1), q); });
  }
}

// ---------------------------------------------------------------------------
// Walking a pair's window/tile intersection.
//
// Per pair: the lanes form `rows` rows of `w` columns (lane l at row
// l / w, column l % w; lanes past rows*w idle); slab s covers intersection
// rows [s*rows, (s+1)*rows), the last one possibly partial.  A lane's tap
// offset for slab s is its slab-0 offset plus the uniform s*rows*stride (the
// buffer load's scalar offset); its LDS byte address advances by rows*LS*8
// per slab.  A lane outside the intersection in a slab loads its tap at an
// out-of-range offset (which reads zero) and points its LDS access at the
// copy's dummy cell.  The pair fields come from readlane and stay in scalar
// registers: nothing below selects between two of them under a vector mask
// (a VALU instruction reads at most one scalar operand), which would move
// them, and every loop-carried copy of them, into vector registers.
//
// Latency: a wavefront loads the records of 64 pairs at once, one per lane,
// two batches ahead (the gridder's visibilities one batch ahead,
// batch_records), and broadcasts them with readlane; the taps of a pair's first group of up to GRP slabs are
// loaded while the previous pair's LDS updates run.  Two slabs per group
// balance the loads kept in flight against the slots a short pair leaves
// empty (measured on the 4096^2 w-projection workload, T = 32: 3 slabs are
// 3-6% slower, 4 slabs 10-15%, 1 slab 13-17%).
constexpr int GRP = 2;

// Wavefronts per gridding workgroup (1..4, a plan's choice through
// rcg_tiled_set_grid_waves; the autotuner searches it).  Each holds a
// private copy of the tile in LDS, so the copies' size sets how many
// workgroups share a CU's 160 KB: at T = 32 three wavefronts (25 KB) let six
// workgroups (18 wavefronts) in, against four of four (16).  The default
// below is the plan's choice until it is set: on the 4096^2 w-projection
// workload 2, 3 and 4 wavefronts are within 1-2% of each other at T = 32,
// one 3-5% slower; at T = 64 three copies leave one workgroup per CU.
template <int T>
constexpr int grid_waves()
{
  return T == 32 ? 3 : 4;
}
// Wavefronts per gridding workgroup a plan may choose (rcg_tiled_set_grid_waves).
constexpr int MAX_GRID_WAVES = 4;

struct Pair {
  int h, rows, nslabs, point, whole, conj, slot;
  int lr;          // lane row (idle lanes: 64)
  unsigned tb0;    // tap byte offset of the lane's slab-0 cell (idle lanes: unused)
  int lb0;         // LDS byte address of the lane's slab-0 cell
  int tstep;       // tap byte step per slab (uniform)
  int lstep;       // LDS byte step per slab (uniform)
};

// base: LDS byte address of the tile copy (row-major, row stride LS)
template <int LS>
__device__ __forceinline__ Pair decode_pair(int px, int py, int pz, int pw, int lane, int base)
{
  Pair p;
  const int stride = pz & 0xffff;
  p.point = pw;
  const int r0 = px & 63, c0 = (px >> 6) & 63;
  const int w = ((px >> 12) & 63) + 1;
  p.rows = ((px >> 18) & 63) + 1;
  p.h = (pz >> 16) & 127;
  p.nslabs = (pz >> 23) & 127;
  p.conj = (px >> 30) & 1;
  p.whole = px < 0;
  p.slot = (px >> 24) & 63;
  // lane / w through the approximate reciprocal: (lane + 0.5) / w is at
  // least 0.5 / w away from an integer, far above its rounding error
  const int lr = (int)(((float)lane + 0.5f) * __builtin_amdgcn_rcpf((float)w));
  // 24-bit multiplies (full rate): lr < 64, w <= 64, stride < 2^16
  const int lc = lane - __mul24(lr, w);
  const bool idle = lr >= p.rows;
  p.lr = idle ? 64 : lr;
  // idle lanes (lr = 64) fail every slab's row test in load_group, which
  // then sends their loads out of range, so tb0 needs no select here
  p.tb0 = (unsigned)(py + __mul24(lr, stride) + lc) * 8u;
  p.lb0 = base + (__mul24(r0 + lr, LS) + c0 + lc) * 8;
  p.tstep = p.rows * stride * 8;
  p.lstep = p.rows * LS * 8;
  return p;
}

// Taps and LDS byte addresses of slabs [s0, s0 + GRP) of pair p: outside
// the intersection the tap reads zero and the address is `dummy`.  One
// compare per slab serves both selects.
__device__ __forceinline__ void load_group(float2 (&t)[GRP], int (&la)[GRP], const Pair &p,
                                           int s0, int dummy, __amdgpu_buffer_rsrc_t rsrc)
{
#pragma unroll
  for (int u = 0; u < GRP; u++) {
    const bool act = p.lr < p.h - (s0 + u) * p.rows;
    t[u] = tap_load(rsrc, act ? p.tb0 : TAP_OOB, (s0 + u) * p.tstep);
    la[u] = act ? p.lb0 + (s0 + u) * p.lstep : dummy;
  }
}

// ---------------------------------------------------------------------------
// gridding: each wavefront owns a private copy of the tile in LDS, so the
// read-modify-write of a cell needs no atomic: within one pair every lane
// and slab has its own cell (lanes outside the intersection only ever add
// zero to the dummy cell), and a wavefront's LDS operations complete in
// program order; the copies are summed at the flush.

template <int T, int GW>
__global__ void __launch_bounds__(64 * GW)
k_grid_tiles(TiledDev d, const float2 *__restrict__ vis,
             const float2 *__restrict__ taps, long long n_taps,
             float2 *__restrict__ grid)
{
  constexpr int LS = T + 1;
  constexpr int CP = T * LS + 2;  // one copy (+ dummy cell)
  constexpr int DUMMY = T * LS;
  __shared__ float2 acc[GW * CP];
  const __amdgpu_buffer_rsrc_t rsrc = tap_rsrc(taps, n_taps);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // LDS byte addresses of the wavefront's copy and of its dummy cell; the
  // dummy address lives in a vector register (vmove) so that the address
  // selects, whose condition is a scalar mask, need no copy of it per group
  // (measured: gridder kernel 2.83 -> 2.80 ms on the 4096^2 w-projection
  // workload)
  const int base = wave * CP * 8, dummy = vmove((wave * CP + DUMMY) * 8);
  for_each_item(d, d.queue, [&](const ItemRange &ir) {
    for (int q = threadIdx.x; q < GW * CP; q += 64 * GW) acc[q] = make_float2(0.f, 0.f);
    __syncthreads();
    const int npw = ir.end - ir.begin > wave ? (ir.end - ir.begin - wave + GW - 1) / GW : 0;
    if (npw > 0) {
      // records of the current batch and the next two; visibilities of the
      // current batch and the next
      int4 rec = batch_records(d, ir, wave, lane, npw, 0, GW);
      int4 rec_n = batch_records(d, ir, wave, lane, npw, 1, GW);
      int4 rec_nn = batch_records(d, ir, wave, lane, npw, 2, GW);
      float2 mv = vis[rec.w], mv_n = vis[rec_n.w];
      // Taps are conjugated for mirrored (w < 0) layers and, in the adjoint,
      // once more for all of them: a pair's conj field becomes, once per
      // pair, the sign mask of the FMA coefficients b and c (0 when the two
      // conjugations cancel).
      auto flip_of = [&](int conj) { return conj != d.adj ? 0 : (int)0x80000000; };
      // The group after (q, sq): the same pair's next slabs, or the next
      // pair's first; past the end, the empty pair (no lane active).
      auto advance = [&](Pair &q, int &sq, int &kq, float &qx, float &qy) {
        sq += GRP;
        if (sq >= q.nslabs) {
          kq++;
          sq = 0;
          if (kq < npw) {
            const int j = kq & 63;
            if (j == 0) {
              // the moves first (vmove), then the loads into the registers
              // they free; rec.w is not used past this point
              rec = make_int4(vmove(rec_n.x), vmove(rec_n.y), vmove(rec_n.z), 0);
              mv = make_float2(vmovef(mv_n.x), vmovef(mv_n.y));
              rec_n = make_int4(vmove(rec_nn.x), vmove(rec_nn.y), vmove(rec_nn.z),
                                vmove(rec_nn.w));
              __builtin_amdgcn_sched_barrier(0);
              mv_n = vis[rec_n.w];
              rec_nn = batch_records(d, ir, wave, lane, npw, (kq >> 6) + 2, GW);
            }
            q = decode_pair<LS>(rl(rec.x, j), rl(rec.y, j), rl(rec.z, j), rl(rec.w, j), lane,
                                base);
            q.conj = flip_of(q.conj);
            qx = rlf(mv.x, j);
            qy = rlf(mv.y, j);
          } else {
            q.h = 0;
          }
        }
      };
      // Taps are issued two groups ahead: the current group's were issued two
      // groups ago, the next group's one group ago.  cur: the current
      // group's visibility, conjugation and pair index; (pn, sn, kn): the
      // next group, whose successor this group issues.
      Pair pn = decode_pair<LS>(rl(rec.x, 0), rl(rec.y, 0), rl(rec.z, 0), rl(rec.w, 0), lane,
                                base);
      pn.conj = flip_of(pn.conj);
      int sn = 0, kn = 0;
      float vxn = rlf(mv.x, 0), vyn = rlf(mv.y, 0);
      float2 ta[GRP], tb[GRP], tc3[GRP];
      int la[GRP], lb[GRP], lc3[GRP];
      load_group(ta, la, pn, sn, dummy, rsrc);
      float vx = vxn, vy = vyn;
      int cj = pn.conj, kc = 0;
      advance(pn, sn, kn, vxn, vyn);
      load_group(tb, lb, pn, sn, dummy, rsrc);
      bool more = true;
      // One group: issue the taps of the group after the next one into t2,
      // then update the current group's cells from t0 while they fly.
      // Called with the three buffers in rotation, so no register copy
      // forces a wait on the loads in flight.
      auto body = [&](float2 (&t0)[GRP], int (&l0)[GRP], float2 (&t2)[GRP], int (&l2)[GRP]) {
        // The current group's cells, read first so that the LDS latency
        // overlaps the choice of the group after next.  A wavefront's LDS
        // operations complete in order, so the reads see the previous
        // group's writes; the cells of one group are distinct, so all its
        // reads may precede its writes.
        float2 x[GRP];
#pragma unroll
        for (int u = 0; u < GRP; u++) x[u] = lds_at(acc, l0[u]);
        Pair pnn = pn;
        int snn = sn, knn = kn;
        float vxnn = vxn, vynn = vyn;
        advance(pnn, snn, knn, vxnn, vynn);
        load_group(t2, l2, pnn, snn, dummy, rsrc);
        __builtin_amdgcn_sched_barrier(0);
        // b = cj ? vy : -vy and c = cj ? -vx : vx as sign-bit flips
        // (scalar integer ops) with cj's mask from flip_of
        const unsigned flip = (unsigned)cj;
        const float a = vx, e = vy;
        const float b = __int_as_float(__float_as_int(vy) ^ flip);
        const float c = __int_as_float(__float_as_int(vx) ^ flip ^ 0x80000000u);
#pragma unroll
        for (int u = 0; u < GRP; u++) {
          x[u].x = fmaf(a, t0[u].x, fmaf(b, t0[u].y, x[u].x));
          x[u].y = fmaf(c, t0[u].y, fmaf(e, t0[u].x, x[u].y));
        }
#pragma unroll
        for (int u = 0; u < GRP; u++) lds_at(acc, l0[u]) = x[u];
        vx = vxn;
        vy = vyn;
        cj = pn.conj;
        kc = kn;
        pn = pnn;
        sn = snn;
        kn = knn;
        vxn = vxnn;
        vyn = vynn;
        more = kc < npw;
      };
      // The three bodies run on every trip (past the end they work on the
      // empty pair and touch only the dummy cell): a loop exit between them
      // would join the back edge with undefined state, which costs vector
      // copies on every group.
      do {
        body(ta, la, tc3, lc3);
        body(tb, lb, ta, la);
        body(tc3, lc3, tb, lb);
      } while (more);
      // Never true (chunk >= 1): a use of both buffers on the loop exits
      // keeps the compiler from sinking the prefetch loads past the exit
      // tests.
      if (d.chunk < 0) {
        float2 z = make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < GRP; u++) {
          z.x += ta[u].x + tb[u].x + tc3[u].x;
          z.y += ta[u].y + tb[u].y + tc3[u].y;
        }
        acc[wave * CP + lane] = z;
      }
    }
    __syncthreads();
    // flush: sum the copies in order
    const int gsz = d.grid_size;
    float2 *gi = grid + ir.goff;
    for (int q = threadIdx.x; q < T * T; q += 64 * GW) {
      const int r = q / T, cc = q - (q / T) * T;
      const int gu = ir.tu0 + r, gv = ir.tv0 + cc;
      if (gu >= gsz || gv >= gsz) continue;
      const int l = r * LS + cc;
      float re = acc[l].x, im = acc[l].y;
#pragma unroll
      for (int c = 1; c < GW; c++) {
        re += acc[c * CP + l].x;
        im += acc[c * CP + l].y;
      }
      float2 *g = gi + (long long)gu * d.pitch + gv;
      if (ir.single) {
        *g = make_float2(re, im);
      } else if (re != 0.f || im != 0.f) {
        if (d.lo) {
          // compensated flush: the exact rounding error of each atomic
          // add (old + x, as the atomic computed it) goes to a second grid
          const float ox = unsafeAtomicAdd(&g->x, re);
          const float oy = unsafeAtomicAdd(&g->y, im);
          float2 *l = d.lo + ir.goff + (long long)gu * d.pitch + gv;
          const float ex = two_sum_err(ox, re), ey = two_sum_err(oy, im);
          if (ex != 0.f) unsafeAtomicAdd(&l->x, ex);
          if (ey != 0.f) unsafeAtomicAdd(&l->y, ey);
        } else {
          unsafeAtomicAdd(&g->x, re);
          unsafeAtomicAdd(&g->y, im);
        }
      }
    }
  });
}

// ---------------------------------------------------------------------------
// FP64 gridding over the same pairs (the gridder of the MS4 host-memory
// entry points).  Lanes cover a pair's window/tile intersection as in
// decode_pair, one slab row per trip.  Each tap's contribution is the
// reference's complex product (scatter_gridder_w_dependent_dyn_1p.cpp:169)
// in FP64 without contraction, the same value the plain FP64 gridder adds;
// only the order of the sums differs.  The workgroup's wavefronts share one
// FP64 copy of the tile and add into it with LDS FP64 atomics (ds_add_f64):
// at T = 32 a copy is 17 KB, so nine workgroups fit a CU.  Measured on the
// 4096^2 w-projection points in 32640 chunks of 306 (1.35 G taps): 2.9 ms,
// against 8.6 ms with a private copy per wavefront (two workgroups per CU)
// and 3.0-4.9 ms with the taps of 2-8 slabs loaded together.  A tap index
// past the table reads nothing and adds nothing.  add: the tile sums are
// added to the grid's contents (otherwise the grid was zeroed and a tile
// with a single work item is stored).

constexpr int F64_WAVES = 4;  // 2: 3.7 ms, 8: 3.2-3.5 ms on the workload above

template <int T>
__global__ void __launch_bounds__(64 * F64_WAVES)
k_grid_tiles_f64(TiledDev d, const double2 *__restrict__ vis,
                 const double2 *__restrict__ taps, long long n_taps,
                 double2 *__restrict__ grid, int add)
{
#pragma clang fp contract(off)
  constexpr int LS = T + 1;
  constexpr int CP = T * LS;
  __shared__ double2 acc[CP];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for_each_item(d, d.queue, [&](const ItemRange &ir) {
    for (int q = threadIdx.x; q < CP; q += 64 * F64_WAVES) acc[q] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int k = ir.begin + wave; k < ir.end; k += F64_WAVES) {
      const int4 rec = d.pairs[k];
      const int r0 = rec.x & 63, c0 = (rec.x >> 6) & 63;
      const int w = ((rec.x >> 12) & 63) + 1;
      const int rows = ((rec.x >> 18) & 63) + 1;
      const int stride = rec.z & 0xffff, h = (rec.z >> 16) & 127;
      const int lr = lane / w, lc = lane - lr * w;
      if (lr >= rows) continue;
      const double2 v = vis[rec.w];
      const double sgn = (rec.x >> 30) & 1 ? -1.0 : 1.0;
      long long ti = (long long)rec.y + (long long)lr * stride + lc;
      const long long tstep = (long long)rows * stride;
      double2 *cell = acc + (r0 + lr) * LS + c0 + lc;
      const int cstep = rows * LS;
      for (int row = lr; row < h; row += rows, ti += tstep, cell += cstep) {
        if (ti < 0 || ti >= n_taps) continue;
        double2 t = taps[ti];
        t.y *= sgn;
        const double pr = v.x * t.x - v.y * t.y;
        const double pi = v.x * t.y + v.y * t.x;
        __hip_atomic_fetch_add(&cell->x, pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&cell->y, pi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    const int gsz = d.grid_size;
    double2 *gi = grid + ir.goff;
    for (int q = threadIdx.x; q < T * T; q += 64 * F64_WAVES) {
      const int r = q / T, cc = q - (q / T) * T;
      const int gu = ir.tu0 + r, gv = ir.tv0 + cc;
      if (gu >= gsz || gv >= gsz) continue;
      const double2 a = acc[r * LS + cc];
      double2 *g = gi + (long long)gu * d.pitch + gv;
      if (ir.single) {
        if (add) {
          double2 o = *g;
          o.x += a.x;
          o.y += a.y;
          *g = o;
        } else {
          *g = a;
        }
      } else if (a.x != 0.0 || a.y != 0.0) {
        unsafeAtomicAdd(&g->x, a.x);
        unsafeAtomicAdd(&g->y, a.y);
      }
    }
  });
}

// FP64 degridding over the same pairs (the degridder of the MS4 host-memory
// entry points, reference scatter_gridder_w_dependent_dyn_1p.cpp:368-431):
// the workgroup stages the tile of the grid in LDS; each wavefront takes
// pairs in turn, its lanes covering the window/tile intersection as in
// k_grid_tiles_f64 and summing the reference's complex products
// grid * tap (conjugated taps for mirrored layers) in FP64 without
// contraction; the 64 lane sums are added across the wavefront.  A pair
// holding a point's whole window stores its visibility; the pairs of a
// window split over tiles add theirs with FP64 atomics into the zeroed
// output (only the order of the additions differs from the reference).
template <int T>
__global__ void __launch_bounds__(64 * F64_WAVES)
k_degrid_tiles_f64(TiledDev d, const double2 *__restrict__ grid,
                   const double2 *__restrict__ taps, long long n_taps,
                   double2 *__restrict__ vis)
{
#pragma clang fp contract(off)
  constexpr int LS = T + 1;
  __shared__ double2 tile[T * LS];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for_each_item(d, d.queue + 8, [&](const ItemRange &ir) {
    const int gsz = d.grid_size;
    const double2 *gi = grid + ir.goff;
    for (int q = threadIdx.x; q < T * T; q += 64 * F64_WAVES) {
      const int r = q / T, c = q - (q / T) * T;
      const int gu = ir.tu0 + r, gv = ir.tv0 + c;
      tile[r * LS + c] = (gu < gsz && gv < gsz) ? gi[(long long)gu * d.pitch + gv]
                                                : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int k = ir.begin + wave; k < ir.end; k += F64_WAVES) {
      const int4 rec = d.pairs[k];
      const int r0 = rec.x & 63, c0 = (rec.x >> 6) & 63;
      const int w = ((rec.x >> 12) & 63) + 1;
      const int rows = ((rec.x >> 18) & 63) + 1;
      const int stride = rec.z & 0xffff, h = (rec.z >> 16) & 127;
      const int lr = lane / w, lc = lane - lr * w;
      double sx = 0.0, sy = 0.0;
      if (lr < rows) {
        const double sgn = (rec.x >> 30) & 1 ? -1.0 : 1.0;
        long long ti = (long long)rec.y + (long long)lr * stride + lc;
        const long long tstep = (long long)rows * stride;
        const double2 *cell = tile + (r0 + lr) * LS + c0 + lc;
        const int cstep = rows * LS;
        for (int row = lr; row < h; row += rows, ti += tstep, cell += cstep) {
          if (ti < 0 || ti >= n_taps) continue;
          double2 t = taps[ti];
          t.y *= sgn;
          const double2 g = *cell;
          sx += g.x * t.x - g.y * t.y;
          sy += g.x * t.y + g.y * t.x;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
      }
      if (lane == 0) {
        double2 *o = vis + rec.w;
        if (rec.x < 0) {  // the whole window lies in this tile
          *o = make_double2(sx, sy);
        } else {
          unsafeAtomicAdd(&o->x, sx);
          unsafeAtomicAdd(&o->y, sy);
        }
      }
    }
  });
}

// grid += lo, cell by cell (the end of a compensated gridding)
__global__ void k_comp_combine(float2 *__restrict__ grid, const float2 *__restrict__ lo,
                               long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float2 l = lo[i];
    if (l.x != 0.f || l.y != 0.f) {
      float2 g = grid[i];
      g.x += l.x;
      g.y += l.y;
      grid[i] = g;
    }
  }
}

// ---------------------------------------------------------------------------
// degridding: the grid tile is staged once in LDS (read-only); per pair the
// lanes accumulate four partial products, which are summed across the
// wavefront four pairs at a time.

template <int T, int DW>
__global__ void __launch_bounds__(64 * DW)
k_degrid_tiles(TiledDev d, const float2 *__restrict__ grid,
               const float2 *__restrict__ taps, long long n_taps,
               float2 *__restrict__ vis)
{
  constexpr int LS = T + 1;
  constexpr int DUMMY = T * LS;
  __shared__ float2 tile[T * LS + 2];
  __shared__ float2 sums[DW][SUMS * 64];
  // output codes of the wavefront's pairs k mod 128: two batches of 64, the
  // current one and the one whose last sums may still be pending
  __shared__ int codes[DW][128];
  const __amdgpu_buffer_rsrc_t rsrc = tap_rsrc(taps, n_taps);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int npts = (int)d.npts;
  float2 *my_sums = sums[wave];
  int *my_codes = codes[wave];
  for_each_item(d, d.queue + 8, [&](const ItemRange &ir) {
    const int gsz = d.grid_size;
    const float2 *gi = grid + ir.goff;
    for (int q = threadIdx.x; q < T * T; q += 64 * DW) {
      const int r = q / T, c = q - (q / T) * T;
      const int gu = ir.tu0 + r, gv = ir.tv0 + c;
      tile[r * LS + c] = (gu < gsz && gv < gsz) ? gi[(long long)gu * d.pitch + gv]
                                                : make_float2(0.f, 0.f);
    }
    if (threadIdx.x == 0) tile[DUMMY] = make_float2(0.f, 0.f);
    __syncthreads();
    const int npw = ir.end - ir.begin > wave ? (ir.end - ir.begin - wave + DW - 1) / DW : 0;
    if (npw > 0) {
      // records of the current batch and the next two
      int4 rec = batch_records(d, ir, wave, lane, npw, 0, DW);
      int4 rec_n = batch_records(d, ir, wave, lane, npw, 1, DW);
      int4 rec_nn = batch_records(d, ir, wave, lane, npw, 2, DW);
      int k = 0, s0 = 0;
      my_codes[lane] = out_code(rec, npts);
      Pair p = decode_pair<LS>(rl(rec.x, 0), rl(rec.y, 0), rl(rec.z, 0), rl(rec.w, 0), lane, 0);
      // +-1 from the conjugation bit by integer operations on the scalar pair
      // field (a select of two float constants becomes a vector instruction)
      auto sign_of = [](int conj) { return __int_as_float(0x3f800000 | (conj << 31)); };
      float sg = sign_of(p.conj);
      float2 ta[GRP], tb[GRP];
      int la[GRP], lb[GRP];
      load_group(ta, la, p, 0, DUMMY * 8, rsrc);
      // P = sum gx*tx, Q = sum gy*ty, R = sum gx*ty, U = sum gy*tx
      float P = 0.f, Q = 0.f, R = 0.f, U = 0.f;
      bool more = true;
      // one group; alternate buffers as in the gridder
      auto body = [&](float2 (&tc)[GRP], int (&lc)[GRP], float2 (&tn)[GRP], int (&ln)[GRP]) {
        // the current group's grid cells, read first so that the LDS
        // latency overlaps the choice of the next group
        float2 g[GRP];
#pragma unroll
        for (int u = 0; u < GRP; u++) g[u] = lds_at(tile, lc[u]);
        Pair pn = p;
        float sgn = sg;
        int kn = k, sn = s0 + GRP;
        const bool last = sn >= p.nslabs;
        if (last) {
          kn = k + 1;
          sn = 0;
          if (kn < npw) {
            const int j = kn & 63;
            if (j == 0) {
              // the slots of this batch last held the batch before the
              // previous one, whose sums are all stored
              // the moves first (vmove), then the load into the registers
              // they free
              rec = make_int4(vmove(rec_n.x), vmove(rec_n.y), vmove(rec_n.z), vmove(rec_n.w));
              rec_n = make_int4(vmove(rec_nn.x), vmove(rec_nn.y), vmove(rec_nn.z),
                                vmove(rec_nn.w));
              my_codes[(kn & 64) + lane] = out_code(rec, npts);
              __builtin_amdgcn_sched_barrier(0);
              rec_nn = batch_records(d, ir, wave, lane, npw, (kn >> 6) + 2, DW);
            }
            pn = decode_pair<LS>(rl(rec.x, j), rl(rec.y, j), rl(rec.z, j), rl(rec.w, j), lane,
                                 0);
            sgn = sign_of(pn.conj);
          } else {
            // past the end: the empty pair, no lane active; it never ends,
            // so no sum is written for it
            pn.h = 0;
            pn.nslabs = 1 << 20;
          }
        }
        more = kn < npw;
        load_group(tn, ln, pn, sn, DUMMY * 8, rsrc);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < GRP; u++) {
          P = fmaf(g[u].x, tc[u].x, P);
          Q = fmaf(g[u].y, tc[u].y, Q);
          R = fmaf(g[u].x, tc[u].y, R);
          U = fmaf(g[u].y, tc[u].x, U);
        }
        if (last) {
          // tap t' = (tx, sg*ty): re = P - sg*Q, im = sg*R + U (sg = +-1, so
          // each fma rounds as the add or subtract would); a point inside
          // one tile is final, a split point leaves its part in its slot for
          // k_degrid_reduce
          my_sums[(k & (SUMS - 1)) * 64 + lane] = make_float2(fmaf(-sg, Q, P), fmaf(sg, R, U));
          if ((k & (SUMS - 1)) == SUMS - 1)
            flush_sums(my_sums, my_codes + (k & 124), vis, d.partial, SUMS, lane);
          P = Q = R = U = 0.f;
        }
        p = pn;
        sg = sgn;
        k = kn;
        s0 = sn;
      };
      // Both bodies run on every trip (the second one past the end works on
      // the empty pair and touches only the dummy cell): a loop exit between
      // them would join the back edge with undefined state, which costs
      // vector copies on every group.
      do {
        body(ta, la, tb, lb);
        body(tb, lb, ta, la);
      } while (more);
      if (npw & (SUMS - 1))
        flush_sums(my_sums, my_codes + ((npw - 1) & 124), vis, d.partial, npw & (SUMS - 1),
                   lane);
      // never true (chunk >= 1): keeps the prefetch loads from sinking
      if (d.chunk < 0) {
        float2 z = make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < GRP; u++) {
          z.x += ta[u].x + tb[u].x;
          z.y += ta[u].y + tb[u].y;
        }
        vis[lane] = z;
      }
    }
  });
}

// Split points: the sum of their per-tile parts in slot order (fixed order,
// so the FP32 result is reproducible).  Out-of-bound points keep the zero
// the entry point wrote.
__global__ void __launch_bounds__(256)
k_degrid_reduce(TiledDev d, const rcg_pg *__restrict__ pg, long long n,
                float2 *__restrict__ vis)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const rcg_pg r = pg[i];
  if (r.cell < 0 || r.supp <= 0) return;
  int np;
  if (d.layout == LAYOUT_HALO) {
    const int ns = (r.supp + d.wmax - 1) >> d.lw;
    np = ns * ns;
  } else {
    int u0, v0, tu0, tu1, tv0, tv1;
    point_tiles(r, d.pitch, d.inv_pitch, d.lt, &u0, &v0, &tu0, &tu1, &tv0, &tv1);
    np = (tu1 - tu0 + 1) * (tv1 - tv0 + 1);
  }
  if (np <= 1 || np > d.mpp) return;  // whole (written by the tile kernel)
  // partials are stored slot-major, so consecutive points read consecutive
  // elements of each slot
  const float2 *q = d.partial + i;
  float2 s = q[0];
  for (int k = 1; k < np; k++) {
    const float2 t = q[(long long)k * d.npts];
    s.x += t.x;
    s.y += t.y;
  }
  vis[i] = s;
}

// ---------------------------------------------------------------------------
// host side

namespace {

int ilog2(int x)
{
  int l = 0;
  while ((1 << l) < x) l++;
  return l;
}

template <class T>
hipError_t dalloc(T **p, size_t n)
{
  return hipMalloc(reinterpret_cast<void **>(p), (n ? n : 1) * sizeof(T));
}

// Persistent workgroups of a kernel: resident workgroups per CU x CUs,
// a multiple of 8 (one share per XCD).
hipError_t persistent_blocks(const void *kernel, int block, int *nb)
{
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  if (e != hipSuccess) return e;
  *nb = std::max(8, (std::max(1, per_cu) * std::max(1, cus)) & ~7);
  return hipSuccess;
}

// The gridding kernel of tile T with gw wavefronts per workgroup.
template <int T>
const void *grid_kernel(int gw)
{
  switch (gw) {
    case 1: return reinterpret_cast<const void *>(&k_grid_tiles<T, 1>);
    case 2: return reinterpret_cast<const void *>(&k_grid_tiles<T, 2>);
    case 3: return reinterpret_cast<const void *>(&k_grid_tiles<T, 3>);
    default: return reinterpret_cast<const void *>(&k_grid_tiles<T, 4>);
  }
}

// The degridding kernel of tile T with dw wavefronts per workgroup.
template <int T>
const void *degrid_kernel(int dw)
{
  switch (dw) {
    case 2: return reinterpret_cast<const void *>(&k_degrid_tiles<T, 2>);
    case 8: return reinterpret_cast<const void *>(&k_degrid_tiles<T, 8>);
    default: return reinterpret_cast<const void *>(&k_degrid_tiles<T, 4>);
  }
}

// Persistent workgroups of the grid kernel (gw wavefronts) and, when
// degrid_blocks is given, of the degrid kernel (dw wavefronts).
template <int T>
hipError_t tile_kernels(int gw, int dw, int *grid_blocks, int *degrid_blocks)
{
  hipError_t e = grid_blocks ? persistent_blocks(grid_kernel<T>(gw), 64 * gw, grid_blocks)
                             : hipSuccess;
  if (e == hipSuccess && degrid_blocks)
    e = persistent_blocks(degrid_kernel<T>(dw), 64 * dw, degrid_blocks);
  return e;
}

template <int T>
void launch_degrid_tiles(int dw, int nb, const TiledDev &d, const float2 *g, const float2 *t,
                         long long nt, float2 *v, hipStream_t st)
{
  switch (dw) {
    case 2:
      hipLaunchKernelGGL((k_degrid_tiles<T, 2>), dim3(nb), dim3(128), 0, st, d, g, t, nt, v);
      break;
    case 8:
      hipLaunchKernelGGL((k_degrid_tiles<T, 8>), dim3(nb), dim3(512), 0, st, d, g, t, nt, v);
      break;
    default:
      hipLaunchKernelGGL((k_degrid_tiles<T, 4>), dim3(nb), dim3(256), 0, st, d, g, t, nt, v);
      break;
  }
}

template <int T, int GW>
void launch_grid_tiles(int nb, const TiledDev &d, const float2 *v, const float2 *t,
                       long long nt, float2 *g, hipStream_t st)
{
  hipLaunchKernelGGL((k_grid_tiles<T, GW>), dim3(nb), dim3(64 * GW), 0, st, d, v, t, nt, g);
}

template <int T>
void launch_grid_tiles(int gw, int nb, const TiledDev &d, const float2 *v, const float2 *t,
                       long long nt, float2 *g, hipStream_t st)
{
  switch (gw) {
    case 1: launch_grid_tiles<T, 1>(nb, d, v, t, nt, g, st); break;
    case 2: launch_grid_tiles<T, 2>(nb, d, v, t, nt, g, st); break;
    case 3: launch_grid_tiles<T, 3>(nb, d, v, t, nt, g, st); break;
    default: launch_grid_tiles<T, 4>(nb, d, v, t, nt, g, st); break;
  }
}

rcg_status check_tap_table(const rcg_gcf *gcf, const char *who)
{
  if (gcf->n_taps <= 0 || gcf->n_taps >= (0x7fffffffLL >> 3)) {
    set_error(RCG_ERR_ARG, "%s: tap table of %lld elements; the FP32 path needs "
              "0 < n_taps < 2^28", who, (long long)gcf->n_taps);
    return RCG_ERR_ARG;
  }
  return RCG_OK;
}

// Events around a plan's main kernel when timing is on.
hipEvent_t *timing_slot(rcg_tiled_plan *pl, int dir)
{
  if (!pl->timing) return nullptr;
  return pl->ev[dir][pl->n_ev[dir] % N_TIMED];
}

}  // namespace

}  // namespace rcg

using namespace rcg;

extern "C" rcg_status rcg_tiled_plan_create_ex(const rcg_geom *geoms, int n_inst, int tile,
                                               int chunk, int max_support, int layout,
                                               rcg_tiled_plan **out)
{
  clear_error();
  if (!geoms || !out || n_inst < 1) {
    set_error(RCG_ERR_ARG, "rcg_tiled_plan_create: null argument or no instance");
    return RCG_ERR_ARG;
  }
  *out = nullptr;
  if (layout != RCG_LAYOUT_PAIRS && layout != RCG_LAYOUT_HALO && layout != RCG_LAYOUT_HALO16) {
    set_error(RCG_ERR_ARG, "layout %d is not one of RCG_LAYOUT_PAIRS, RCG_LAYOUT_HALO, "
              "RCG_LAYOUT_HALO16", layout);
    return RCG_ERR_ARG;
  }
  long long n = 0;
  std::vector<int> pt_off(n_inst + 1), blk_off(n_inst + 1);
  for (int k = 0; k < n_inst; k++) {
    rcg_status s = check_geom(&geoms[k]);
    if (s != RCG_OK) return s;
    if (geoms[k].grid_size != geoms[0].grid_size || geoms[k].grid_pitch != geoms[0].grid_pitch) {
      set_error(RCG_ERR_ARG, "instance %d: every instance needs the grid shape of the first", k);
      return RCG_ERR_ARG;
    }
    pt_off[k] = (int)n;
    n += (long long)geoms[k].n_chunks * geoms[k].ts_ch;
    if (n > INT_MAX) {
      set_error(RCG_ERR_ARG, "more than 2^31-1 points in one plan");
      return RCG_ERR_ARG;
    }
  }
  pt_off[n_inst] = (int)n;
  const bool halo = layout != RCG_LAYOUT_PAIRS;
  if (tile != 16 && tile != 32 && (halo || tile != 64)) {
    set_error(RCG_ERR_ARG, "tile must be 16, 32%s (got %d)", halo ? "" : " or 64", tile);
    return RCG_ERR_ARG;
  }
  if (chunk < 1 || chunk > (halo ? MAX_HALO_ITEM : (1 << 20)) || max_support < 1 ||
      max_support > MAX_SUPPORT) {
    set_error(RCG_ERR_ARG, "chunk %d / max_support %d out of range%s", chunk, max_support,
              halo ? " (the halo layout takes chunk <= 1024)" : "");
    return RCG_ERR_ARG;
  }
  const int size = geoms[0].grid_size, pitch = geoms[0].grid_pitch;
  const int ntx = (size + tile - 1) / tile;
  const long long tpi = (long long)ntx * ntx;
  if (tpi * n_inst > MAX_PLAN_TILES) {
    set_error(RCG_ERR_ARG, "%lld tiles in the plan exceed %d: use a larger tile or split "
              "the batch", tpi * n_inst, MAX_PLAN_TILES);
    return RCG_ERR_ARG;
  }
  // records per point: tiles a window overlaps (pairs) or sub-windows (halo)
  const int wmax = max_support <= 16 || layout == RCG_LAYOUT_HALO16 ? 16 : 32;
  int max_pairs_pt;
  if (halo) {
    const int nsub = (max_support + wmax - 1) / wmax;
    max_pairs_pt = nsub * nsub;
  } else {
    const int per_dim = (max_support - 1 + tile - 1) / tile + 1;
    max_pairs_pt = per_dim * per_dim;
  }
  if (max_pairs_pt > 64) {  // the record holds the slot in 6 bits
    set_error(RCG_ERR_UNSUPPORTED,
              "max_support %d needs %d records per point with tile %d; at most 64 are supported",
              max_support, max_pairs_pt, tile);
    return RCG_ERR_UNSUPPORTED;
  }
  if (n * max_pairs_pt > INT_MAX) {
    set_error(RCG_ERR_ARG, "%lld points x %d records per point exceed the record index range: "
              "split the batch", n, max_pairs_pt);
    return RCG_ERR_ARG;
  }
  // Binning: LDS histograms per workgroup when the instance's tiles fit in
  // LDS (MAX_BIN_TILES), global atomics per record otherwise; 4 points per
  // thread.
  const int ppt = 4;
  const bool gbins = tpi > MAX_BIN_TILES;
  int nblk = 0;
  for (int k = 0; k < n_inst; k++) {
    blk_off[k] = nblk;
    nblk += ceil_div(pt_off[k + 1] - pt_off[k], (long long)BIN_THREADS * ppt);
  }
  blk_off[n_inst] = nblk;

  rcg_tiled_plan *pl = new (std::nothrow) rcg_tiled_plan();
  if (!pl) {
    set_error(RCG_ERR_NOMEM, "host allocation failed");
    return RCG_ERR_NOMEM;
  }
  pl->geoms.assign(geoms, geoms + n_inst);
  pl->n = n;
  pl->max_support = max_support;
  pl->max_pairs_pt = max_pairs_pt;
  pl->bin_blocks = nblk;
  TiledDev &d = pl->dev;
  d.layout = halo ? LAYOUT_HALO : LAYOUT_PAIRS;
  d.wmax = wmax;
  d.lw = ilog2(wmax);
  d.gbins = gbins ? 1 : 0;
  d.T = tile;
  d.lt = ilog2(tile);
  d.ntx = ntx;
  d.tpi = (int)tpi;
  d.n_inst = n_inst;
  d.n_tiles = (int)(tpi * n_inst);
  d.chunk = chunk;
  d.pitch = pitch;
  d.grid_size = size;
  d.mpp = max_pairs_pt;
  d.ppt = ppt;
  d.gstride = (long long)size * pitch;
  d.pair_cap = n * max_pairs_pt;
  d.npts = n;
  d.inv_pitch = 1.0 / pitch;
  const int n_seg = ceil_div(d.n_tiles, SCAN_SEG);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bin_fill),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bin_count),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e == hipSuccess && !halo) {
    switch (tile) {
      case 16:
        pl->grid_waves = grid_waves<16>();
        e = tile_kernels<16>(pl->grid_waves, pl->degrid_waves, &pl->grid_blocks,
                             &pl->degrid_blocks);
        break;
      case 32:
        pl->grid_waves = grid_waves<32>();
        e = tile_kernels<32>(pl->grid_waves, pl->degrid_waves, &pl->grid_blocks,
                             &pl->degrid_blocks);
        break;
      default:
        pl->grid_waves = grid_waves<64>();
        e = tile_kernels<64>(pl->grid_waves, pl->degrid_waves, &pl->grid_blocks,
                             &pl->degrid_blocks);
        break;
    }
  }
  if (e == hipSuccess && halo &&
      halo_blocks(tile, wmax, &pl->grid_blocks, &pl->degrid_blocks) != RCG_OK)
    e = hipErrorInvalidValue;
  if (e == hipSuccess) e = dalloc(&pl->geoms_dev, (size_t)n_inst);
  if (e == hipSuccess) e = dalloc(&pl->offs, (size_t)2 * (n_inst + 1));
  if (e == hipSuccess) e = dalloc(&pl->pg, (size_t)n);
  if (e == hipSuccess) e = dalloc(&d.tile_count, (size_t)d.n_tiles);
  if (e == hipSuccess) e = dalloc(&d.tile_start, (size_t)d.n_tiles);
  if (e == hipSuccess) e = dalloc(&d.tile_cursor, (size_t)d.n_tiles);
  if (e == hipSuccess) e = dalloc(&d.item_start, (size_t)d.n_tiles);
  if (e == hipSuccess) e = dalloc(&d.totals, 4);
  if (e == hipSuccess) e = dalloc(&d.seg, (size_t)2 * n_seg);
  if (e == hipSuccess) e = dalloc(&d.queue, (size_t)N_QUEUE);
  if (e == hipSuccess) e = dalloc(&pl->taps, 1);
  if (e == hipSuccess) e = dalloc(&pl->err, 1);
  if (e == hipSuccess) e = dalloc(&d.pairs, (size_t)d.pair_cap);
  if (e == hipSuccess) e = dalloc(&d.partial, (size_t)d.pair_cap);
  if (e == hipSuccess && halo) {
    // item buffers of the halo gridder: every work item up to the bound on
    // their number (records / chunk + tiles), within a 2 GiB budget; items
    // past it flush with atomics
    const long long cells = halo_region_cells(tile, wmax);
    const long long bound = d.pair_cap / chunk + d.n_tiles;
    const long long budget = (2LL << 30) / (cells * (long long)sizeof(float2));
    d.item_cap = (int)std::min(bound, (long long)INT_MAX);
    d.ibuf_cap = (int)std::min(bound, budget);
    e = dalloc(&d.ibuf, (size_t)(d.ibuf_cap * cells));
    if (e == hipSuccess) e = dalloc(&d.runs, (size_t)d.item_cap * 64);
    if (e == hipSuccess) e = dalloc(&d.vsort, (size_t)d.pair_cap);
  }
  if (e == hipSuccess)
    e = hipMemcpy(pl->geoms_dev, geoms, sizeof(rcg_geom) * n_inst, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(pl->offs, pt_off.data(), sizeof(int) * (n_inst + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(pl->offs + n_inst + 1, blk_off.data(), sizeof(int) * (n_inst + 1),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    rcg_tiled_plan_destroy(pl);
    set_error(RCG_ERR_NOMEM, "plan set-up failed: %s", hipGetErrorString(e));
    return RCG_ERR_NOMEM;
  }
  d.pt_off = pl->offs;
  d.blk_off = pl->offs + n_inst + 1;
  *out = pl;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_plan_create_batch(const rcg_geom *geoms, int n_inst, int tile,
                                                  int chunk, int max_support,
                                                  rcg_tiled_plan **out)
{
  return rcg_tiled_plan_create_ex(geoms, n_inst, tile, chunk, max_support, RCG_LAYOUT_PAIRS,
                                  out);
}

extern "C" rcg_status rcg_tiled_plan_create(const rcg_geom *geom, int tile, int chunk,
                                            int max_support, rcg_tiled_plan **out)
{
  return rcg_tiled_plan_create_batch(geom, 1, tile, chunk, max_support, out);
}

extern "C" rcg_status rcg_tiled_plan_destroy(rcg_tiled_plan *pl)
{
  if (!pl) return RCG_OK;
  TiledDev &d = pl->dev;
  void *ptrs[] = {pl->geoms_dev, pl->offs, pl->pg, d.tile_count, d.tile_start,
                  d.tile_cursor, d.item_start, d.totals, d.seg, d.queue, pl->taps, pl->err,
                  d.pairs, d.partial, pl->lo, d.ibuf, d.runs, d.vsort};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int dir = 0; dir < 2; dir++)
    for (int k = 0; k < N_TIMED; k++)
      for (int j = 0; j < 2; j++)
        if (pl->ev[dir][k][j]) (void)hipEventDestroy(pl->ev[dir][k][j]);
  delete pl;
  return RCG_OK;
}

namespace {
rcg_status bin_passes(rcg_tiled_plan *pl, hipStream_t st);
}  // namespace

extern "C" rcg_status rcg_tiled_bin(rcg_tiled_plan *pl, const rcg_gcf *gcf, void *stream)
{
  clear_error();
  if (!pl || !gcf) {
    set_error(RCG_ERR_ARG, "rcg_tiled_bin: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_gcf(gcf);
  if (s != RCG_OK) return s;
  hipStream_t st = (hipStream_t)stream;
  TiledDev &d = pl->dev;
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, st, d.tile_count, 0,
                     (long long)d.n_tiles);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.totals, 0, 4LL);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st,
                     reinterpret_cast<int *>(pl->taps), 0, 2LL);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, pl->err, 0, 1LL);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  s = launch_pregrid_batch(pl->geoms_dev, d.pt_off, d.n_inst, pl->n, gcf, pl->pg, pl->taps,
                           pl->err, st);
  if (s != RCG_OK) return s;
  return bin_passes(pl, st);
}

rcg_status rcg::tiled_bin_records(rcg_tiled_plan *pl, hipStream_t st)
{
  TiledDev &d = pl->dev;
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, st, d.tile_count, 0,
                     (long long)d.n_tiles);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.totals, 0, 4LL);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st,
                     reinterpret_cast<int *>(pl->taps), 0, 2LL);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, pl->err, 0, 1LL);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  return bin_passes(pl, st);
}

namespace {

// Counting sort of the plan's records into per-tile pair lists: LDS
// histograms, three-kernel scan, LDS-ranked fill.
rcg_status bin_passes(rcg_tiled_plan *pl, hipStream_t st)
{
  TiledDev &d = pl->dev;
  const size_t lds_count = d.gbins ? 0 : (size_t)d.tpi * sizeof(int);
  const int n_seg = ceil_div(d.n_tiles, SCAN_SEG);
  hipLaunchKernelGGL(k_bin_count, dim3(pl->bin_blocks), dim3(BIN_THREADS), lds_count, st,
                     pl->pg, d, pl->err);
  hipLaunchKernelGGL(k_scan_reduce, dim3(n_seg), dim3(SCAN_SEG), 0, st, d);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_SEG), 0, st, d, n_seg);
  hipLaunchKernelGGL(k_scan_apply, dim3(n_seg), dim3(SCAN_SEG), 0, st, d);
  hipLaunchKernelGGL(k_bin_fill, dim3(pl->bin_blocks), dim3(BIN_THREADS), lds_count, st,
                     pl->pg, d, pl->err);
  RCG_HIP_TRY(hipGetLastError());
  if (d.layout == LAYOUT_HALO) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.queue + 16, 0, 1LL);
    RCG_HIP_TRY(hipGetLastError());
    return halo_sort_items(d, st);
  }
  return RCG_OK;
}

}  // namespace

namespace {

rcg_status grid_tiled(rcg_tiled_plan *pl, const rcg_gcf *gcf, const rcg_c64 *vis, rcg_c64 *grid,
                      int adjoint, void *stream)
{
  clear_error();
  if (!pl || !gcf || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_grid_tiled_f32: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_tap_table(gcf, "rcg_grid_tiled_f32");
  if (s != RCG_OK) return s;
  hipStream_t st = (hipStream_t)stream;
  TiledDev d = pl->dev;
  d.adj = adjoint;
  d.lo = pl->compensated ? pl->lo : nullptr;
  hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st,
                     reinterpret_cast<float *>(grid), d.gstride * d.n_inst * 2);
  if (d.lo)
    hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st,
                       reinterpret_cast<float *>(d.lo), d.gstride * d.n_inst * 2);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.queue, 0, 8LL);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  const float2 *v = reinterpret_cast<const float2 *>(vis);
  const float2 *t = reinterpret_cast<const float2 *>(gcf->data);
  const long long nt = gcf->n_taps;
  float2 *g = reinterpret_cast<float2 *>(grid);
  const dim3 nb(pl->grid_blocks);
  hipEvent_t *ev = timing_slot(pl, 0);
  if (ev) RCG_HIP_TRY(hipEventRecord(ev[0], st));
  if (d.layout == LAYOUT_HALO) {
    s = halo_grid(d, pl->grid_blocks, v, t, nt, g, st);
    if (s != RCG_OK) return s;
  } else {
    switch (d.T) {
      case 16: launch_grid_tiles<16>(pl->grid_waves, pl->grid_blocks, d, v, t, nt, g, st); break;
      case 32: launch_grid_tiles<32>(pl->grid_waves, pl->grid_blocks, d, v, t, nt, g, st); break;
      default: launch_grid_tiles<64>(pl->grid_waves, pl->grid_blocks, d, v, t, nt, g, st); break;
    }
  }
  RCG_HIP_TRY(hipGetLastError());
  if (ev) {
    RCG_HIP_TRY(hipEventRecord(ev[1], st));
    pl->n_ev[0]++;
  }
  if (d.lo) {
    hipLaunchKernelGGL(k_comp_combine, dim3(2048), dim3(256), 0, st, g, d.lo,
                       d.gstride * d.n_inst);
    RCG_HIP_TRY(hipGetLastError());
  }
  return RCG_OK;
}

}  // namespace

extern "C" rcg_status rcg_tiled_compensated(rcg_tiled_plan *pl, int enable)
{
  clear_error();
  if (!pl) {
    set_error(RCG_ERR_ARG, "rcg_tiled_compensated: null plan");
    return RCG_ERR_ARG;
  }
  if (enable && !pl->lo) {
    const size_t cells = (size_t)pl->dev.gstride * pl->dev.n_inst;
    if (hipMalloc(reinterpret_cast<void **>(&pl->lo), cells * sizeof(float2)) != hipSuccess) {
      pl->lo = nullptr;
      set_error(RCG_ERR_NOMEM, "rcg_tiled_compensated: %zu cells", cells);
      return RCG_ERR_NOMEM;
    }
  }
  pl->compensated = enable != 0;
  return RCG_OK;
}

extern "C" rcg_status rcg_grid_tiled_f32(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                         const rcg_c64 *vis, rcg_c64 *grid, void *stream)
{
  return grid_tiled(pl, gcf, vis, grid, 0, stream);
}

extern "C" rcg_status rcg_grid_tiled_adjoint_f32(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                                 const rcg_c64 *vis, rcg_c64 *grid, void *stream)
{
  return grid_tiled(pl, gcf, vis, grid, 1, stream);
}

extern "C" rcg_status rcg_degrid_tiled_f32(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                           const rcg_c64 *grid, rcg_c64 *vis,
                                           void *stream)
{
  clear_error();
  if (!pl || !gcf || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_degrid_tiled_f32: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_tap_table(gcf, "rcg_degrid_tiled_f32");
  if (s != RCG_OK) return s;
  hipStream_t st = (hipStream_t)stream;
  const TiledDev &d = pl->dev;
  hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(256), 0, st,
                     reinterpret_cast<float *>(vis), pl->n * 2);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.queue + 8, 0, 8LL);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  const float2 *g = reinterpret_cast<const float2 *>(grid);
  const float2 *t = reinterpret_cast<const float2 *>(gcf->data);
  const long long nt = gcf->n_taps;
  float2 *v = reinterpret_cast<float2 *>(vis);
  const dim3 nb(pl->degrid_blocks);
  hipEvent_t *ev = timing_slot(pl, 1);
  if (ev) RCG_HIP_TRY(hipEventRecord(ev[0], st));
  if (d.layout == LAYOUT_HALO) {
    s = halo_degrid(d, pl->degrid_blocks, g, t, nt, v, st);
    if (s != RCG_OK) return s;
  } else {
    switch (d.T) {
      case 16: launch_degrid_tiles<16>(pl->degrid_waves, pl->degrid_blocks, d, g, t, nt, v, st); break;
      case 32: launch_degrid_tiles<32>(pl->degrid_waves, pl->degrid_blocks, d, g, t, nt, v, st); break;
      default: launch_degrid_tiles<64>(pl->degrid_waves, pl->degrid_blocks, d, g, t, nt, v, st); break;
    }
  }
  RCG_HIP_TRY(hipGetLastError());
  if (ev) {
    RCG_HIP_TRY(hipEventRecord(ev[1], st));
    pl->n_ev[1]++;
  }
  hipLaunchKernelGGL(k_degrid_reduce, dim3((unsigned)((pl->n + 255) / 256)), dim3(256), 0, st,
                     d, pl->pg, pl->n, v);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

namespace {

template <int T>
rcg_status launch_grid_f64(const TiledDev &d, const double2 *v, const double2 *t, long long nt,
                           double2 *g, int add, hipStream_t st)
{
  int nb = 0;
  RCG_HIP_TRY(persistent_blocks(reinterpret_cast<const void *>(&k_grid_tiles_f64<T>),
                                64 * F64_WAVES, &nb));
  hipLaunchKernelGGL(k_grid_tiles_f64<T>, dim3(nb), dim3(64 * F64_WAVES), 0, st, d, v, t, nt, g,
                     add);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

rcg_status grid_f64(rcg_tiled_plan *pl, const double2 *v, const double2 *t, long long nt,
                    double2 *g, int add, hipStream_t st)
{
  TiledDev d = pl->dev;
  if (d.layout != LAYOUT_PAIRS) {
    set_error(RCG_ERR_UNSUPPORTED, "FP64 tiled gridding needs a plan of RCG_LAYOUT_PAIRS");
    return RCG_ERR_UNSUPPORTED;
  }
  d.adj = 0;
  d.lo = nullptr;
  if (!add)
    hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st, reinterpret_cast<float *>(g),
                       d.gstride * d.n_inst * 4);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.queue, 0, 8LL);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  hipEvent_t *ev = timing_slot(pl, 0);
  if (ev) RCG_HIP_TRY(hipEventRecord(ev[0], st));
  rcg_status s;
  switch (d.T) {
    case 16: s = launch_grid_f64<16>(d, v, t, nt, g, add, st); break;
    case 32: s = launch_grid_f64<32>(d, v, t, nt, g, add, st); break;
    default: s = launch_grid_f64<64>(d, v, t, nt, g, add, st); break;
  }
  if (s != RCG_OK) return s;
  if (ev) {
    RCG_HIP_TRY(hipEventRecord(ev[1], st));
    pl->n_ev[0]++;
  }
  return RCG_OK;
}

template <int T>
rcg_status launch_degrid_f64(const TiledDev &d, const double2 *g, const double2 *t,
                             long long nt, double2 *v, hipStream_t st)
{
  int nb = 0;
  RCG_HIP_TRY(persistent_blocks(reinterpret_cast<const void *>(&k_degrid_tiles_f64<T>),
                                64 * F64_WAVES, &nb));
  hipLaunchKernelGGL(k_degrid_tiles_f64<T>, dim3(nb), dim3(64 * F64_WAVES), 0, st, d, g, t, nt,
                     v);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

rcg_status degrid_f64(rcg_tiled_plan *pl, const double2 *g, const double2 *t, long long nt,
                      double2 *v, hipStream_t st)
{
  TiledDev d = pl->dev;
  if (d.layout != LAYOUT_PAIRS) {
    set_error(RCG_ERR_UNSUPPORTED, "FP64 tiled degridding needs a plan of RCG_LAYOUT_PAIRS");
    return RCG_ERR_UNSUPPORTED;
  }
  // zeroed output: points left out stay 0, split windows add into it
  hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st, reinterpret_cast<float *>(v),
                     d.npts * 4);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.queue + 8, 0, 8LL);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  hipEvent_t *ev = timing_slot(pl, 1);
  if (ev) RCG_HIP_TRY(hipEventRecord(ev[0], st));
  rcg_status s;
  switch (d.T) {
    case 16: s = launch_degrid_f64<16>(d, g, t, nt, v, st); break;
    case 32: s = launch_degrid_f64<32>(d, g, t, nt, v, st); break;
    default: s = launch_degrid_f64<64>(d, g, t, nt, v, st); break;
  }
  if (s != RCG_OK) return s;
  if (ev) {
    RCG_HIP_TRY(hipEventRecord(ev[1], st));
    pl->n_ev[1]++;
  }
  return RCG_OK;
}

}  // namespace

extern "C" rcg_status rcg_tiled_set_grid_waves(rcg_tiled_plan *pl, int waves)
{
  clear_error();
  if (!pl) {
    set_error(RCG_ERR_ARG, "rcg_tiled_set_grid_waves: null plan");
    return RCG_ERR_ARG;
  }
  if (pl->dev.layout != LAYOUT_PAIRS) {
    set_error(RCG_ERR_UNSUPPORTED, "rcg_tiled_set_grid_waves: the halo layout's gridder runs "
              "one wavefront per workgroup");
    return RCG_ERR_UNSUPPORTED;
  }
  if (waves < 0 || waves > MAX_GRID_WAVES) {
    set_error(RCG_ERR_ARG, "rcg_tiled_set_grid_waves: waves %d not in [0, %d]", waves,
              MAX_GRID_WAVES);
    return RCG_ERR_ARG;
  }
  const int T = pl->dev.T;
  if (waves == 0) waves = T == 16 ? grid_waves<16>() : T == 32 ? grid_waves<32>() : grid_waves<64>();
  int nb = 0;
  hipError_t e = T == 16 ? tile_kernels<16>(waves, 0, &nb, nullptr)
               : T == 32 ? tile_kernels<32>(waves, 0, &nb, nullptr)
                         : tile_kernels<64>(waves, 0, &nb, nullptr);
  RCG_HIP_TRY(e);
  pl->grid_waves = waves;
  pl->grid_blocks = nb;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_set_degrid_waves(rcg_tiled_plan *pl, int waves)
{
  clear_error();
  if (!pl) {
    set_error(RCG_ERR_ARG, "rcg_tiled_set_degrid_waves: null plan");
    return RCG_ERR_ARG;
  }
  if (pl->dev.layout != LAYOUT_PAIRS) {
    set_error(RCG_ERR_UNSUPPORTED, "rcg_tiled_set_degrid_waves: the halo layout's degridder "
              "runs one wavefront per workgroup");
    return RCG_ERR_UNSUPPORTED;
  }
  if (waves == 0) waves = 4;
  if (waves != 2 && waves != 4 && waves != 8) {
    set_error(RCG_ERR_ARG, "rcg_tiled_set_degrid_waves: waves %d not 2, 4 or 8 (0: 4)", waves);
    return RCG_ERR_ARG;
  }
  const int T = pl->dev.T;
  int nb = 0;
  hipError_t e = T == 16 ? tile_kernels<16>(0, waves, nullptr, &nb)
               : T == 32 ? tile_kernels<32>(0, waves, nullptr, &nb)
                         : tile_kernels<64>(0, waves, nullptr, &nb);
  RCG_HIP_TRY(e);
  pl->degrid_waves = waves;
  pl->degrid_blocks = nb;
  return RCG_OK;
}

extern "C" rcg_status rcg_degrid_tiled_f64(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                           const rcg_c128 *grid, rcg_c128 *vis, void *stream)
{
  clear_error();
  if (!pl || !gcf || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_degrid_tiled_f64: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_gcf(gcf);
  if (s != RCG_OK) return s;
  if (gcf->n_taps <= 0) {
    set_error(RCG_ERR_ARG, "rcg_degrid_tiled_f64: empty tap table");
    return RCG_ERR_ARG;
  }
  return degrid_f64(pl, reinterpret_cast<const double2 *>(grid),
                    reinterpret_cast<const double2 *>(gcf->data), gcf->n_taps,
                    reinterpret_cast<double2 *>(vis), (hipStream_t)stream);
}

extern "C" rcg_status rcg_grid_tiled_f64(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                         const rcg_c128 *vis, rcg_c128 *grid, void *stream)
{
  clear_error();
  if (!pl || !gcf || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_grid_tiled_f64: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_gcf(gcf);
  if (s != RCG_OK) return s;
  if (gcf->n_taps <= 0) {
    set_error(RCG_ERR_ARG, "rcg_grid_tiled_f64: empty tap table");
    return RCG_ERR_ARG;
  }
  return grid_f64(pl, reinterpret_cast<const double2 *>(vis),
                  reinterpret_cast<const double2 *>(gcf->data), gcf->n_taps,
                  reinterpret_cast<double2 *>(grid), 0, (hipStream_t)stream);
}

rcg_status rcg::tiled_grid_f64_add(rcg_tiled_plan *pl, const double2 *taps, long long n_taps,
                                   const double2 *vis, double2 *grid, hipStream_t st)
{
  if (n_taps <= 0) {
    set_error(RCG_ERR_ARG, "tiled FP64 gridding: empty tap table");
    return RCG_ERR_ARG;
  }
  return grid_f64(pl, vis, taps, n_taps, grid, 1, st);
}

extern "C" rcg_status rcg_tiled_stats(rcg_tiled_plan *pl, unsigned long long *taps,
                                      long long *n_pairs, long long *n_items,
                                      int *err, void *stream)
{
  clear_error();
  if (!pl) {
    set_error(RCG_ERR_ARG, "rcg_tiled_stats: null plan");
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  int tot[4] = {0, 0, 0, 0};
  unsigned long long tp = 0;
  int e = 0;
  RCG_HIP_TRY(hipMemcpyAsync(tot, pl->dev.totals, sizeof(tot), hipMemcpyDeviceToHost, st));
  RCG_HIP_TRY(hipMemcpyAsync(&tp, pl->taps, sizeof(tp), hipMemcpyDeviceToHost, st));
  RCG_HIP_TRY(hipMemcpyAsync(&e, pl->err, sizeof(e), hipMemcpyDeviceToHost, st));
  RCG_HIP_TRY(hipStreamSynchronize(st));
  if (taps) *taps = tp;
  if (n_pairs) *n_pairs = tot[0];
  if (n_items) *n_items = tot[1];
  if (err) *err = e;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_timing(rcg_tiled_plan *pl, int enable)
{
  clear_error();
  if (!pl) {
    set_error(RCG_ERR_ARG, "rcg_tiled_timing: null plan");
    return RCG_ERR_ARG;
  }
  if (enable && !pl->ev[0][0][0]) {
    for (int dir = 0; dir < 2; dir++)
      for (int k = 0; k < N_TIMED; k++)
        for (int j = 0; j < 2; j++) RCG_HIP_TRY(hipEventCreate(&pl->ev[dir][k][j]));
  }
  pl->timing = enable != 0;
  pl->n_ev[0] = pl->n_ev[1] = 0;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_kernel_times(rcg_tiled_plan *pl, int which, float *ms,
                                             int max, int *count)
{
  clear_error();
  if (!pl || !ms || !count || which < 0 || which > 1 || max < 0) {
    set_error(RCG_ERR_ARG, "rcg_tiled_kernel_times: bad argument");
    return RCG_ERR_ARG;
  }
  const int rec = pl->n_ev[which];
  const int n = std::min(std::min(rec, N_TIMED), max);
  for (int k = 0; k < n; k++) {
    hipEvent_t *e = pl->ev[which][(rec - n + k) % N_TIMED];
    RCG_HIP_TRY(hipEventSynchronize(e[1]));
    RCG_HIP_TRY(hipEventElapsedTime(&ms[k], e[0], e[1]));
  }
  *count = n;
  pl->n_ev[which] = 0;
  return RCG_OK;
}

// ---------------------------------------------------------------------------
// graphs

struct rcg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" rcg_status rcg_graph_destroy(rcg_graph *g)
{
  if (!g) return RCG_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_graph_create(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                             const rcg_c64 *vis, rcg_c64 *grid,
                                             rcg_c64 *vis_out, int stages, rcg_graph **out)
{
  clear_error();
  if (!pl || !gcf || !out || (stages & ~7) || !stages ||
      ((stages & RCG_STAGE_GRID) && (!vis || !grid)) ||
      ((stages & RCG_STAGE_DEGRID) && (!grid || !vis_out))) {
    set_error(RCG_ERR_ARG, "rcg_tiled_graph_create: bad argument");
    return RCG_ERR_ARG;
  }
  *out = nullptr;
  rcg_graph *g = new (std::nothrow) rcg_graph();
  if (!g) {
    set_error(RCG_ERR_NOMEM, "host allocation failed");
    return RCG_ERR_NOMEM;
  }
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    if (st) (void)hipStreamDestroy(st);
    delete g;
    set_error(RCG_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
    return RCG_ERR_HIP;
  }
  const bool timing = pl->timing;
  pl->timing = false;  // no event records inside the graph
  rcg_status s = RCG_OK;
  if (stages & RCG_STAGE_BIN) s = rcg_tiled_bin(pl, gcf, st);
  if (s == RCG_OK && (stages & RCG_STAGE_GRID)) s = rcg_grid_tiled_f32(pl, gcf, vis, grid, st);
  if (s == RCG_OK && (stages & RCG_STAGE_DEGRID))
    s = rcg_degrid_tiled_f32(pl, gcf, grid, vis_out, st);
  pl->timing = timing;
  e = hipStreamEndCapture(st, &g->graph);
  if (s == RCG_OK && e != hipSuccess) {
    set_error(RCG_ERR_HIP, "stream capture: %s", hipGetErrorString(e));
    s = RCG_ERR_HIP;
  }
  if (s == RCG_OK) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error(RCG_ERR_HIP, "graph instantiation: %s", hipGetErrorString(e));
      s = RCG_ERR_HIP;
    }
  }
  (void)hipStreamDestroy(st);
  if (s != RCG_OK) {
    rcg_graph_destroy(g);
    return s;
  }
  *out = g;
  return RCG_OK;
}

extern "C" rcg_status rcg_graph_launch(rcg_graph *g, void *stream)
{
  clear_error();
  if (!g || !g->exec) {
    set_error(RCG_ERR_ARG, "rcg_graph_launch: null graph");
    return RCG_ERR_ARG;
  }
  RCG_HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return RCG_OK;
}
