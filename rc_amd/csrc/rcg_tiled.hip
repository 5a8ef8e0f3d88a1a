// rcg_tiled.hip -- FP32 gridder/degridder over uv tiles held in LDS.
//
// Same mathematics as the reference's scatter gridder and its transpose
// degridder (MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p
// .cpp:146-176), organised for CDNA4:
//
//  1. pregrid (FP64, exact; rcg_pregrid.hip) -> one record per point;
//  2. binning: the grid is cut into T x T tiles; every in-bound point is
//     listed in each tile its S x S window overlaps (a point near a tile
//     edge appears in up to four tiles when S <= T).  A counting sort builds
//     the per-tile lists without host synchronisation: LDS histograms per
//     workgroup -> global counts -> one-workgroup scan -> LDS-ranked fill;
//  3. work items: each tile's list is cut into chunks of at most `chunk`
//     pairs; workgroup b takes item b (XCD-aware order) and exits when b is
//     past the device-side item count, so the launch size is an upper
//     bound known on the host and nothing waits on the host;
//  4. gridding: the workgroup accumulates its items' taps into an FP32 tile
//     in LDS (ds_add_f32), lanes of a wavefront covering the rows and
//     columns of the window/tile intersection, GCF rows read coalesced;
//     the tile is stored to HBM with plain stores when the tile has a single
//     item, with FP32 atomics (non-zero cells only) otherwise;
//  5. degridding: the workgroup stages its grid tile in LDS, each wavefront
//     forms the partial sum of one pair over the intersection and reduces it
//     across lanes; points whose window lies in one tile are stored, split
//     points are added atomically.
#include <climits>

#include "rcg_internal.h"
#include "rcg_tiled.h"

namespace rcg {

// Pair record: one (point, tile) incidence, 16 bytes.
//   x: u0 << 16 | v0 (window top-left cell, both < 32768)
//   y: tap0 (element offset of the recentred window)
//   z: stride | supp << 16 | conj << 31
//   w: point index
__device__ __forceinline__ int4 make_pair(int u0, int v0, const rcg_pg &r, int point)
{
  int4 p;
  p.x = (u0 << 16) | v0;
  p.y = r.tap0;
  p.z = (int)(uint16_t)r.stride | ((int)r.supp << 16) |
        ((r.flags & PG_CONJ) ? (int)0x80000000 : 0);
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
// Each workgroup (BIN_THREADS threads) takes BIN_PPT consecutive points per
// thread.  Tile ranges of a point: rows [u0 >> lt, (u0 + S - 1) >> lt],
// columns likewise.

constexpr int BIN_THREADS = 1024;
constexpr int BIN_PPT = 4;

__device__ __forceinline__ void point_tiles(const rcg_pg &r, int pitch, int lt,
                                            int *u0, int *v0, int *tu0,
                                            int *tu1, int *tv0, int *tv1)
{
  *u0 = r.cell / pitch;
  *v0 = r.cell - *u0 * pitch;
  const int S = r.supp;
  *tu0 = *u0 >> lt;
  *tu1 = (*u0 + S - 1) >> lt;
  *tv0 = *v0 >> lt;
  *tv1 = (*v0 + S - 1) >> lt;
}

__global__ void __launch_bounds__(BIN_THREADS)
k_bin_count(const rcg_pg *__restrict__ pg, long long n, int pitch, int lt,
            int ntx, int n_tiles, int max_pairs_pt, int *__restrict__ tile_count,
            int *__restrict__ err)
{
  extern __shared__ int hist[];
  for (int t = threadIdx.x; t < n_tiles; t += BIN_THREADS) hist[t] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * BIN_THREADS * BIN_PPT;
  for (int k = 0; k < BIN_PPT; k++) {
    const long long i = base + (long long)k * BIN_THREADS + threadIdx.x;
    if (i >= n) break;
    const rcg_pg r = pg[i];
    if (r.cell < 0 || r.supp <= 0) continue;
    int u0, v0, tu0, tu1, tv0, tv1;
    point_tiles(r, pitch, lt, &u0, &v0, &tu0, &tu1, &tv0, &tv1);
    if ((tu1 - tu0 + 1) * (tv1 - tv0 + 1) > max_pairs_pt) {
      atomicOr(err, ERR_CAPACITY);  // support above the plan's max_support
      continue;
    }
    for (int a = tu0; a <= tu1; a++)
      for (int b = tv0; b <= tv1; b++) atomicAdd(&hist[a * ntx + b], 1);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n_tiles; t += BIN_THREADS)
    if (hist[t]) atomicAdd(&tile_count[t], hist[t]);
}

// One workgroup: exclusive scans of the tile counts (pairs) and of the
// per-tile item counts; totals[0] = pairs, totals[1] = items; clears the
// fill cursors.
constexpr int SCAN_THREADS = 1024;

__global__ void __launch_bounds__(SCAN_THREADS)
k_bin_scan(const int *__restrict__ tile_count, int n_tiles, int chunk,
           int *__restrict__ tile_start, int *__restrict__ item_start,
           int *__restrict__ tile_cursor, int *__restrict__ totals)
{
  __shared__ long long s_pairs[SCAN_THREADS];
  __shared__ long long s_items[SCAN_THREADS];
  const int per = (n_tiles + SCAN_THREADS - 1) / SCAN_THREADS;
  const int t0 = threadIdx.x * per;
  long long lp = 0, li = 0;
  for (int t = t0; t < t0 + per && t < n_tiles; t++) {
    const int c = tile_count[t];
    lp += c;
    li += (c + chunk - 1) / chunk;
  }
  s_pairs[threadIdx.x] = lp;
  s_items[threadIdx.x] = li;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 partial sums
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    long long ap = 0, ai = 0;
    if (threadIdx.x >= off) {
      ap = s_pairs[threadIdx.x - off];
      ai = s_items[threadIdx.x - off];
    }
    __syncthreads();
    s_pairs[threadIdx.x] += ap;
    s_items[threadIdx.x] += ai;
    __syncthreads();
  }
  long long rp = s_pairs[threadIdx.x] - lp;
  long long ri = s_items[threadIdx.x] - li;
  for (int t = t0; t < t0 + per && t < n_tiles; t++) {
    const int c = tile_count[t];
    tile_start[t] = (int)rp;
    item_start[t] = (int)ri;
    tile_cursor[t] = 0;
    rp += c;
    ri += (c + chunk - 1) / chunk;
  }
  if (threadIdx.x == SCAN_THREADS - 1) {
    totals[0] = (int)s_pairs[SCAN_THREADS - 1];
    totals[1] = (int)s_items[SCAN_THREADS - 1];
  }
}

// Fill: LDS ranks within the workgroup, one global atomic per touched tile
// for the workgroup's base, then every pair is written to its slot.
__global__ void __launch_bounds__(BIN_THREADS)
k_bin_fill(const rcg_pg *__restrict__ pg, long long n, int pitch, int lt,
           int ntx, int n_tiles, int max_pairs_pt,
           const int *__restrict__ tile_start, int *__restrict__ tile_cursor,
           int4 *__restrict__ pairs, long long pair_cap, int *__restrict__ err)
{
  extern __shared__ int lds[];
  int *hist = lds;                                        // n_tiles
  uint16_t *rank = reinterpret_cast<uint16_t *>(lds + n_tiles);  // per pair
  for (int t = threadIdx.x; t < n_tiles; t += BIN_THREADS) hist[t] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * BIN_THREADS * BIN_PPT;
  for (int k = 0; k < BIN_PPT; k++) {
    const long long i = base + (long long)k * BIN_THREADS + threadIdx.x;
    if (i >= n) break;
    const rcg_pg r = pg[i];
    if (r.cell < 0 || r.supp <= 0) continue;
    int u0, v0, tu0, tu1, tv0, tv1;
    point_tiles(r, pitch, lt, &u0, &v0, &tu0, &tu1, &tv0, &tv1);
    if ((tu1 - tu0 + 1) * (tv1 - tv0 + 1) > max_pairs_pt) continue;
    int j = 0;
    for (int a = tu0; a <= tu1; a++)
      for (int b = tv0; b <= tv1; b++, j++)
        rank[((k * BIN_THREADS) + threadIdx.x) * max_pairs_pt + j] =
            (uint16_t)atomicAdd(&hist[a * ntx + b], 1);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n_tiles; t += BIN_THREADS) {
    const int c = hist[t];
    hist[t] = c ? tile_start[t] + atomicAdd(&tile_cursor[t], c) : 0;
  }
  __syncthreads();
  for (int k = 0; k < BIN_PPT; k++) {
    const long long i = base + (long long)k * BIN_THREADS + threadIdx.x;
    if (i >= n) break;
    const rcg_pg r = pg[i];
    if (r.cell < 0 || r.supp <= 0) continue;
    int u0, v0, tu0, tu1, tv0, tv1;
    point_tiles(r, pitch, lt, &u0, &v0, &tu0, &tu1, &tv0, &tv1);
    if ((tu1 - tu0 + 1) * (tv1 - tv0 + 1) > max_pairs_pt) continue;
    const int4 pr = make_pair(u0, v0, r, (int)i);
    int j = 0;
    for (int a = tu0; a <= tu1; a++)
      for (int b = tv0; b <= tv1; b++, j++) {
        const long long pos =
            (long long)hist[a * ntx + b] +
            rank[((k * BIN_THREADS) + threadIdx.x) * max_pairs_pt + j];
        if (pos < pair_cap) pairs[pos] = pr;
        else atomicOr(err, ERR_CAPACITY);
      }
  }
}

// ---------------------------------------------------------------------------
// work-item lookup: the largest t with item_start[t] <= item.  Tiles after
// the owning tile start past `item`, so this is the owner (empty tiles
// before it share smaller indices).
__device__ __forceinline__ int find_tile(const int *__restrict__ item_start,
                                         const int *__restrict__ tile_count,
                                         int n_tiles, int item)
{
  int lo = 0, hi = n_tiles - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (item_start[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int xcd_item(int b, int total)
{
  // Workgroups are dispatched round-robin over the 8 XCDs (b mod 8); give
  // each XCD a contiguous range of the `total` items so a tile's items and
  // neighbouring tiles share an L2.  Blocks past the ranges return -1.
  const int per = (total + 7) >> 3;
  const int k = b >> 3;
  return k < per ? (b & 7) * per + k : -1;
}

struct ItemRange {
  int tile, begin, end, single;
};

__device__ __forceinline__ bool item_range(const TiledDev &d, ItemRange *ir)
{
  const int total = d.totals[1];
  const int item = xcd_item(blockIdx.x, total);
  if (item < 0 || item >= total) return false;
  const int t = find_tile(d.item_start, d.tile_count, d.n_tiles, item);
  const int k = item - d.item_start[t];
  const int cnt = d.tile_count[t];
  ir->tile = t;
  ir->begin = d.tile_start[t] + k * d.chunk;
  ir->end = d.tile_start[t] + min(cnt, (k + 1) * d.chunk);
  ir->single = cnt <= d.chunk;
  return true;
}

// ---------------------------------------------------------------------------
// Walking a pair's window/tile intersection.
//
// A wavefront loads the records (and, for gridding, the visibilities) of up
// to 64 pairs at once, one per lane, and then walks them in order, the pair
// fields broadcast from their lane with readlane (uniform, scalar).  For a
// pair whose intersection is h x w cells, lane l takes elements
// e = l, l + 64, l + 128, ... of the row-major flattening; each step moves
// its (row, col) by (64 / w, 64 % w) with at most one wrap, and its tap and
// LDS offsets by precomputed increments, so the loop has no multiplies and
// no branches: a lane past the end loads its tap through a buffer
// descriptor at an offset beyond the table (which reads 0) and points its
// LDS access at a pad cell.  JOB_U steps are issued before they are
// consumed, so JOB_U tap loads per lane are in flight.

constexpr int JOB_U = 4;
constexpr unsigned TAP_OOB = 0x7ffffff0u;  // byte offset past any tap table

__device__ __forceinline__ int rl(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ float rlf(float v, int j)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t tap_rsrc(const float2 *taps, long long n)
{
  // raw buffer over the tap table; reads past num_records return 0
  const long long bytes = n * 8 < 0x7fffffffLL ? n * 8 : 0x7fffffffLL;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(taps), (short)0,
                                           (int)bytes, 0x00020000);
}

__device__ __forceinline__ float2 tap_load(__amdgpu_buffer_rsrc_t r, unsigned off)
{
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return make_float2(__int_as_float(v[0]), __int_as_float(v[1]));
}

struct Walk {
  // uniform
  int h, w, K, point, whole, conj;
  int dr, dc, tstep, twrap, lstep, lwrap;
  // per lane: row, column, tap byte offset, LDS element index
  int r, c, tb, li;
};

template <int LS>
__device__ __forceinline__ Walk walk_pair(const int4 &my, int j, int tu0, int tv0,
                                          int T, int lane)
{
  Walk k;
  const int px = rl(my.x, j), py = rl(my.y, j), pz = rl(my.z, j);
  k.point = rl(my.w, j);
  const int u0 = px >> 16, v0 = px & 0xffff;
  const int S = (pz >> 16) & 0x7fff;
  const int stride = pz & 0xffff;
  k.conj = pz < 0;
  const int ra = max(u0, tu0), rb = min(u0 + S, tu0 + T);
  const int ca = max(v0, tv0), cb = min(v0 + S, tv0 + T);
  k.h = rb - ra;
  k.w = cb - ca;
  k.whole = (k.h == S) && (k.w == S);
  k.K = (k.h * k.w + 63) >> 6;
  // integer quotients by w through the approximate reciprocal: (x + 0.5)/w
  // is at least 0.5/w from an integer, far above the rounding error
  const float rw = __builtin_amdgcn_rcpf((float)k.w);
  k.dr = (int)(64.5f * rw);
  k.dc = 64 - k.dr * k.w;
  k.tstep = (k.dr * stride + k.dc) * 8;
  k.twrap = (stride - k.w) * 8;
  k.lstep = k.dr * LS + k.dc;
  k.lwrap = LS - k.w;
  k.r = (int)(((float)lane + 0.5f) * rw);
  k.c = lane - k.r * k.w;
  k.tb = (py + (ra - u0 + k.r) * stride + (ca - v0) + k.c) * 8;
  k.li = (ra - tu0 + k.r) * LS + (ca - tv0) + k.c;
  return k;
}

__device__ __forceinline__ void walk_step(Walk &k)
{
  k.c += k.dc;
  const bool m = k.c >= k.w;
  k.c = m ? k.c - k.w : k.c;
  k.r += k.dr + (m ? 1 : 0);
  k.tb += k.tstep + (m ? k.twrap : 0);
  k.li += k.lstep + (m ? k.lwrap : 0);
}

// ---------------------------------------------------------------------------
// gridding: each wavefront owns a private copy of the tile in LDS, so the
// read-modify-write of a cell needs no atomic: the cells of one pair's steps
// are distinct, and a wavefront's LDS operations complete in program order;
// the four copies are summed at the flush.  Column T of every row is padding
// that absorbs the (zero) updates of lanes past the end of a pair.

template <int T>
__global__ void __launch_bounds__(256)
k_grid_tiles(TiledDev d, const float2 *__restrict__ vis,
             const float2 *__restrict__ taps, long long n_taps,
             float2 *__restrict__ grid)
{
  constexpr int LS = T + 1;
  constexpr int CP = T * LS;  // one copy
  __shared__ float2 acc[4 * CP];
  ItemRange ir;
  if (!item_range(d, &ir)) return;
  for (int q = threadIdx.x; q < 4 * CP; q += 256) acc[q] = make_float2(0.f, 0.f);
  __syncthreads();
  const __amdgpu_buffer_rsrc_t rs = tap_rsrc(taps, n_taps);
  const int tu0 = (ir.tile / d.ntx) * T, tv0 = (ir.tile % d.ntx) * T;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float2 *my_acc = acc + wave * CP;
  const int pad = (lane & 31) * LS + T;
  for (int base = ir.begin + wave * 64; base < ir.end; base += 256) {
    const int nb = min(64, ir.end - base);
    int4 my = make_int4(0, 0, 0, 0);
    float2 mv = make_float2(0.f, 0.f);
    if (lane < nb) {
      my = d.pairs[base + lane];
      mv = vis[my.w];
    }
    for (int j = 0; j < nb; j++) {
      Walk k = walk_pair<LS>(my, j, tu0, tv0, T, lane);
      // contribution of tap t (conjugated for mirrored w):
      //   re = a*t.x + b*t.y, im = c*t.y + e*t.x
      const float vx = rlf(mv.x, j), vy = rlf(mv.y, j);
      const float a = vx, b = k.conj ? vy : -vy, c = k.conj ? -vx : vx, e = vy;
      for (int s0 = 0; s0 < k.K; s0 += JOB_U) {
        int li[JOB_U];
        float2 t[JOB_U];
#pragma unroll
        for (int u = 0; u < JOB_U; u++) {
          const bool act = (s0 + u < k.K) && (k.r < k.h);
          li[u] = act ? k.li : pad;
          t[u] = tap_load(rs, act ? (unsigned)k.tb : TAP_OOB);
          walk_step(k);
        }
        float2 x[JOB_U];
#pragma unroll
        for (int u = 0; u < JOB_U; u++) x[u] = my_acc[li[u]];
#pragma unroll
        for (int u = 0; u < JOB_U; u++) {
          x[u].x = fmaf(a, t[u].x, fmaf(b, t[u].y, x[u].x));
          x[u].y = fmaf(c, t[u].y, fmaf(e, t[u].x, x[u].y));
        }
#pragma unroll
        for (int u = 0; u < JOB_U; u++) my_acc[li[u]] = x[u];
      }
    }
  }
  __syncthreads();
  // flush: sum the four copies
  const int gsz = d.grid_size;
  for (int q = threadIdx.x; q < T * T; q += 256) {
    const int r = q / T, cc = q - (q / T) * T;
    const int gu = tu0 + r, gv = tv0 + cc;
    if (gu >= gsz || gv >= gsz) continue;
    const int l = r * LS + cc;
    const float2 s0 = acc[l], s1 = acc[CP + l], s2 = acc[2 * CP + l], s3 = acc[3 * CP + l];
    const float re = (s0.x + s1.x) + (s2.x + s3.x);
    const float im = (s0.y + s1.y) + (s2.y + s3.y);
    float2 *g = grid + (long long)gu * d.pitch + gv;
    if (ir.single) {
      *g = make_float2(re, im);
    } else if (re != 0.f || im != 0.f) {
      unsafeAtomicAdd(&g->x, re);
      unsafeAtomicAdd(&g->y, im);
    }
  }
}

// ---------------------------------------------------------------------------
// degridding: the grid tile is staged once in LDS (read-only); per pair the
// lanes accumulate four partial products and reduce across the wavefront.

__device__ __forceinline__ float wave_sum(float x)
{
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

template <int T>
__global__ void __launch_bounds__(256)
k_degrid_tiles(TiledDev d, const float2 *__restrict__ grid,
               const float2 *__restrict__ taps, long long n_taps,
               float2 *__restrict__ vis)
{
  constexpr int LS = T + 1;
  __shared__ float2 tile[T * LS];
  ItemRange ir;
  if (!item_range(d, &ir)) return;
  const int tu0 = (ir.tile / d.ntx) * T, tv0 = (ir.tile % d.ntx) * T;
  const int gsz = d.grid_size;
  for (int q = threadIdx.x; q < T * T; q += 256) {
    const int r = q / T, c = q - (q / T) * T;
    const int gu = tu0 + r, gv = tv0 + c;
    tile[r * LS + c] = (gu < gsz && gv < gsz) ? grid[(long long)gu * d.pitch + gv]
                                              : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const __amdgpu_buffer_rsrc_t rs = tap_rsrc(taps, n_taps);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int base = ir.begin + wave * 64; base < ir.end; base += 256) {
    const int nb = min(64, ir.end - base);
    int4 my = make_int4(0, 0, 0, 0);
    if (lane < nb) my = d.pairs[base + lane];
    for (int j = 0; j < nb; j++) {
      Walk k = walk_pair<LS>(my, j, tu0, tv0, T, lane);
      // P = sum gx*tx, Q = sum gy*ty, R = sum gx*ty, U = sum gy*tx
      float P = 0.f, Q = 0.f, R = 0.f, U = 0.f;
      for (int s0 = 0; s0 < k.K; s0 += JOB_U) {
        float2 t[JOB_U], g[JOB_U];
#pragma unroll
        for (int u = 0; u < JOB_U; u++) {
          const bool act = (s0 + u < k.K) && (k.r < k.h);
          t[u] = tap_load(rs, act ? (unsigned)k.tb : TAP_OOB);
          g[u] = tile[act ? k.li : 0];
          walk_step(k);
        }
#pragma unroll
        for (int u = 0; u < JOB_U; u++) {
          P = fmaf(g[u].x, t[u].x, P);
          Q = fmaf(g[u].y, t[u].y, Q);
          R = fmaf(g[u].x, t[u].y, R);
          U = fmaf(g[u].y, t[u].x, U);
        }
      }
      // tap t' = (tx, sg*ty): re = P - sg*Q, im = sg*R + U
      float re = k.conj ? P + Q : P - Q;
      float im = k.conj ? U - R : U + R;
      re = wave_sum(re);
      im = wave_sum(im);
      if (lane == 0) {
        float2 *o = vis + k.point;
        if (k.whole) {
          *o = make_float2(re, im);
        } else {
          unsafeAtomicAdd(&o->x, re);
          unsafeAtomicAdd(&o->y, im);
        }
      }
    }
  }
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

}  // namespace

rcg_status tiled_blocks(const rcg_tiled_plan *pl, int *nb)
{
  // upper bound on work items: one per tile plus one per `chunk` pairs
  long long ub = (long long)pl->dev.n_tiles + pl->dev.pair_cap / pl->dev.chunk + 1;
  ub = ((ub + 7) & ~7LL) + 8;
  if (ub > INT_MAX) {
    set_error(RCG_ERR_ARG, "too many work items (%lld)", ub);
    return RCG_ERR_ARG;
  }
  *nb = (int)ub;
  return RCG_OK;
}

}  // namespace rcg

using namespace rcg;

extern "C" rcg_status rcg_tiled_plan_create(const rcg_geom *geom, int tile,
                                            int chunk, int max_support,
                                            rcg_tiled_plan **out)
{
  clear_error();
  if (!geom || !out) {
    set_error(RCG_ERR_ARG, "rcg_tiled_plan_create: null argument");
    return RCG_ERR_ARG;
  }
  *out = nullptr;
  rcg_status s = check_geom(geom);
  if (s != RCG_OK) return s;
  if (tile != 32 && tile != 64 && tile != 128) {
    set_error(RCG_ERR_ARG, "tile must be 32, 64 or 128 (got %d)", tile);
    return RCG_ERR_ARG;
  }
  if (chunk < 1 || max_support < 1 || max_support > MAX_SUPPORT) {
    set_error(RCG_ERR_ARG, "chunk %d / max_support %d out of range", chunk, max_support);
    return RCG_ERR_ARG;
  }
  const int ntx = (geom->grid_size + tile - 1) / tile;
  const long long n_tiles = (long long)ntx * ntx;
  if (n_tiles > MAX_BIN_TILES) {
    set_error(RCG_ERR_ARG, "%lld tiles exceed the binning limit %d: use a larger tile",
              n_tiles, MAX_BIN_TILES);
    return RCG_ERR_ARG;
  }
  const int per_dim = (max_support - 1 + tile - 1) / tile + 1;
  const int max_pairs_pt = per_dim * per_dim;
  const long long n = (long long)geom->n_chunks * geom->ts_ch;
  rcg_tiled_plan *pl = new (std::nothrow) rcg_tiled_plan();
  if (!pl) {
    set_error(RCG_ERR_NOMEM, "host allocation failed");
    return RCG_ERR_NOMEM;
  }
  pl->geom = *geom;
  pl->n = n;
  pl->max_support = max_support;
  pl->max_pairs_pt = max_pairs_pt;
  TiledDev &d = pl->dev;
  d.T = tile;
  d.lt = ilog2(tile);
  d.ntx = ntx;
  d.n_tiles = (int)n_tiles;
  d.chunk = chunk;
  d.pitch = geom->grid_pitch;
  d.grid_size = geom->grid_size;
  d.pair_cap = n * max_pairs_pt;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bin_fill),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     160 * 1024);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bin_count),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e == hipSuccess) e = dalloc(&pl->pg, (size_t)n);
  if (e == hipSuccess) e = dalloc(&d.tile_count, (size_t)n_tiles);
  if (e == hipSuccess) e = dalloc(&d.tile_start, (size_t)n_tiles);
  if (e == hipSuccess) e = dalloc(&d.tile_cursor, (size_t)n_tiles);
  if (e == hipSuccess) e = dalloc(&d.item_start, (size_t)n_tiles);
  if (e == hipSuccess) e = dalloc(&d.totals, 4);
  if (e == hipSuccess) e = dalloc(&pl->taps, 1);
  if (e == hipSuccess) e = dalloc(&pl->err, 1);
  if (e == hipSuccess) e = dalloc(&d.pairs, (size_t)d.pair_cap);
  if (e != hipSuccess) {
    rcg_tiled_plan_destroy(pl);
    set_error(RCG_ERR_NOMEM, "device allocation failed: %s", hipGetErrorString(e));
    return RCG_ERR_NOMEM;
  }
  *out = pl;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_plan_destroy(rcg_tiled_plan *pl)
{
  if (!pl) return RCG_OK;
  TiledDev &d = pl->dev;
  void *ptrs[] = {pl->pg, d.tile_count, d.tile_start, d.tile_cursor,
                  d.item_start, d.totals, pl->taps, pl->err, d.pairs};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete pl;
  return RCG_OK;
}

extern "C" rcg_status rcg_tiled_bin(rcg_tiled_plan *pl, const rcg_geom *geom,
                                    const rcg_gcf *gcf, void *stream)
{
  clear_error();
  if (!pl || !geom || !gcf) {
    set_error(RCG_ERR_ARG, "rcg_tiled_bin: null argument");
    return RCG_ERR_ARG;
  }
  if ((long long)geom->n_chunks * geom->ts_ch != pl->n ||
      geom->grid_size != pl->geom.grid_size || geom->grid_pitch != pl->geom.grid_pitch) {
    set_error(RCG_ERR_ARG, "geometry does not match the plan");
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  TiledDev &d = pl->dev;
  const long long n = pl->n;
  hipLaunchKernelGGL(k_fill_i32, dim3(64), dim3(256), 0, st, d.tile_count, 0,
                     (long long)d.n_tiles);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d.totals, 0, 4LL);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st,
                     reinterpret_cast<int *>(pl->taps), 0, 2LL);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, pl->err, 0, 1LL);
  RCG_HIP_TRY(hipGetLastError());
  rcg_status s = rcg_pregrid(geom, gcf, pl->pg, pl->taps, pl->err, stream);
  if (s != RCG_OK) return s;
  if (n == 0) return RCG_OK;
  const int nb = ceil_div(n, (long long)BIN_THREADS * BIN_PPT);
  const size_t lds_count = (size_t)d.n_tiles * sizeof(int);
  hipLaunchKernelGGL(k_bin_count, dim3(nb), dim3(BIN_THREADS), lds_count, st,
                     pl->pg, n, d.pitch, d.lt, d.ntx, d.n_tiles, pl->max_pairs_pt,
                     d.tile_count, pl->err);
  RCG_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(SCAN_THREADS), 0, st, d.tile_count,
                     d.n_tiles, d.chunk, d.tile_start, d.item_start, d.tile_cursor,
                     d.totals);
  RCG_HIP_TRY(hipGetLastError());
  const size_t lds_fill = lds_count + (size_t)BIN_THREADS * BIN_PPT *
                                          pl->max_pairs_pt * sizeof(uint16_t);
  if (lds_fill > 160 * 1024) {
    set_error(RCG_ERR_ARG, "binning needs %zu bytes of LDS", lds_fill);
    return RCG_ERR_ARG;
  }
  hipLaunchKernelGGL(k_bin_fill, dim3(nb), dim3(BIN_THREADS), lds_fill, st, pl->pg,
                     n, d.pitch, d.lt, d.ntx, d.n_tiles, pl->max_pairs_pt,
                     d.tile_start, d.tile_cursor, d.pairs, d.pair_cap, pl->err);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_grid_tiled_f32(rcg_tiled_plan *pl, const rcg_gcf *gcf,
                                         const rcg_c64 *vis, rcg_c64 *grid,
                                         void *stream)
{
  clear_error();
  if (!pl || !gcf || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_grid_tiled_f32: null argument");
    return RCG_ERR_ARG;
  }
  if (gcf->n_taps <= 0 || gcf->n_taps >= (0x7fffffffLL >> 3)) {
    set_error(RCG_ERR_ARG, "tap table of %lld elements: the FP32 path needs 0 < n_taps < 2^28",
              (long long)gcf->n_taps);
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const TiledDev &d = pl->dev;
  const long long cells = (long long)d.grid_size * d.pitch * 2;
  hipLaunchKernelGGL(k_fill_f32, dim3(2048), dim3(256), 0, st,
                     reinterpret_cast<float *>(grid), cells);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  int nb;
  rcg_status s = tiled_blocks(pl, &nb);
  if (s != RCG_OK) return s;
  const float2 *v = reinterpret_cast<const float2 *>(vis);
  const float2 *t = reinterpret_cast<const float2 *>(gcf->data);
  const long long nt = gcf->n_taps;
  float2 *g = reinterpret_cast<float2 *>(grid);
  switch (d.T) {
    case 32: hipLaunchKernelGGL(k_grid_tiles<32>, dim3(nb), dim3(256), 0, st, d, v, t, nt, g); break;
    case 64: hipLaunchKernelGGL(k_grid_tiles<64>, dim3(nb), dim3(256), 0, st, d, v, t, nt, g); break;
    default:
      set_error(RCG_ERR_UNSUPPORTED, "tile %d not supported by the FP32 gridder", d.T);
      return RCG_ERR_UNSUPPORTED;
  }
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
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
  if (gcf->n_taps <= 0 || gcf->n_taps >= (0x7fffffffLL >> 3)) {
    set_error(RCG_ERR_ARG, "tap table of %lld elements: the FP32 path needs 0 < n_taps < 2^28",
              (long long)gcf->n_taps);
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const TiledDev &d = pl->dev;
  hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(256), 0, st,
                     reinterpret_cast<float *>(vis), pl->n * 2);
  RCG_HIP_TRY(hipGetLastError());
  if (pl->n == 0) return RCG_OK;
  int nb;
  rcg_status s = tiled_blocks(pl, &nb);
  if (s != RCG_OK) return s;
  const float2 *g = reinterpret_cast<const float2 *>(grid);
  const float2 *t = reinterpret_cast<const float2 *>(gcf->data);
  const long long nt = gcf->n_taps;
  float2 *v = reinterpret_cast<float2 *>(vis);
  switch (d.T) {
    case 32: hipLaunchKernelGGL(k_degrid_tiles<32>, dim3(nb), dim3(256), 0, st, d, g, t, nt, v); break;
    case 64: hipLaunchKernelGGL(k_degrid_tiles<64>, dim3(nb), dim3(256), 0, st, d, g, t, nt, v); break;
    default:
      set_error(RCG_ERR_UNSUPPORTED, "tile %d not supported by the FP32 degridder", d.T);
      return RCG_ERR_UNSUPPORTED;
  }
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
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
