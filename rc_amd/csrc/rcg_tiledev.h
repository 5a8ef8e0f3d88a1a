// rcg_tiledev.h -- device helpers shared by the tiled gridding kernels:
// persistent work-item queues, tap loads through a raw buffer, record
// streaming and the wavefront sums of the degridders.
#pragma once

#include "rcg_internal.h"
#include "rcg_tiled.h"

namespace rcg {

// ---------------------------------------------------------------------------
// work items

// The largest t with item_start[t] <= item.  Tiles after the owning tile
// start past `item`, so this is the owner (empty tiles before it share
// smaller indices).
__device__ __forceinline__ int find_tile(const int *__restrict__ item_start, int n_tiles,
                                         int item)
{
  int lo = 0, hi = n_tiles - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (item_start[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Next work item of a workgroup (called by one thread).  The `total` items
// are cut into 8 contiguous ranges, one per XCD: workgroups are dispatched
// round-robin over the XCDs, so workgroup b runs on XCD b mod 8 and draws
// from range b mod 8 first, then from the others in turn (`tried` counts
// the ranges found empty).  Returns -1 when every range is used up; each
// counter only grows, so every workgroup ends after at most 8 failed draws.
__device__ __forceinline__ int draw_item(int *__restrict__ q, int total, int home, int &tried)
{
  const int per = (total + 7) >> 3;
  while (tried < 8) {
    const int x = (home + tried) & 7;
    const int k = atomicAdd(&q[x], 1);
    if (k < per && x * per + k < total) return x * per + k;
    tried++;
  }
  return -1;
}

struct ItemRange {
  int item;                // index of the work item
  int begin, end, single;  // pairs [begin, end); the tile has one item
  int tu0, tv0;            // first grid row and column of the tile
  long long goff;          // element offset of the instance's grid
};

__device__ __forceinline__ ItemRange item_of(const TiledDev &d, int item)
{
  const int t = find_tile(d.item_start, d.n_tiles, item);
  const int k = item - d.item_start[t];
  const int cnt = d.tile_count[t];
  const int inst = t / d.tpi, lt = t - inst * d.tpi;
  ItemRange ir;
  ir.item = item;
  ir.begin = d.tile_start[t] + k * d.chunk;
  ir.end = d.tile_start[t] + min(cnt, (k + 1) * d.chunk);
  ir.single = cnt <= d.chunk;
  ir.tu0 = (lt / d.ntx) * d.T;
  ir.tv0 = (lt % d.ntx) * d.T;
  ir.goff = (long long)inst * d.gstride;
  return ir;
}

// Work loop of a persistent workgroup: fn(item range) for every item it
// draws.  The trailing barrier keeps the next draw (and the next item's
// LDS writes) behind every wavefront's use of the current one.
template <class F>
__device__ __forceinline__ void for_each_item(const TiledDev &d, int *__restrict__ q, F fn)
{
  __shared__ int s_item;
  const int total = d.totals[1];
  int tried = 0;
  while (true) {
    if (threadIdx.x == 0) s_item = draw_item(q, total, blockIdx.x & 7, tried);
    __syncthreads();
    const int item = s_item;
    if (item < 0) break;
    fn(item_of(d, item));
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Taps and records.

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

// voff: per-lane byte offset (checked against the table size); soff:
// uniform byte offset added to it.
__device__ __forceinline__ float2 tap_load(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return make_float2(__int_as_float(v[0]), __int_as_float(v[1]));
}

__device__ __forceinline__ float2 &lds_at(float2 *lds, int byte)
{
  return *reinterpret_cast<float2 *>(reinterpret_cast<char *>(lds) + byte);
}

// A register move the compiler keeps where it is written: the batch
// rotation moves the records (and visibilities) of the coming batches into
// the registers of the earlier ones before it issues the next loads, so
// those loads can land in the registers the moves free.  Left to the
// compiler, the moves become copies at the end of the block, after the
// loads, which then target temporaries whose copies wait for them.
__device__ __forceinline__ int vmove(int x)
{
  int r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

__device__ __forceinline__ float vmovef(float x)
{
  return __int_as_float(vmove(__float_as_int(x)));
}

// A wavefront's pairs are ir.begin + wave + nw k (nw wavefronts in the
// workgroup), k < npw, each cut into groups of GRP slabs.  Their records come
// 64 at a time, one per lane (batch b: pairs 64 b .. 64 b + 63), and are
// broadcast with readlane.  The kernels load records two batches ahead (and
// the gridder the visibilities one batch ahead, from records that arrived a
// batch earlier), so nothing loaded at a batch boundary is used there: a
// use would wait for the load, and a wavefront's loads complete in order, so
// that wait would also drain the taps in flight.  Lanes past the last pair
// load the last pair again (never decoded), which keeps the load free of
// branches.
__device__ __forceinline__ int4 batch_records(const TiledDev &d, const ItemRange &ir, int wave,
                                              int lane, int npw, int b, int nw)
{
  const int k = min(b * 64 + lane, npw - 1);
  return d.pairs[ir.begin + wave + nw * k];
}

// Error of the rounded sum fl(a + b) (Knuth's TwoSum): fl(a + b) + err =
// a + b exactly, in round-to-nearest without overflow.
__device__ __forceinline__ float two_sum_err(float a, float b)
{
#pragma clang fp contract(off)
  const float s = a + b;
  const float bp = s - a;
  const float ap = s - bp;
  return (a - ap) + (b - bp);
}

// Sums of 4 pairs at once.  Pair k of a wavefront (in the order it walks
// them) has its output code in codes[k % 128], written by lane k % 64 when
// its batch of 64 records becomes current, and when it ends every lane
// leaves its part of the pair's sum in row k % 4 of the wavefront's LDS
// buffer.  After
// pairs 4a..4a+3 (or at the end of an item) lane l sums entries c, c + 16,
// c + 32, c + 48 of row l / 16 (c = l % 16), then a DPP prefix sum over its
// row of 16 lanes (row_shr 1, 2, 4, 8 with zero fill) leaves the pair's
// total in the row's last lane, which stores it.  The order of the
// additions is fixed, so the result is reproducible.
constexpr int SUMS = 4;
// Output addresses are typed global, so the final store is a global (not
// flat) store and does not hold up the wavefront's LDS waits; the (re, im)
// pair is stored as one 64-bit word, re in the low half.
typedef __attribute__((address_space(1))) unsigned long long gu64;

__device__ __forceinline__ float row_sum(float x)
{
  int v = __float_as_int(x);
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
  v = __float_as_int(x);
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));
  v = __float_as_int(x);
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));
  v = __float_as_int(x);
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));
  return x;
}

// Output of a pair, from its record: ~point for a point whose window lies
// in one tile (its visibility), slot * npts + point for a split point (its
// partial); below 2^31 because plans hold at most 2^31 - 1 pairs.
__device__ __forceinline__ int out_code(const int4 &rec, int npts)
{
  const int whole = rec.x >> 31;  // all ones for a whole pair
  const int split = ((rec.x >> 24) & 63) * npts + rec.w;
  return (~rec.w & whole) | (split & ~whole);
}

// n: pairs in the buffer (1..SUMS), whose output codes are code[0..n);
// entries of the other rows are stale and their sums are not stored
__device__ __forceinline__ void flush_sums(const float2 *buf, const int *code, float2 *vis,
                                           float2 *partial, int n, int lane)
{
  // the entries come from other lanes of the wavefront: keep the compiler
  // from moving LDS accesses across this point (the hardware runs a
  // wavefront's LDS instructions in order)
  __builtin_amdgcn_wave_barrier();
  const int q = lane >> 4, c = lane & 15;
  const float2 *b = buf + q * 64 + c;
  const float2 a0 = b[0], a1 = b[16], a2 = b[32], a3 = b[48];
  const float re = row_sum((a0.x + a1.x) + (a2.x + a3.x));
  const float im = row_sum((a0.y + a1.y) + (a2.y + a3.y));
  if (c == 15 && q < n) {
    const int o = code[q];
    gu64 *dst = (gu64 *)(o < 0 ? vis + ~o : partial + o);
    *dst = (unsigned long long)__float_as_uint(im) << 32 | __float_as_uint(re);
  }
  __builtin_amdgcn_wave_barrier();
}

}  // namespace rcg
