// rcg_flow.hip -- the MS6 dna-flow visibility binning kernels on the GPU.
//
// Restates MS6/programs/Kernel/Binning.hs:
//   binSizer (:34-97)  w range [low, high] over the points inside the uv
//                      bounding box of the regions (a fold from (0, 0)),
//                      high inflated by (high - low) * 1e-4, then per region
//                      and w bin (start, end, count);
//   binner  (:99-181)  every point appended, in input order, to the output
//                      vector of the (u, v, w) box that the map lookups find,
//                      then zero padding of the unfilled rows.
// Both are Haskell loops over host vectors in the reference; here each is
// one call over host memory, with the per-point work on the GPU.  The
// region and box tables are nested Data.Map lookups there (lookupLE: the
// greatest key <= x); here they are flattened into sorted key arrays with
// offsets, searched by binary search on the device.
//
// binner's output order within a box is input order: the points are sorted
// by box with a stable radix sort of (box, index) pairs.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "rcg_internal.h"

namespace rcg {
namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() { if (s) (void)hipStreamDestroy(s); }
};

#define FL_TRY(expr)                                                          \
  do {                                                                        \
    hipError_t e__ = (expr);                                                  \
    if (e__ != hipSuccess) {                                                  \
      set_error(RCG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e__)); \
      return -1;                                                              \
    }                                                                         \
  } while (0)

// Index in [lo, hi) of the greatest a[j] <= x, or -1 (x NaN: -1).
__device__ __forceinline__ int le_search(const double *__restrict__ a, int lo, int hi, double x)
{
  int r = -1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) {
      r = mid;
      lo = mid + 1;
    } else {
      hi = mid;
    }
  }
  return r;
}

struct BBox {
  double umin, umax, vmin, vmax;
};

__device__ __forceinline__ bool in_bbox(const BBox &b, double u, double v)
{
  return u >= b.umin && u < b.umax && v >= b.vmin && v < b.vmax;
}

// binSizer's fold (Binning.hs:51-56) as a reduction: per block the minimum
// and maximum of 0 and the w of its in-box points, and whether one of them
// is NaN.  part[3 * block + {0, 1, 2}] = low, high, NaN flag.
__global__ void __launch_bounds__(256)
k_fs_range(const double *__restrict__ vis, int n, int width, BBox bb, double *__restrict__ part)
{
  __shared__ double s_lo[4], s_hi[4];
  __shared__ int s_nan;
  if (threadIdx.x == 0) s_nan = 0;
  __syncthreads();
  double lo = 0.0, hi = 0.0;
  int nan = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double *r = vis + (long long)i * width;
    if (!in_bbox(bb, r[0], r[1])) continue;
    const double w = r[2];
    if (w != w) nan = 1;
    else {
      lo = w <= lo ? w : lo;
      hi = w <= hi ? hi : w;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
    lo = a <= lo ? a : lo;
    hi = b <= hi ? hi : b;
  }
  if (nan) atomicOr(&s_nan, 1);
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      lo = s_lo[k] <= lo ? s_lo[k] : lo;
      hi = s_hi[k] <= hi ? hi : s_hi[k];
    }
    part[3 * blockIdx.x] = lo;
    part[3 * blockIdx.x + 1] = hi;
    part[3 * blockIdx.x + 2] = s_nan;
  }
}

// Region lookup of binSizer (Binning.hs:80-81): ukeys[nu] sorted u lows;
// u key j owns v keys [voff[j], voff[j + 1]) of vkeys (sorted) with region
// ids vreg.  count[region * bins + bin] += 1 (Binning.hs:79-84).  Nearby
// points share a counter, so each wavefront adds once per distinct counter.
__global__ void __launch_bounds__(256)
k_fs_count(const double *__restrict__ vis, int n, int width, BBox bb, double low, double high,
           int bins, const double *__restrict__ ukeys, int nu, const int *__restrict__ voff,
           const double *__restrict__ vkeys, const int *__restrict__ vreg,
           unsigned long long *__restrict__ count)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  long long key = -1;
  if (i < n) {
    const double *r = vis + (long long)i * width;
    const double u = r[0], v = r[1], w = r[2];
    if (in_bbox(bb, u, v) && w >= low && w <= high) {
      const int iu = le_search(ukeys, 0, nu, u);
      const int iv = iu < 0 ? -1 : le_search(vkeys, voff[iu], voff[iu + 1], v);
      if (iv >= 0) {
        const double t = (w - low) / (high - low) * (double)bins;
        // GHC's floor :: Double -> Int maps NaN to minBound, clamped to bin 0
        long long b = t == t ? (long long)floor(t) : LLONG_MIN;
        b = b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
        key = (long long)vreg[iv] * bins + b;
      }
    }
  }
  unsigned long long pending = __ballot(key >= 0);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const long long k = __shfl(key, leader);
    const unsigned long long m = __ballot(key == k);
    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&count[k], (unsigned long long)__popcll(m));
    pending &= ~m;
  }
}

// binner's lookup chain (Binning.hs:133-142): w key, then v key, then u key,
// then the range test.  key[i] = the box, or n_boxes when none; val[i] = i.
struct BoxEntry {
  double ul, uh, vl, vh, wl, wh;
  int box;
};

__global__ void __launch_bounds__(256)
k_fb_target(const double *__restrict__ vis, int n, const double *__restrict__ wkeys, int nw,
            const int *__restrict__ wvoff, const double *__restrict__ vkeys,
            const int *__restrict__ vuoff, const double *__restrict__ ukeys,
            const int *__restrict__ uent, const BoxEntry *__restrict__ ent, int n_boxes,
            unsigned *__restrict__ key, unsigned *__restrict__ val)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *r = vis + 5LL * i;
  const double u = r[0], v = r[1], w = r[2];
  int box = n_boxes;
  const int iw = le_search(wkeys, 0, nw, w);
  if (iw >= 0) {
    const int iv = le_search(vkeys, wvoff[iw], wvoff[iw + 1], v);
    if (iv >= 0) {
      const int iu = le_search(ukeys, vuoff[iv], vuoff[iv + 1], u);
      if (iu >= 0) {
        const BoxEntry e = ent[uent[iu]];
        if (e.ul <= u && u < e.uh && e.vl <= v && v < e.vh && e.wl <= w && w < e.wh)
          box = e.box;
      }
    }
  }
  key[i] = (unsigned)box;
  val[i] = (unsigned)i;
}

// Rows per box from the sorted keys: the run of box k is [first, end);
// first[k] and end[k] stay 0 for a box without rows.
__global__ void __launch_bounds__(256)
k_fb_runs(const unsigned *__restrict__ key, int n, int n_boxes, int *__restrict__ first,
          int *__restrict__ end)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned k = key[j];
  if (k >= (unsigned)n_boxes) return;
  if (j == 0 || key[j - 1] != k) first[k] = j;
  if (j == n - 1 || key[j + 1] != k) end[k] = j + 1;
}

// rows[j] = vis[val[j]] for the kept points, in sorted order.
__global__ void __launch_bounds__(256)
k_fb_gather(const double *__restrict__ vis, const unsigned *__restrict__ val, int m,
            double *__restrict__ rows)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 5LL * m) return;
  const long long j = t / 5;
  rows[t] = vis[5LL * val[j] + (t - 5 * j)];
}

int check_vis(const double *vis, int n, int width, int min_width)
{
  if (n < 0 || width < min_width || (long long)n * width > (1LL << 31)) {
    set_error(RCG_ERR_ARG, "bad visibility array: %d rows of width %d", n, width);
    return -1;
  }
  if (n > 0 && !vis) {
    set_error(RCG_ERR_ARG, "null visibility array");
    return -1;
  }
  return 0;
}

}  // namespace
}  // namespace rcg

using namespace rcg;

extern "C" int rcg_flow_bin_sizes(const double *vis, int n, int width, const double *regions,
                                  int n_regions, int bins, void *out, double *low_high)
{
  clear_error();
  if (check_vis(vis, n, width, 3)) return -1;
  if (n_regions <= 0 || !regions || bins <= 0 || !out) {
    set_error(RCG_ERR_ARG, "need at least one region, one bin and an output table");
    return -1;
  }
  BBox bb;
  bb.umin = bb.vmin = INFINITY;
  bb.umax = bb.vmax = -INFINITY;
  for (int r = 0; r < n_regions; r++) {
    const double *g = regions + 4 * r;
    bb.umin = std::min(bb.umin, g[0]);
    bb.umax = std::max(bb.umax, g[1]);
    bb.vmin = std::min(bb.vmin, g[2]);
    bb.vmax = std::max(bb.vmax, g[3]);
  }
  // u low -> (v low -> region); a later region with the same lows replaces
  // an earlier one (Map.fromListWith Map.union, Binning.hs:76-79)
  std::map<double, std::map<double, int>> lut;
  for (int r = 0; r < n_regions; r++) lut[regions[4 * r]][regions[4 * r + 2]] = r;
  std::vector<double> hu, hv;
  std::vector<int> hvoff, hvreg;
  for (auto &ue : lut) {
    hu.push_back(ue.first);
    hvoff.push_back((int)hv.size());
    for (auto &ve : ue.second) {
      hv.push_back(ve.first);
      hvreg.push_back(ve.second);
    }
  }
  hvoff.push_back((int)hv.size());

  double low = 0.0, high = 0.0;
  std::vector<unsigned long long> cnt((size_t)n_regions * bins, 0);
  if (n > 0) {
    StreamGuard sg;
    FL_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    hipStream_t st = sg.s;
    const int nb = std::min(ceil_div(n, 256), 1024);
    DevBuf d_vis, d_part, d_u, d_voff, d_v, d_vreg, d_cnt;
    FL_TRY(d_vis.alloc((size_t)n * width * sizeof(double)));
    FL_TRY(d_part.alloc((size_t)nb * 3 * sizeof(double)));
    FL_TRY(hipMemcpyAsync(d_vis.p, vis, (size_t)n * width * sizeof(double),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fs_range, dim3(nb), dim3(256), 0, st, d_vis.as<double>(), n, width, bb,
                       d_part.as<double>());
    FL_TRY(hipGetLastError());
    std::vector<double> part((size_t)nb * 3);
    FL_TRY(hipMemcpyAsync(part.data(), d_part.p, part.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    FL_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < nb; b++) {
      if (part[3 * b + 2] != 0.0) {
        set_error(RCG_ERR_ARG, "a point inside the regions has w = NaN");
        return -1;
      }
      low = part[3 * b] <= low ? part[3 * b] : low;
      high = part[3 * b + 1] <= high ? high : part[3 * b + 1];
    }
    high = high + (high - low) * 0.0001;  // Binning.hs:57
    FL_TRY(d_u.alloc(hu.size() * sizeof(double)));
    FL_TRY(d_voff.alloc(hvoff.size() * sizeof(int)));
    FL_TRY(d_v.alloc(hv.size() * sizeof(double)));
    FL_TRY(d_vreg.alloc(hvreg.size() * sizeof(int)));
    FL_TRY(d_cnt.alloc(cnt.size() * sizeof(unsigned long long)));
    FL_TRY(hipMemcpyAsync(d_u.p, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice,
                          st));
    FL_TRY(hipMemcpyAsync(d_voff.p, hvoff.data(), hvoff.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    FL_TRY(hipMemcpyAsync(d_v.p, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice,
                          st));
    FL_TRY(hipMemcpyAsync(d_vreg.p, hvreg.data(), hvreg.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    FL_TRY(hipMemcpyAsync(d_cnt.p, cnt.data(), cnt.size() * sizeof(unsigned long long),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fs_count, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                       d_vis.as<double>(), n, width, bb, low, high, bins, d_u.as<double>(),
                       (int)hu.size(), d_voff.as<int>(), d_v.as<double>(), d_vreg.as<int>(),
                       d_cnt.as<unsigned long long>());
    FL_TRY(hipGetLastError());
    FL_TRY(hipMemcpyAsync(cnt.data(), d_cnt.p, cnt.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st));
    FL_TRY(hipStreamSynchronize(st));
  } else {
    high = high + (high - low) * 0.0001;
  }
  // per region: bins triples (start, end as doubles; count as int64),
  // Binning.hs:58-69
  unsigned char *o = static_cast<unsigned char *>(out);
  for (int r = 0; r < n_regions; r++)
    for (int b = 0; b < bins; b++) {
#pragma clang fp contract(off)
      const double s0 = low + (double)b * (high - low) / (double)bins;
      const double s1 = low + (double)(b + 1) * (high - low) / (double)bins;
      const long long c = (long long)cnt[(size_t)r * bins + b];
      unsigned char *slot = o + ((size_t)r * bins + b) * 24;
      memcpy(slot, &s0, 8);
      memcpy(slot + 8, &s1, 8);
      memcpy(slot + 16, &c, 8);
    }
  if (low_high) {
    low_high[0] = low;
    low_high[1] = high;
  }
  return 0;
}

extern "C" int rcg_flow_binner(const double *vis, int n, int width, const double *boxes,
                               int n_boxes, const int *bin_off, const double *bin_range,
                               const long long *bin_size, double *const *out,
                               long long *filled)
{
  clear_error();
  if (width != 5) {
    set_error(RCG_ERR_ARG, "wBinner: unexpected data width %d (5 expected)", width);
    return -1;
  }
  if (check_vis(vis, n, width, 5)) return -1;
  if (n_boxes <= 0 || !boxes || !bin_off || !out || n_boxes >= INT_MAX) {
    set_error(RCG_ERR_ARG, "need at least one output box, its bins and its output pointer");
    return -1;
  }
  const int n_bins = bin_off[n_boxes];
  if (bin_off[0] != 0 || n_bins < 0 || (n_bins > 0 && (!bin_range || !bin_size))) {
    set_error(RCG_ERR_ARG, "bad bin offsets");
    return -1;
  }
  std::vector<long long> cap((size_t)n_boxes, 0), pad((size_t)n_boxes, 0);
  for (int b = 0; b < n_boxes; b++) {
    if (bin_off[b + 1] < bin_off[b]) {
      set_error(RCG_ERR_ARG, "bin offsets of box %d decrease", b);
      return -1;
    }
    for (int k = bin_off[b]; k < bin_off[b + 1]; k++) {
      if (bin_size[k] < 0) {
        set_error(RCG_ERR_ARG, "negative bin size");
        return -1;
      }
      cap[(size_t)b] += bin_size[k];
      pad[(size_t)b] = std::max(pad[(size_t)b], bin_size[k]);
    }
    if (cap[(size_t)b] > 0 && !out[b]) {
      set_error(RCG_ERR_ARG, "null output of box %d", b);
      return -1;
    }
  }
  // w low -> v low -> u low -> (ranges, box): the first box and bin with a
  // given triple of lows wins (Map.unionsWith (Map.unionWith Map.union),
  // Binning.hs:130-131)
  std::vector<BoxEntry> ents;
  std::map<double, std::map<double, std::map<double, int>>> lut;
  for (int b = 0; b < n_boxes; b++)
    for (int k = bin_off[b]; k < bin_off[b + 1]; k++) {
      const double *g = boxes + 4 * b;
      auto &um = lut[bin_range[2 * k]][g[2]];
      if (um.count(g[0])) continue;
      um[g[0]] = (int)ents.size();
      ents.push_back({g[0], g[1], g[2], g[3], bin_range[2 * k], bin_range[2 * k + 1], b});
    }
  std::vector<double> hw, hv, hu;
  std::vector<int> hwvoff, hvuoff, huent;
  for (auto &we : lut) {
    hw.push_back(we.first);
    hwvoff.push_back((int)hv.size());
    for (auto &ve : we.second) {
      hv.push_back(ve.first);
      hvuoff.push_back((int)hu.size());
      for (auto &ue : ve.second) {
        hu.push_back(ue.first);
        huent.push_back(ue.second);
      }
    }
  }
  hwvoff.push_back((int)hv.size());
  hvuoff.push_back((int)hu.size());

  std::vector<unsigned> count((size_t)n_boxes, 0);
  std::unique_ptr<double[]> rows;
  if (n > 0 && !ents.empty()) {
    StreamGuard sg;
    FL_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    hipStream_t st = sg.s;
    DevBuf d_vis, d_w, d_wv, d_v, d_vu, d_u, d_ue, d_ent, d_key, d_val, d_key2, d_val2, d_runs,
        d_tmp, d_rows;
    FL_TRY(d_vis.alloc((size_t)n * 5 * sizeof(double)));
    FL_TRY(d_w.alloc(hw.size() * sizeof(double)));
    FL_TRY(d_wv.alloc(hwvoff.size() * sizeof(int)));
    FL_TRY(d_v.alloc(hv.size() * sizeof(double)));
    FL_TRY(d_vu.alloc(hvuoff.size() * sizeof(int)));
    FL_TRY(d_u.alloc(hu.size() * sizeof(double)));
    FL_TRY(d_ue.alloc(huent.size() * sizeof(int)));
    FL_TRY(d_ent.alloc(ents.size() * sizeof(BoxEntry)));
    FL_TRY(d_key.alloc((size_t)n * sizeof(unsigned)));
    FL_TRY(d_val.alloc((size_t)n * sizeof(unsigned)));
    FL_TRY(d_key2.alloc((size_t)n * sizeof(unsigned)));
    FL_TRY(d_val2.alloc((size_t)n * sizeof(unsigned)));
    FL_TRY(d_runs.alloc((size_t)n_boxes * 2 * sizeof(int)));
    FL_TRY(hipMemcpyAsync(d_vis.p, vis, (size_t)n * 5 * sizeof(double), hipMemcpyHostToDevice,
                          st));
    FL_TRY(hipMemcpyAsync(d_w.p, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice,
                          st));
    FL_TRY(hipMemcpyAsync(d_wv.p, hwvoff.data(), hwvoff.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    FL_TRY(hipMemcpyAsync(d_v.p, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice,
                          st));
    FL_TRY(hipMemcpyAsync(d_vu.p, hvuoff.data(), hvuoff.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    FL_TRY(hipMemcpyAsync(d_u.p, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice,
                          st));
    FL_TRY(hipMemcpyAsync(d_ue.p, huent.data(), huent.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    FL_TRY(hipMemcpyAsync(d_ent.p, ents.data(), ents.size() * sizeof(BoxEntry),
                          hipMemcpyHostToDevice, st));
    FL_TRY(hipMemsetAsync(d_runs.p, 0, (size_t)n_boxes * 2 * sizeof(int), st));
    hipLaunchKernelGGL(k_fb_target, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                       d_vis.as<double>(), n, d_w.as<double>(), (int)hw.size(), d_wv.as<int>(),
                       d_v.as<double>(), d_vu.as<int>(), d_u.as<double>(), d_ue.as<int>(),
                       d_ent.as<BoxEntry>(), n_boxes, d_key.as<unsigned>(),
                       d_val.as<unsigned>());
    FL_TRY(hipGetLastError());
    // stable sort by box: keys 0..n_boxes need end_bit bits
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= (unsigned long long)n_boxes) end_bit++;
    size_t tmp_bytes = 0;
    FL_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_key.as<unsigned>(),
                                              d_key2.as<unsigned>(), d_val.as<unsigned>(),
                                              d_val2.as<unsigned>(), n, 0, end_bit, st));
    FL_TRY(d_tmp.alloc(tmp_bytes));
    FL_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp.p, tmp_bytes, d_key.as<unsigned>(),
                                              d_key2.as<unsigned>(), d_val.as<unsigned>(),
                                              d_val2.as<unsigned>(), n, 0, end_bit, st));
    hipLaunchKernelGGL(k_fb_runs, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                       d_key2.as<unsigned>(), n, n_boxes, d_runs.as<int>(),
                       d_runs.as<int>() + n_boxes);
    FL_TRY(hipGetLastError());
    std::vector<int> runs((size_t)n_boxes * 2);
    FL_TRY(hipMemcpyAsync(runs.data(), d_runs.p, runs.size() * sizeof(int),
                          hipMemcpyDeviceToHost, st));
    FL_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < n_boxes; b++)
      count[(size_t)b] = (unsigned)(runs[(size_t)n_boxes + b] - runs[(size_t)b]);
    long long kept = 0;
    for (int b = 0; b < n_boxes; b++) {
      if ((long long)count[(size_t)b] > cap[(size_t)b]) {
        set_error(RCG_ERR_RANGE, "box %d receives %u rows but holds %lld", b,
                  count[(size_t)b], cap[(size_t)b]);
        return -1;
      }
      kept += count[(size_t)b];
    }
    if (kept > 0) {
      rows.reset(new (std::nothrow) double[(size_t)kept * 5]);
      if (!rows) {
        set_error(RCG_ERR_NOMEM, "no host memory for %lld rows", kept);
        return -1;
      }
      FL_TRY(d_rows.alloc((size_t)kept * 5 * sizeof(double)));
      hipLaunchKernelGGL(k_fb_gather, dim3(ceil_div(5 * kept, 256)), dim3(256), 0, st,
                         d_vis.as<double>(), d_val2.as<unsigned>(), (int)kept,
                         d_rows.as<double>());
      FL_TRY(hipGetLastError());
      FL_TRY(hipMemcpyAsync(rows.get(), d_rows.p, (size_t)kept * 5 * sizeof(double),
                            hipMemcpyDeviceToHost, st));
      FL_TRY(hipStreamSynchronize(st));
    }
  }
  // each box's rows, then the zero padding of Binning.hs:156-178: every bin
  // of the box zeroes rows [filled, its size)
  size_t at = 0;
  for (int b = 0; b < n_boxes; b++) {
    const size_t c = count[(size_t)b];
    if (c) memcpy(out[b], rows.get() + at * 5, c * 5 * sizeof(double));
    at += c;
    if (pad[(size_t)b] > (long long)c)
      memset(out[b] + c * 5, 0, (size_t)(pad[(size_t)b] - (long long)c) * 5 * sizeof(double));
    if (filled) filled[b] = (long long)c;
  }
  return 0;
}
