// rcg_binner.hip -- the MS4 uv binner (bin / binm) on the GPU.
//
// Restates MS4/kernel/cpu/binner/binner.cpp:100-189 (doit<OVER = 8,
// DIVIDERS = 32, mirror>), whose only caller is the GPU gather-gridder
// plugin (MS4/dna-programs/Kernel/GPU/GatherGrid.hs:20-44):
//   - every point is pregridded without translation (pregridPoint<8, mirror>
//     under __BINNER, MS4/kernel/common/common.h:96-171);
//   - its bin is (qu, qv) = (p.u / ds, p.v / ds) with ds = grid_size / 32 in
//     C's truncating division (binner.cpp:135-137), which must lie in
//     [0, 32)^2 (binner.cpp:153-155, an assert there, an error here);
//   - the record is translated by grid_size/2 - supp/2 with
//     supp = gcfsupps[w_plane] and carries supp in place of the plane
//     (binner.cpp:143-151); the amplitude is rotated by exp(2 pi i w scale)
//     in FP64 (binner.cpp:157-158);
//   - the point is copied into at most one horizontal neighbour, one
//     vertical neighbour and the diagonal between them when its remainder
//     lies within margin = supp/2 of a bin edge (binner.cpp:161-196);
//   - each bin keeps its copies in input order (one append stream per bin,
//     binner.cpp:87-90).
// The result is one pinned host allocation: a table of per-bin offsets and
// counts, then every bin's rotated amplitudes, then every bin's records
// (binner.cpp:38-85).  The function returns the address just past the
// table, as the reference does.
//
// GPU mapping: the records, rotated amplitudes and bin codes are computed
// per point; a stable counting sort places the copies.  Points are cut into
// segments of SEG; one wavefront per segment counts its copies per bin in LDS
// (counts stored bin-major, [bin][segment]); a scan per bin over its segments
// and one over the bin totals give every (bin, segment) its first slot; the
// fill pass walks each segment again in input order, 64 points at a time,
// and ranks the copies of one bin at a time with ballots, so that each bin
// receives its copies in input order.
#include <climits>
#include <cstring>
#include <vector>

#include "rcg_internal.h"

namespace rcg {
namespace {

constexpr int DIVS = 32;                     // MS4/kernel/common/metrix.h:10
constexpr int NBINS = DIVS * DIVS;
constexpr int SEG = 4096;                    // points per counting segment
// sizeof(binTable) = 2 + 3 * 32 * 32 ints, rounded up to 32 bytes
// (binner.cpp:18-28)
constexpr int TABLE_BYTES = (int)((sizeof(int) * (2 + 3 * NBINS) + 31) / 32 * 32);

struct BinTable {
  int data_size;      // whole allocation in bytes (table included)
  int data_off;       // TABLE_BYTES
  int vis_off[NBINS]; // byte offsets from the end of the table
  int pre_off[NBINS];
  int n_items[NBINS];
};
static_assert(TABLE_BYTES == 12320, "binTable size");

// Error bits of k_bn_check
constexpr int BN_REPR = 1;   // u, v or w plane outside int16
constexpr int BN_BIN = 2;    // home bin outside [0, 32)^2
constexpr int BN_INDEX = 4;  // layer index outside int16

struct BnPoint {
  int u, v, wp, idx;
  bool repr;
};

// pregridPoint<8, mirror> without translation (common.h:120-166 under
// __BINNER); every step one IEEE FP64 operation.
__device__ __forceinline__ BnPoint bn_pregrid(rcg_double3 p, double scale, double wstep,
                                              bool mirror)
{
#pragma clang fp contract(off)
  BnPoint r;
  const double us = p.u * scale, vs = p.v * scale;
  const double wr = round(p.w / wstep);
  const double fu = floor(us), fv = floor(vs);
  r.repr = fu >= -32768.0 && fu <= 32767.0 && fv >= -32768.0 && fv <= 32767.0 &&
           wr >= -32768.0 && wr <= 32767.0;
  if (!r.repr) {
    r.u = r.v = r.wp = r.idx = 0;
    return r;
  }
  r.u = (int)fu;
  r.v = (int)fv;
  r.wp = (int)wr;
  const int ou = (int)floor(8.0 * (us - fu));
  const int ov = (int)floor(8.0 * (vs - fv));
  r.idx = (mirror && r.wp < 0) ? -((-r.wp * 8 + ou) * 8 + ov) : (r.wp * 8 + ou) * 8 + ov;
  return r;
}

// Validity and w-plane range (out[0] = min, out[1] = max, preset to
// INT_MAX / INT_MIN).
__global__ void __launch_bounds__(256)
k_bn_check(const rcg_double3 *__restrict__ uvw, int n, double scale, double wstep, int mirror,
           int ds, int *__restrict__ range, int *__restrict__ err)
{
  int lo = INT_MAX, hi = INT_MIN, e = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const BnPoint p = bn_pregrid(uvw[i], scale, wstep, mirror != 0);
    if (!p.repr) {
      e |= BN_REPR;
      continue;
    }
    const int qu = p.u / ds, qv = p.v / ds;
    if (qu < 0 || qv < 0 || qu >= DIVS || qv >= DIVS) e |= BN_BIN;
    if (p.idx < -32768 || p.idx > 32767) e |= BN_INDEX;
    lo = min(lo, p.wp);
    hi = max(hi, p.wp);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
    e |= __shfl_xor(e, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (lo != INT_MAX) atomicMin(&range[0], lo);
    if (hi != INT_MIN) atomicMax(&range[1], hi);
    if (e) atomicOr(err, e);
  }
}

// Per point: the translated record (u, v, layer index, support as int16),
// the rotated amplitude and the bin code home | (du + 1) << 10 | (dv + 1) << 12.
// supp[w - pmin] = gcfsupps[w].
__global__ void __launch_bounds__(256)
k_bn_point(const rcg_double3 *__restrict__ uvw, const double2 *__restrict__ amp, int n,
           double scale, double wstep, int mirror, int grid_size, const int *__restrict__ supp,
           int pmin, short4 *__restrict__ rec, double2 *__restrict__ rot, int *__restrict__ code)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const rcg_double3 x = uvw[i];
  const BnPoint p = bn_pregrid(x, scale, wstep, mirror != 0);
  const int ds = grid_size / DIVS;
  const int qu = p.u / ds, ru = p.u % ds;
  const int qv = p.v / ds, rv = p.v % ds;
  const int s = supp[p.wp - pmin];
  const short t = (short)(grid_size / 2 - s / 2);
  rec[i] = make_short4((short)(p.u + t), (short)(p.v + t), (short)p.idx, (short)s);
  const double ph = 2.0 * M_PI * (x.w * scale);
  double sn, cs;
  sincos(ph, &sn, &cs);
  const double2 a = amp[i];
  rot[i] = make_double2(a.x * cs - a.y * sn, a.x * sn + a.y * cs);
  int du = 0, dv = 0;
  const int margin = s / 2;
  if (margin > 0) {
    if (ru < margin && qu >= 1) du = -1;
    else if (ru > ds - margin && qu < DIVS - 1) du = 1;
    if (rv < margin && qv >= 1) dv = -1;
    else if (rv > ds - margin && qv < DIVS - 1) dv = 1;
  }
  code[i] = (qu * DIVS + qv) | (du + 1) << 10 | (dv + 1) << 12;
}

__device__ __forceinline__ void bn_bins(int c, int b[4])
{
  const int home = c & (NBINS - 1);
  const int du = ((c >> 10) & 3) - 1, dv = ((c >> 12) & 3) - 1;
  b[0] = home;
  b[1] = du ? home + du * DIVS : -1;
  b[2] = dv ? home + dv : -1;
  b[3] = (du && dv) ? home + du * DIVS + dv : -1;
}

// One wavefront per segment: counts[bin * nseg + seg].
__global__ void __launch_bounds__(64)
k_bn_count(const int *__restrict__ code, int n, int nseg, int *__restrict__ counts)
{
  __shared__ int hist[NBINS];
  const int lane = threadIdx.x, seg = blockIdx.x;
  for (int b = lane; b < NBINS; b += 64) hist[b] = 0;
  __syncthreads();
  const int end = min(n, (seg + 1) * SEG);
  for (int i = seg * SEG + lane; i < end; i += 64) {
    int b[4];
    bn_bins(code[i], b);
    for (int k = 0; k < 4; k++)
      if (b[k] >= 0) atomicAdd(&hist[b[k]], 1);
  }
  __syncthreads();
  for (int b = lane; b < NBINS; b += 64) counts[(long long)b * nseg + seg] = hist[b];
}

// Exclusive scan of v over a 256-thread workgroup; *tot = the sum.
__device__ __forceinline__ int bn_block_scan(int v, int *sh, int *tot)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int a = __shfl_up(x, off);
    if (lane >= off) x += a;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wave; w++) base += sh[w];
  *tot = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return base + x - v;
}

// One workgroup per bin: offs[bin][seg] = first slot of the segment's copies
// within the bin, totals[bin] = the bin's copies.
__global__ void __launch_bounds__(256)
k_bn_scan_bins(const int *__restrict__ counts, int nseg, int *__restrict__ offs,
               int *__restrict__ totals)
{
  __shared__ int sh[4];
  const long long row = (long long)blockIdx.x * nseg;
  int carry = 0;
  for (int s0 = 0; s0 < nseg; s0 += 256) {
    const int s = s0 + threadIdx.x;
    const int v = s < nseg ? counts[row + s] : 0;
    int tot;
    const int e = bn_block_scan(v, sh, &tot);
    if (s < nseg) offs[row + s] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// One workgroup of 1024 threads: start[bin] = exclusive scan of the bin
// totals, start[1024] = all copies.
__global__ void __launch_bounds__(NBINS)
k_bn_scan_top(const int *__restrict__ totals, int *__restrict__ start)
{
  __shared__ int sh[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int v = totals[threadIdx.x];
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int a = __shfl_up(x, off);
    if (lane >= off) x += a;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wave; w++) base += sh[w];
  start[threadIdx.x] = base + x - v;
  if (threadIdx.x == NBINS - 1) start[NBINS] = base + x;
}

// One wavefront per segment, 64 points per step in input order.  Per step
// the copies still to place are taken one bin at a time (the bin of the
// first lane with a copy left): the lanes holding a copy of that bin are
// ranked by lane, i.e. by input order, so each bin receives its copies in
// input order.  A point has at most one copy per bin.
__global__ void __launch_bounds__(64)
k_bn_fill(const int *__restrict__ code, const short4 *__restrict__ rec,
          const double2 *__restrict__ rot, int n, int nseg, const int *__restrict__ offs,
          const int *__restrict__ start, double2 *__restrict__ vis_out,
          short4 *__restrict__ pre_out)
{
  __shared__ int cur[NBINS];
  const int lane = threadIdx.x, seg = blockIdx.x;
  for (int b = lane; b < NBINS; b += 64) cur[b] = start[b] + offs[(long long)b * nseg + seg];
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1;
  const int end = min(n, (seg + 1) * SEG);
  for (int i0 = seg * SEG; i0 < end; i0 += 64) {
    const int i = i0 + lane;
    int b[4] = {-1, -1, -1, -1};
    short4 r = make_short4(0, 0, 0, 0);
    double2 a = make_double2(0.0, 0.0);
    if (i < end) {
      bn_bins(code[i], b);
      r = rec[i];
      a = rot[i];
    }
    int pend = (b[0] >= 0) | (b[1] >= 0) << 1 | (b[2] >= 0) << 2 | (b[3] >= 0) << 3;
    for (;;) {
      const unsigned long long act = __ballot(pend != 0);
      if (!act) break;
      const int leader = __ffsll((long long)act) - 1;
      const int mine = (pend & 1) ? b[0] : (pend & 2) ? b[1] : (pend & 4) ? b[2] : b[3];
      const int B = __shfl(mine, leader);
      int k = -1;
      for (int q = 3; q >= 0; q--)
        if (((pend >> q) & 1) && b[q] == B) k = q;
      const unsigned long long m = __ballot(k >= 0);
      const int base = cur[B];
      if (k >= 0) {
        const int pos = base + __popcll(m & below);
        vis_out[pos] = a;
        pre_out[pos] = r;
        pend &= ~(1 << k);
      }
      __builtin_amdgcn_wave_barrier();
      if (lane == leader) cur[B] = base + __popcll(m);
      __builtin_amdgcn_wave_barrier();
    }
  }
}

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

#define BN_TRY(expr)                                                          \
  do {                                                                        \
    hipError_t e__ = (expr);                                                  \
    if (e__ != hipSuccess) {                                                  \
      set_error(RCG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e__)); \
      return nullptr;                                                         \
    }                                                                         \
  } while (0)

// A pinned host block of `bytes` with the table at its start; returns the
// address past the table, or null.
void *alloc_result(size_t bytes, char **base)
{
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
    set_error(RCG_ERR_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
    return nullptr;
  }
  *base = static_cast<char *>(p);
  return *base + TABLE_BYTES;
}

void *bin_call(int n, double scale, double wstep, const rcg_c128 *amps,
               const rcg_double3 *uvws, const int *gcfsupps, int grid_size, bool mirror)
{
  clear_error();
  if (n < 0 || n > (1 << 29)) {
    set_error(RCG_ERR_ARG, "num_points %d outside [0, 2^29]", n);
    return nullptr;
  }
  if (grid_size < DIVS || grid_size > 32767) {
    set_error(RCG_ERR_ARG, "grid_size %d outside [32, 32767] (grid_size / 32 cells per bin)",
              grid_size);
    return nullptr;
  }
  if (n > 0 && (!amps || !uvws || !gcfsupps)) {
    set_error(RCG_ERR_ARG, "null amplitude, uvw or support table");
    return nullptr;
  }
  const int ds = grid_size / DIVS;
  if (n == 0) {
    char *base = nullptr;
    void *data = alloc_result(TABLE_BYTES, &base);
    if (!data) return nullptr;
    BinTable *t = reinterpret_cast<BinTable *>(base);
    memset(base, 0, TABLE_BYTES);
    t->data_size = TABLE_BYTES;
    t->data_off = TABLE_BYTES;
    return data;
  }

  StreamGuard sg;
  BN_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  hipStream_t st = sg.s;
  DevBuf d_uvw, d_amp, d_range, d_err;
  BN_TRY(d_uvw.alloc((size_t)n * sizeof(rcg_double3)));
  BN_TRY(d_amp.alloc((size_t)n * sizeof(rcg_c128)));
  BN_TRY(d_range.alloc(3 * sizeof(int)));
  BN_TRY(hipMemcpyAsync(d_uvw.p, uvws, (size_t)n * sizeof(rcg_double3),
                        hipMemcpyHostToDevice, st));
  BN_TRY(hipMemcpyAsync(d_amp.p, amps, (size_t)n * sizeof(rcg_c128), hipMemcpyHostToDevice,
                        st));
  int hr[3] = {INT_MAX, INT_MIN, 0};
  BN_TRY(hipMemcpyAsync(d_range.p, hr, sizeof(hr), hipMemcpyHostToDevice, st));
  const int nb = std::min(ceil_div(n, 256), 2048);
  hipLaunchKernelGGL(k_bn_check, dim3(nb), dim3(256), 0, st, d_uvw.as<rcg_double3>(), n,
                     scale, wstep, mirror ? 1 : 0, ds, d_range.as<int>(),
                     d_range.as<int>() + 2);
  BN_TRY(hipGetLastError());
  BN_TRY(hipMemcpyAsync(hr, d_range.p, sizeof(hr), hipMemcpyDeviceToHost, st));
  BN_TRY(hipStreamSynchronize(st));
  if (hr[2] & BN_REPR) {
    set_error(RCG_ERR_RANGE, "a point's u*scale, v*scale or w/wstep leaves the int16 range");
    return nullptr;
  }
  if (hr[2] & BN_BIN) {
    set_error(RCG_ERR_RANGE,
              "a point's untranslated cell lies outside bins [0, 32)^2 of %d cells (the "
              "reference asserts)", ds);
    return nullptr;
  }
  if (hr[2] & BN_INDEX) {
    set_error(RCG_ERR_RANGE, "a point's GCF layer index leaves the int16 range");
    return nullptr;
  }

  // supports of the planes in use, read through the caller's centred table
  const int pmin = hr[0], pmax = hr[1];
  std::vector<int> hs((size_t)(pmax - pmin + 1));
  for (int p = pmin; p <= pmax; p++) hs[(size_t)(p - pmin)] = gcfsupps[p];
  const int nseg = ceil_div(n, SEG);
  DevBuf d_supp, d_rec, d_rot, d_code, d_counts, d_offs, d_tot, d_start;
  BN_TRY(d_supp.alloc(hs.size() * sizeof(int)));
  BN_TRY(d_rec.alloc((size_t)n * sizeof(short4)));
  BN_TRY(d_rot.alloc((size_t)n * sizeof(double2)));
  BN_TRY(d_code.alloc((size_t)n * sizeof(int)));
  BN_TRY(d_counts.alloc((size_t)NBINS * nseg * sizeof(int)));
  BN_TRY(d_offs.alloc((size_t)NBINS * nseg * sizeof(int)));
  BN_TRY(d_tot.alloc(NBINS * sizeof(int)));
  BN_TRY(d_start.alloc((NBINS + 1) * sizeof(int)));
  BN_TRY(hipMemcpyAsync(d_supp.p, hs.data(), hs.size() * sizeof(int), hipMemcpyHostToDevice,
                        st));
  hipLaunchKernelGGL(k_bn_point, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                     d_uvw.as<rcg_double3>(), d_amp.as<double2>(), n, scale, wstep,
                     mirror ? 1 : 0, grid_size, d_supp.as<int>(), pmin, d_rec.as<short4>(),
                     d_rot.as<double2>(), d_code.as<int>());
  BN_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bn_count, dim3(nseg), dim3(64), 0, st, d_code.as<int>(), n, nseg,
                     d_counts.as<int>());
  BN_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bn_scan_bins, dim3(NBINS), dim3(256), 0, st, d_counts.as<int>(), nseg,
                     d_offs.as<int>(), d_tot.as<int>());
  BN_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bn_scan_top, dim3(1), dim3(NBINS), 0, st, d_tot.as<int>(),
                     d_start.as<int>());
  BN_TRY(hipGetLastError());
  std::vector<int> start(NBINS + 1);
  BN_TRY(hipMemcpyAsync(start.data(), d_start.p, start.size() * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  BN_TRY(hipStreamSynchronize(st));
  const long long total = start[NBINS];
  const long long bytes = TABLE_BYTES + total * (long long)(sizeof(rcg_c128) + 8);
  if (bytes > INT_MAX) {
    set_error(RCG_ERR_RANGE, "%lld copies need %lld bytes; the table's offsets are int",
              total, bytes);
    return nullptr;
  }
  const long long vis_bytes = total * (long long)sizeof(rcg_c128);
  DevBuf d_stage;
  BN_TRY(d_stage.alloc((size_t)(bytes - TABLE_BYTES)));
  char *stage = d_stage.as<char>();
  hipLaunchKernelGGL(k_bn_fill, dim3(nseg), dim3(64), 0, st, d_code.as<int>(),
                     d_rec.as<short4>(), d_rot.as<double2>(), n, nseg, d_offs.as<int>(),
                     d_start.as<int>(), reinterpret_cast<double2 *>(stage),
                     reinterpret_cast<short4 *>(stage + vis_bytes));
  BN_TRY(hipGetLastError());

  char *base = nullptr;
  void *data = alloc_result((size_t)bytes, &base);
  if (!data) return nullptr;
  BinTable *t = reinterpret_cast<BinTable *>(base);
  memset(base, 0, TABLE_BYTES);
  t->data_size = (int)bytes;
  t->data_off = TABLE_BYTES;
  for (int b = 0; b < NBINS; b++) {
    // empty bins: the reference leaves the offsets unset; here they point
    // where the bin would start
    t->n_items[b] = start[b + 1] - start[b];
    t->vis_off[b] = (int)(start[b] * (long long)sizeof(rcg_c128));
    t->pre_off[b] = (int)(vis_bytes + start[b] * 8LL);
  }
  hipError_t e = hipMemcpyAsync(data, stage, (size_t)(bytes - TABLE_BYTES),
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_error(RCG_ERR_HIP, "copy of the binned data failed: %s", hipGetErrorString(e));
    (void)hipHostFree(base);
    return nullptr;
  }
  return data;
}

}  // namespace
}  // namespace rcg

extern "C" void *bin(int num_points, double scale, double wstep, rcg_c128 *amps,
                     rcg_double3 *uvws, int gcfsupps[], int grid_size)
{
  return rcg::bin_call(num_points, scale, wstep, amps, uvws, gcfsupps, grid_size, false);
}

extern "C" void *binm(int num_points, double scale, double wstep, rcg_c128 *amps,
                      rcg_double3 *uvws, int gcfsupps[], int grid_size)
{
  return rcg::bin_call(num_points, scale, wstep, amps, uvws, gcfsupps, grid_size, true);
}

extern "C" void rcg_bin_free(void *data)
{
  if (data) (void)hipHostFree(static_cast<char *>(data) - rcg::TABLE_BYTES);
}
