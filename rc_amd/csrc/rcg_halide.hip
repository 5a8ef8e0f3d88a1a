// rcg_halide.hip -- the MS6 kernel entry points (Halide buffer_t ABI) on the
// GPU: kern_scatter / kern_degrid and their _gpu1 variants.
//
// Semantics restated from the Halide generators
// MS6/kernel/cpu/gridding/scatter.cpp:17-145 and degrid.cpp:17-124 (one
// oversampled GCF, round-based pixel mapping, u on the fast grid axis,
// grid added to, degrid subtracting in place), with the dispatch and
// return codes of scatter1.cpp:9-19, degrid1.cpp:9-18,
// MS6/kernel/gpu/gridding/kern_scatter_gpu1.cpp:22-40 and
// kern_degrid_gpu1.cpp:21-41.  FP64 throughout, as the reference.
//
// GPU structure: one wavefront per point, lanes over the S*S window
// (row-major in (y, x)); gridding adds with FP64 atomics, degridding reduces
// across the wavefront and subtracts.
//
// nvGridder (MS6/kernel/nvidia/gridder/halide_api.cu:4-29) is the same
// scatter with that gridder's own conventions (see nv_scatter below).
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "rcg_internal.h"
#include "../../include/rcgrid_halide.h"

namespace rcg {
namespace {

constexpr int H_OVER = 8;

struct HB {
  int ext[4], str[4], mn[4];
};

HB hb_of(const buffer_t *b)
{
  HB h;
  for (int i = 0; i < 4; i++) {
    h.ext[i] = b->extent[i];
    h.str[i] = b->stride[i];
    h.mn[i] = b->min[i];
  }
  return h;
}

// Elements spanned by a buffer (strides are positive here).
long long span_of(const buffer_t *b)
{
  long long s = 1;
  for (int i = 0; i < 4; i++)
    if (b->extent[i] > 0) s += (long long)(b->extent[i] - 1) * b->stride[i];
  return s;
}

struct Pix {
  int x0, y0, over;
  bool ok;
};

// scatter.cpp:70-79 / degrid.cpp:59-67: pixel by round, oversampling offset
// by round of the fraction above floor, clamped to [0, 7]; Halide's round
// takes halves to even (rint).
__device__ __forceinline__ Pix h_pixel(double u, double v, double scale, int grid_size, int S,
                                       const HB &g, int margin)
{
#pragma clang fp contract(off)
  Pix p;
  const double us = u * scale, vs = v * scale;
  const int ou = min(max((int)rint(H_OVER * (us - floor(us))), 0), H_OVER - 1);
  const int ov = min(max((int)rint(H_OVER * (vs - floor(vs))), 0), H_OVER - 1);
  p.x0 = (int)(rint(us) + (double)(grid_size / 2 - S / 2));
  p.y0 = (int)(rint(vs) + (double)(grid_size / 2 - S / 2));
  p.over = ou + H_OVER * ov;
  const int m = max(0, (margin - S) / 2);
  p.ok = p.x0 >= g.mn[1] + m && p.x0 <= g.mn[1] + g.ext[1] - S - 1 - m &&
         p.y0 >= g.mn[2] + m && p.y0 <= g.mn[2] + g.ext[2] - S - 1 - m;
  return p;
}

__global__ void __launch_bounds__(256)
k_h_scatter(const double *__restrict__ vis, HB vb, const double *__restrict__ gcf, HB gb,
            double *__restrict__ uvg, HB ub, double scale, int grid_size, int margin, int S)
{
  const long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= vb.ext[1]) return;
  const double *r = vis + t * vb.str[1];
  const double u = r[0], v = r[vb.str[0]], re = r[3 * vb.str[0]], im = r[4 * vb.str[0]];
  const Pix p = h_pixel(u, v, scale, grid_size, S, ub, margin);
  if (!p.ok) return;
  for (int k = lane; k < S * S; k += 64) {
    const int sy = k / S, sx = k - (k / S) * S;
    const double *g = gcf + (long long)sx * gb.str[1] + (long long)sy * gb.str[2] +
                      (long long)p.over * gb.str[3];
    const double gr = g[0], gi = g[gb.str[0]];
    const double pr = re * gr - im * gi, pi = re * gi + im * gr;
    double *c = uvg + (long long)(p.x0 + sx - ub.mn[1]) * ub.str[1] +
                (long long)(p.y0 + sy - ub.mn[2]) * ub.str[2];
    unsafeAtomicAdd(c, pr);
    unsafeAtomicAdd(c + ub.str[0], pi);
  }
}

__global__ void __launch_bounds__(256)
k_h_degrid(const double *__restrict__ gcf, HB gb, const double *__restrict__ uvg, HB ub,
           const double *__restrict__ vis, HB vb, double *__restrict__ out, HB ob,
           double scale, int grid_size, int margin, int S)
{
  const long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= vb.ext[1]) return;
  const double *r = vis + t * vb.str[1];
  double *o = out + t * ob.str[1];
  const Pix p = h_pixel(r[0], r[vb.str[0]], scale, grid_size, S, ub, margin);
  double sr = 0.0, si = 0.0;
  if (p.ok) {
    for (int k = lane; k < S * S; k += 64) {
      const int sy = k / S, sx = k - (k / S) * S;
      const double *g = gcf + (long long)sx * gb.str[1] + (long long)sy * gb.str[2] +
                        (long long)p.over * gb.str[3];
      const double *c = uvg + (long long)(p.x0 + sx - ub.mn[1]) * ub.str[1] +
                        (long long)(p.y0 + sy - ub.mn[2]) * ub.str[2];
      const double gr = g[0], gi = g[gb.str[0]];
      const double cr = c[0], ci = c[ub.str[0]];
      sr += cr * gr - ci * gi;
      si += cr * gi + ci * gr;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_xor(sr, off);
      si += __shfl_xor(si, off);
    }
  }
  if (lane < 5) {
    double x = r[lane * vb.str[0]];
    if (lane == 3) x -= sr;
    if (lane == 4) x -= si;
    o[lane * ob.str[0]] = x;
  }
}

// Support of a GCF buffer (gcf_common.h:8-14): extent {2, S, S, ...}.
int gcf_size(const buffer_t *g)
{
  const int s = g->extent[1];
  return (g->extent[0] == 2 && g->extent[2] == s) ? s : -1;
}

bool check_vis(const buffer_t *b, const char *who)
{
  if (!b || b->elem_size != 8 || b->extent[0] != 5 || b->stride[0] < 1 || b->stride[1] < 5 ||
      b->extent[1] < 0 || (b->extent[1] > 0 && !b->host && !b->dev)) {
    set_error(RCG_ERR_ARG, "%s: vis must be doubles of extent {5, n}", who);
    return false;
  }
  return true;
}

bool check_gcf(const buffer_t *b, int S, const char *who)
{
  if (!b || b->elem_size != 8 || b->extent[3] < 64 || b->stride[0] < 1 ||
      b->stride[1] < 2 * b->stride[0] || b->stride[2] < S * b->stride[1] ||
      b->stride[3] < S * b->stride[2] || (!b->host && !b->dev)) {
    set_error(RCG_ERR_ARG, "%s: gcf must be doubles of extent {2, S, S, 64}", who);
    return false;
  }
  return true;
}

bool check_uvg(const buffer_t *b, const char *who)
{
  if (!b || b->elem_size != 8 || b->extent[0] != 2 || b->stride[0] < 1 ||
      b->stride[1] < 2 * b->stride[0] || b->stride[2] < b->extent[1] * b->stride[1] ||
      b->extent[1] < 1 || b->extent[2] < 1 || (!b->host && !b->dev)) {
    set_error(RCG_ERR_ARG, "%s: uvg must be doubles of extent {2, X, Y}", who);
    return false;
  }
  return true;
}

// A buffer's data on the device: the caller's device pointer (gpu mode,
// dev set, host copy not newer) or a copy of the host data.
struct DevData {
  double *p = nullptr;
  bool own = false;
};

hipError_t to_device(const buffer_t *b, bool gpu_mode, bool upload, hipStream_t st, DevData *d)
{
  if (gpu_mode && b->dev && !b->host_dirty) {
    d->p = reinterpret_cast<double *>(b->dev);
    d->own = false;
    return hipSuccess;
  }
  const size_t bytes = (size_t)span_of(b) * sizeof(double);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->p), bytes);
  if (e != hipSuccess) {
    d->p = nullptr;
    return e;
  }
  d->own = true;
  if (upload && b->host) e = hipMemcpyAsync(d->p, b->host, bytes, hipMemcpyHostToDevice, st);
  return e;
}

void release(DevData *ds, int n)
{
  for (int i = 0; i < n; i++)
    if (ds[i].own && ds[i].p) (void)hipFree(ds[i].p);
}

// Halide's device_free after the reference's _gpu1 calls: no device copy
// is left behind (caller-owned device memory is not freed here).
void drop_device(buffer_t *b)
{
  b->dev = 0;
  b->dev_dirty = false;
  b->host_dirty = false;
}

int scatter(double scale, int grid_size, int margin, buffer_t *vis, buffer_t *gcf, buffer_t *uvg,
            bool gpu_mode, int bad_size)
{
  clear_error();
  const char *who = gpu_mode ? "kern_scatter_gpu1" : "kern_scatter";
  const int S = gcf ? gcf_size(gcf) : -1;
  if (S != 8 && S != 16 && S != 32 && S != 64) return bad_size;
  if (!check_vis(vis, who) || !check_gcf(gcf, S, who) || !check_uvg(uvg, who)) return -1;
  if (!uvg->host) {
    set_error(RCG_ERR_ARG, "%s: uvg needs host memory for the result", who);
    return -1;
  }
  hipStream_t st = nullptr;
  DevData d[3];
  int rc = 0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      to_device(vis, gpu_mode, true, st, &d[0]) != hipSuccess ||
      to_device(gcf, gpu_mode, true, st, &d[1]) != hipSuccess ||
      to_device(uvg, gpu_mode, true, st, &d[2]) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "%s: device buffers", who);
    rc = -1;
  } else {
    const long long n = vis->extent[1];
    if (n > 0) {
      const long long blocks = (n * 64 + 255) / 256;
      hipLaunchKernelGGL(k_h_scatter, dim3((unsigned)blocks), dim3(256), 0, st, d[0].p,
                         hb_of(vis), d[1].p, hb_of(gcf), d[2].p, hb_of(uvg), scale, grid_size,
                         margin, S);
    }
    const size_t bytes = (size_t)span_of(uvg) * sizeof(double);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(uvg->host, d[2].p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error(RCG_ERR_HIP, "%s: kernel or copy failed", who);
      rc = -1;
    }
  }
  if (st) (void)hipStreamSynchronize(st);
  release(d, 3);
  if (st) (void)hipStreamDestroy(st);
  if (gpu_mode) {
    drop_device(vis);
    drop_device(gcf);
    drop_device(uvg);
  }
  return rc;
}

int degrid(double scale, int grid_size, int margin, buffer_t *gcf, buffer_t *uvg, buffer_t *vis,
           buffer_t *vis_out, bool gpu_mode, int bad_size)
{
  clear_error();
  const char *who = gpu_mode ? "kern_degrid_gpu1" : "kern_degrid";
  const int S = gcf ? gcf_size(gcf) : -1;
  if (S != 8 && S != 16 && S != 32 && S != 64) return bad_size;
  if (!check_vis(vis, who) || !check_vis(vis_out, who) || !check_gcf(gcf, S, who) ||
      !check_uvg(uvg, who))
    return -1;
  if (vis_out->extent[1] != vis->extent[1] || !vis_out->host) {
    set_error(RCG_ERR_ARG, "%s: vis_out must match vis and have host memory", who);
    return -1;
  }
  hipStream_t st = nullptr;
  DevData d[4];
  int rc = 0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      to_device(gcf, gpu_mode, true, st, &d[0]) != hipSuccess ||
      to_device(uvg, gpu_mode, true, st, &d[1]) != hipSuccess ||
      to_device(vis, gpu_mode, true, st, &d[2]) != hipSuccess ||
      to_device(vis_out, false, false, st, &d[3]) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "%s: device buffers", who);
    rc = -1;
  } else {
    const long long n = vis->extent[1];
    if (n > 0) {
      const long long blocks = (n * 64 + 255) / 256;
      hipLaunchKernelGGL(k_h_degrid, dim3((unsigned)blocks), dim3(256), 0, st, d[0].p,
                         hb_of(gcf), d[1].p, hb_of(uvg), d[2].p, hb_of(vis), d[3].p,
                         hb_of(vis_out), scale, grid_size, margin, S);
    }
    const size_t bytes = (size_t)span_of(vis_out) * sizeof(double);
    if (hipGetLastError() != hipSuccess ||
        (n > 0 && hipMemcpyAsync(vis_out->host, d[3].p, bytes, hipMemcpyDeviceToHost, st) !=
                      hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error(RCG_ERR_HIP, "%s: kernel or copy failed", who);
      rc = -1;
    }
  }
  if (st) (void)hipStreamSynchronize(st);
  release(d, 4);
  if (st) (void)hipStreamDestroy(st);
  if (gpu_mode) {
    drop_device(gcf);
    drop_device(uvg);
    drop_device(vis);
    drop_device(vis_out);
  }
  return rc;
}

// ---------------------------------------------------------------------------
// nvGridder: MS6/kernel/nvidia/gridder (grid_gpu.cu:37-51 vis2ints,
// 83-155 grid_kernel_gather, 159-243 host side; binsort.cpp:6-32).
//   vis   5 doubles per point (x, y, z, re, im), read from `host` as a
//         contiguous array of extent[1] points, and sorted there in place;
//   gcf   one 16 x 16 GCF per oversampling offset, [8 sub_y + sub_x][b][a]
//         complex doubles, read from `host`;
//   uvg   grid_size x grid_size complex doubles, x (from u) the fast axis,
//         written (not added to) from `host` on;
//   pixel xs = x * scale, main_x = floor(xs), sub_x = floor(8 (xs - main_x)),
//         likewise y; no shift to the grid centre;
//   taps  pixel (main_x + a - 7, main_y + b - 7) += vis * gcf[over][b][a] for
//         a, b in [0, 16), where it lies inside the grid.
// The reference gathers per pixel over visibilities found through
// bookmarks of its sorted order; for non-negative pixel coordinates and
// consistent sort and bookmark blocks (scale 1, grids up to 2048) that is
// this sum.  For negative coordinates it indexes the GCF with a negative
// sub-pixel offset (C division and remainder truncate toward zero) and
// reads outside the table; here floor is used throughout.

constexpr int NV_GCF = 16;
constexpr int NV_TAPS = NV_GCF * NV_GCF;

__global__ void k_nv_zero(double *__restrict__ p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// pos: x, y of point t at pos[t * ps], pos[t * ps + 1]; val: re, im at
// val[t * vs], val[t * vs + 1]
__global__ void __launch_bounds__(256)
k_nv_scatter(const double *__restrict__ pos, int ps, const double *__restrict__ val, int vs,
             long long n, const double2 *__restrict__ gcf, double *__restrict__ out,
             double scale, int img)
{
#pragma clang fp contract(off)
  const long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= n) return;
  const double xs = pos[t * ps] * scale, ys = pos[t * ps + 1] * scale;
  const double fx = floor(xs), fy = floor(ys);
  // no tap of a point outside these bounds lands in the grid (also NaN)
  if (!(fx >= -8.0 && fx <= img + 6.0 && fy >= -8.0 && fy <= img + 6.0)) return;
  const int mx = (int)fx, my = (int)fy;
  const int sx = (int)floor(H_OVER * (xs - fx)), sy = (int)floor(H_OVER * (ys - fy));
  const double2 *g = gcf + (H_OVER * sy + sx) * NV_TAPS;
  const double vr = val[t * vs], vi = val[t * vs + 1];
  for (int k = lane; k < NV_TAPS; k += 64) {
    const int px = mx + (k & (NV_GCF - 1)) - 7, py = my + k / NV_GCF - 7;
    if (px < 0 || px >= img || py < 0 || py >= img) continue;
    const double2 c = g[k];
    double *o = out + 2 * ((long long)py * img + px);
    unsafeAtomicAdd(o, vr * c.x - vi * c.y);
    unsafeAtomicAdd(o + 1, vr * c.y + c.x * vi);
  }
}

// The reference's in-place host sort (binsort.cpp:6-32): by 8-pixel block
// of the unscaled coordinates, block key (gy * 2055) / 8 + gx in C integer
// arithmetic (gx = floor(x) / 8 truncated, IMG_SIZE 2048), then by
// 8 (x - floor x) + (y - floor y).  Its comparator is not a strict weak
// order on equal keys (std::sort leaves their order unspecified); equal
// keys keep their input order here.
struct NvKey {
  long long blk;
  double sub;
};

int sat_int(double v)
{
  if (!(v == v)) return 0;
  return v < (double)INT_MIN ? INT_MIN : v > (double)INT_MAX ? INT_MAX : (int)v;
}

NvKey nv_key(const double *rec)
{
  const int bs = NV_GCF / 2;
  const int mx = sat_int(std::floor(rec[0])), my = sat_int(std::floor(rec[1]));
  const long long gx = mx / bs, gy = my / bs;
  NvKey k;
  k.blk = gy * (2048 + bs - 1) / bs + gx;
  k.sub = H_OVER * (rec[0] - mx) + (rec[1] - my);
  return k;
}

// records of `rec` doubles, (x, y) first
void nv_sort(double *recs, long long n, int rec)
{
  if (n < 2) return;
  std::vector<long long> idx(n);
  std::vector<NvKey> key(n);
  for (long long i = 0; i < n; i++) {
    idx[i] = i;
    key[i] = nv_key(recs + rec * i);
  }
  std::stable_sort(idx.begin(), idx.end(), [&](long long a, long long b) {
    return key[a].blk < key[b].blk || (key[a].blk == key[b].blk && key[a].sub < key[b].sub);
  });
  std::vector<double> tmp(recs, recs + rec * n);
  for (long long i = 0; i < n; i++)
    for (int c = 0; c < rec; c++) recs[rec * i + c] = tmp[rec * idx[i] + c];
}

// Grids n points into out (img x img complex doubles, host, overwritten).
// pos / val: host arrays of n records of ps / vs doubles, (x, y) and
// (re, im) first.  shared: 5-double records holding both, val unused.
int nv_grid_host(const char *who, double scale, double *out, const double *pos, int ps,
                 const double *val, int vs, bool shared, long long n, int img,
                 const double *gcf)
{
  const long long cells = (long long)img * img;
  hipStream_t st = nullptr;
  double *dp = nullptr, *dv = nullptr, *dout = nullptr;
  double2 *dg = nullptr;
  const size_t pbytes = (size_t)std::max(n, 1LL) * ps * sizeof(double);
  const size_t vbytes = (size_t)std::max(n, 1LL) * vs * sizeof(double);
  const size_t gbytes = sizeof(double2) * H_OVER * H_OVER * NV_TAPS;
  const size_t obytes = sizeof(double) * 2 * (size_t)cells;
  int rc = 0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&dp), pbytes) != hipSuccess ||
      (!shared && hipMalloc(reinterpret_cast<void **>(&dv), vbytes) != hipSuccess) ||
      hipMalloc(reinterpret_cast<void **>(&dg), gbytes) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&dout), obytes) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "%s: device buffers", who);
    rc = -1;
  } else {
    hipError_t e = hipMemcpyAsync(dg, gcf, gbytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n > 0)
      e = hipMemcpyAsync(dp, pos, (size_t)n * ps * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n > 0 && !shared)
      e = hipMemcpyAsync(dv, val, (size_t)n * vs * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      const long long nd = 2 * cells;
      const int zb = (int)std::min<long long>((nd + 255) / 256, 65536);
      hipLaunchKernelGGL(k_nv_zero, dim3(zb), dim3(256), 0, st, dout, nd);
      e = hipGetLastError();
    }
    if (e == hipSuccess && n > 0) {
      const long long blocks = (n * 64 + 255) / 256;
      const double *vals = shared ? dp + 3 : dv;
      hipLaunchKernelGGL(k_nv_scatter, dim3((unsigned)blocks), dim3(256), 0, st, dp, ps, vals,
                         vs, n, dg, dout, scale, img);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, obytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error(RCG_ERR_HIP, "%s: kernel or copy failed", who);
      rc = -1;
    }
  }
  if (st) (void)hipStreamSynchronize(st);
  if (dp) (void)hipFree(dp);
  if (dv) (void)hipFree(dv);
  if (dg) (void)hipFree(dg);
  if (dout) (void)hipFree(dout);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

int nv_gridder(double scale, int grid_size, buffer_t *vis, buffer_t *gcf, buffer_t *uvg)
{
  clear_error();
  const char *who = "nvGridder";
  if (!vis || !gcf || !uvg || grid_size < 1) {
    set_error(RCG_ERR_ARG, "%s: null buffer or grid size < 1", who);
    return -1;
  }
  const long long n = vis->extent[1];
  const long long cells = (long long)grid_size * grid_size;
  if (vis->elem_size != 8 || n < 0 || (n > 0 && !vis->host) || vis->extent[0] != 5 ||
      vis->stride[0] != 1 || (n > 1 && vis->stride[1] != 5)) {
    set_error(RCG_ERR_ARG, "%s: vis must be contiguous doubles of extent {5, n}", who);
    return -1;
  }
  if (gcf->elem_size != 8 || !gcf->host || span_of(gcf) < 2LL * H_OVER * H_OVER * NV_TAPS) {
    set_error(RCG_ERR_ARG, "%s: gcf must hold 64 16x16 complex doubles in host memory", who);
    return -1;
  }
  if (uvg->elem_size != 8 || !uvg->host || span_of(uvg) < 2 * cells) {
    set_error(RCG_ERR_ARG, "%s: uvg must hold grid_size^2 complex doubles in host memory", who);
    return -1;
  }
  double *vh = reinterpret_cast<double *>(vis->host);
  nv_sort(vh, n, 5);
  const int rc = nv_grid_host(who, scale, reinterpret_cast<double *>(uvg->host), vh, 5, nullptr,
                              5, true, n, grid_size,
                              reinterpret_cast<const double *>(gcf->host));
  // halide_api.cu:17-26
  vis->dev = 0;
  vis->host_dirty = false;
  gcf->dev = 0;
  gcf->host_dirty = false;
  uvg->dev = 0;
  uvg->host_dirty = false;
  uvg->dev_dirty = false;
  return rc;
}

bool nv_raw_args(const char *who, const void *out, const void *in, size_t npts, size_t img,
                 const void *gcf)
{
  if (!out || !gcf || (npts > 0 && !in) || img < 1 || img > 46340 ||
      npts > (size_t)LLONG_MAX / 64) {
    set_error(RCG_ERR_ARG, "%s: null pointer or size out of range", who);
    return false;
  }
  return true;
}

}  // namespace
}  // namespace rcg

extern "C" int nvGridder(const double scale, const int32_t grid_size, buffer_t *vis,
                         buffer_t *gcf, buffer_t *uvg)
{
  return rcg::nv_gridder(scale, grid_size, vis, gcf, uvg);
}

extern "C" void bin_sort3(double *in, int npts)
{
  rcg::clear_error();
  if (in && npts > 1) rcg::nv_sort(in, npts, 3);
}

extern "C" void bin_sort5(double *in, int npts)
{
  rcg::clear_error();
  if (in && npts > 1) rcg::nv_sort(in, npts, 5);
}

extern "C" void gridGPUc(double scale, double *out, double *in_c, size_t npts, size_t img_dim,
                         double *gcf)
{
  rcg::clear_error();
  if (!rcg::nv_raw_args("gridGPUc", out, in_c, npts, img_dim, gcf)) return;
  (void)rcg::nv_grid_host("gridGPUc", scale, out, in_c, 5, nullptr, 5, true, (long long)npts,
                          (int)img_dim, gcf);
}

extern "C" void gridGPUs(double scale, double *out, double *in, double *in_vals, size_t npts,
                         size_t img_dim, double *gcf)
{
  rcg::clear_error();
  if (!rcg::nv_raw_args("gridGPUs", out, in, npts, img_dim, gcf) || (npts > 0 && !in_vals)) {
    if (npts > 0 && !in_vals) rcg::set_error(RCG_ERR_ARG, "gridGPUs: null in_vals");
    return;
  }
  (void)rcg::nv_grid_host("gridGPUs", scale, out, in, 3, in_vals, 2, false, (long long)npts,
                          (int)img_dim, gcf);
}

extern "C" int kern_scatter(const double scale, const int32_t grid_size, const int32_t margin_size,
                            buffer_t *vis, buffer_t *gcf, buffer_t *uvg)
{
  return rcg::scatter(scale, grid_size, margin_size, vis, gcf, uvg, false, -444);
}

extern "C" int kern_degrid(const double scale, const int32_t grid_size, const int32_t margin_size,
                           buffer_t *gcf, buffer_t *uvg, buffer_t *vis, buffer_t *vis_out)
{
  return rcg::degrid(scale, grid_size, margin_size, gcf, uvg, vis, vis_out, false, -555);
}

extern "C" int kern_scatter_gpu1(const double scale, const int32_t grid_size,
                                 const int32_t margin_size, buffer_t *vis, buffer_t *gcf,
                                 buffer_t *uvg)
{
  return rcg::scatter(scale, grid_size, margin_size, vis, gcf, uvg, true, -222);
}

extern "C" int kern_degrid_gpu1(const double scale, const int32_t grid_size,
                                const int32_t margin_size, buffer_t *gcf, buffer_t *uvg,
                                buffer_t *vis, buffer_t *vis_out)
{
  return rcg::degrid(scale, grid_size, margin_size, gcf, uvg, vis, vis_out, true, -333);
}
