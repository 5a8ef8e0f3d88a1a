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
#include "rcg_tiled.h"
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

// Records of the tiled FP64 gridder for kern_scatter on dense buffers: per
// point the window's first grid cell (rows = y, columns = x, cell = (y0 -
// min_y) * pitch + (x0 - min_x)) and the first tap of its oversampling
// layer (tap (x, y) at over * ostr + y * tstr + x complex elements), or cell
// -1 for a point k_h_scatter would skip; and the visibility packed as a
// complex double.
__global__ void __launch_bounds__(256)
k_h_records(const double *__restrict__ vis, HB vb, HB ub, double scale, int grid_size,
            int margin, int S, int pitch, int tstr, long long ostr, rcg_pg *__restrict__ pg,
            double2 *__restrict__ v2, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *r = vis + i * vb.str[1];
  const Pix p = h_pixel(r[0], r[vb.str[0]], scale, grid_size, S, ub, margin);
  rcg_pg rec;
  rec.cell = -1;
  rec.tap0 = 0;
  rec.stride = (int16_t)tstr;
  rec.supp = (int16_t)S;
  rec.flags = 0;
  if (p.ok) {
    rec.cell = (p.y0 - ub.mn[2]) * pitch + (p.x0 - ub.mn[1]);
    rec.tap0 = (int)(p.over * ostr);
  }
  pg[i] = rec;
  v2[i] = make_double2(r[3 * vb.str[0]], r[4 * vb.str[0]]);
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

struct PlanHold {
  rcg_tiled_plan *p = nullptr;
  double2 *v2 = nullptr;
  ~PlanHold()
  {
    if (v2) (void)hipFree(v2);
    if (p) (void)rcg_tiled_plan_destroy(p);
  }
};

// kern_scatter through the tiled FP64 gridder: binned pairs, one FP64 tile
// per workgroup in LDS, tile sums added to the grid.  Taken for a square
// grid window (extent x = extent y) with interleaved complex doubles (re/im
// stride 1, x stride 2, even y stride) and a GCF of the same kind, where a
// plan fits (at most MAX_PLAN_TILES tiles of 32 or 64, 2^31 pairs).  Returns 1 when
// the points were gridded, 0 when the caller should run k_h_scatter, -1 on
// a failure after work was queued (error set).
int scatter_tiled(double scale, int grid_size, int margin, const buffer_t *vis,
                  const buffer_t *gcf, const buffer_t *uvg, const DevData *d, int S, long long n,
                  hipStream_t st)
{
  const int X = uvg->extent[1];
  if (X != uvg->extent[2] || uvg->stride[0] != 1 || uvg->stride[1] != 2 ||
      (uvg->stride[2] & 1) || gcf->stride[0] != 1 || gcf->stride[1] != 2 ||
      (gcf->stride[2] & 1) || (gcf->stride[3] & 1) || gcf->stride[2] / 2 > 32767 ||
      span_of(gcf) >= 0x7fffffffLL || n > INT_MAX)
    return 0;
  const int pitch = uvg->stride[2] / 2;
  // The plan's geometry: n points in chunks of one on the X x X window.  Its
  // uvw and chunk supports are read only by the pregrid of rcg_tiled_bin,
  // which this route replaces with k_h_records; they point at the
  // visibility records to pass the plan's argument checks.
  rcg_geom g{};
  g.uvw = reinterpret_cast<const rcg_double3 *>(d[0].p);
  g.chunk_supps = reinterpret_cast<const int32_t *>(d[0].p);
  g.n_chunks = (int)n;
  g.ts_ch = 1;
  g.scale = scale;
  g.wstep = 1.0;
  g.grid_size = X;
  g.grid_pitch = pitch;
  PlanHold h;
  for (int tile : {32, 64}) {
    const long long nt = (X + tile - 1) / tile;
    if (nt * nt > MAX_PLAN_TILES) continue;
    if (rcg_tiled_plan_create(&g, tile, 1024, S, &h.p) == RCG_OK) break;
    h.p = nullptr;
  }
  clear_error();
  (void)hipGetLastError();  // a failed plan allocation leaves its error behind
  if (!h.p) return 0;
  if (hipMalloc(reinterpret_cast<void **>(&h.v2), (size_t)n * sizeof(double2)) != hipSuccess) {
    h.v2 = nullptr;
    (void)hipGetLastError();
    return 0;
  }
  hipLaunchKernelGGL(k_h_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d[0].p,
                     hb_of(vis), hb_of(uvg), scale, grid_size, margin, S, pitch,
                     gcf->stride[2] / 2, (long long)(gcf->stride[3] / 2), h.p->pg, h.v2, n);
  if (hipGetLastError() != hipSuccess) {
    set_error(RCG_ERR_HIP, "kern_scatter: record kernel launch failed");
    return -1;
  }
  if (tiled_bin_records(h.p, st) != RCG_OK ||
      tiled_grid_f64_add(h.p, reinterpret_cast<const double2 *>(d[1].p), span_of(gcf) / 2, h.v2,
                         reinterpret_cast<double2 *>(d[2].p), st) != RCG_OK)
    return -1;
  int err = 0;
  if (rcg_tiled_stats(h.p, nullptr, nullptr, nullptr, &err, st) != RCG_OK) return -1;
  if (err) {
    set_error(RCG_ERR_HIP, "kern_scatter: gridding plan error %d", err);
    return -1;
  }
  return 1;
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
    int tiled = 0;
    if (n > 0) tiled = scatter_tiled(scale, grid_size, margin, vis, gcf, uvg, d, S, n, st);
    if (tiled < 0) {
      rc = -1;
    } else if (n > 0 && !tiled) {
      const long long blocks = (n * 64 + 255) / 256;
      hipLaunchKernelGGL(k_h_scatter, dim3((unsigned)blocks), dim3(256), 0, st, d[0].p,
                         hb_of(vis), d[1].p, hb_of(gcf), d[2].p, hb_of(uvg), scale, grid_size,
                         margin, S);
    }
    const size_t bytes = (size_t)span_of(uvg) * sizeof(double);
    if (rc == 0 && (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(uvg->host, d[2].p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)) {
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

// A point's taps land at x = mx + a - 7, y = my + b - 7 for a, b in [0, 16):
// all of them inside the img x img grid.
__device__ __forceinline__ bool nv_inside(int mx, int my, int img)
{
  return mx >= 7 && mx + 8 < img && my >= 7 && my + 8 < img;
}

// Pixel m and sub-pixel offset s of a scaled coordinate xs:
// m = floor(xs), s = floor(8 (xs - m)), as the reference packs them into
// 8 m + s (MS6/kernel/nvidia/gridder/grid_gpu.cu:37-51).  For xs a hair
// below an integer the difference rounds to 1.0 and s to 8, which the
// packing carries into pixel m + 1 with s = 0 (e.g. xs = -1e-300: pixel 0,
// offset 0); the same carry here.  (For other negative xs the reference's
// unpacking by C division, grid_gpu.cu:117-125, would index before the
// table; these stay at the floor pixel.)
__device__ __forceinline__ void nv_split(double xs, double fx, int &m, int &s)
{
  m = (int)fx;
  s = (int)floor(H_OVER * (xs - fx));
  if (s >= H_OVER) {
    m += 1;
    s -= H_OVER;
  }
}

// Records of the tiled FP64 gridder for nvGridder: per point whose window
// lies inside the grid, its first cell ((my - 7) * img + mx - 7, rows = y)
// and the first tap of its 16 x 16 layer (8 sub_y + sub_x) * 256, tap (a,
// b) at + 16 b + a; cell -1 otherwise (k_nv_scatter grids those).  The
// visibility is packed as a complex double.
__global__ void __launch_bounds__(256)
k_nv_records(const double *__restrict__ pos, int ps, const double *__restrict__ val, int vs,
             long long n, double scale, int img, rcg_pg *__restrict__ pg,
             double2 *__restrict__ v2)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double xs = pos[t * ps] * scale, ys = pos[t * ps + 1] * scale;
  const double fx = floor(xs), fy = floor(ys);
  rcg_pg rec;
  rec.cell = -1;
  rec.tap0 = 0;
  rec.stride = NV_GCF;
  rec.supp = NV_GCF;
  rec.flags = 0;
  if (fx >= -8.0 && fx <= img + 6.0 && fy >= -8.0 && fy <= img + 6.0) {
    int mx, my, sx, sy;
    nv_split(xs, fx, mx, sx);
    nv_split(ys, fy, my, sy);
    if (nv_inside(mx, my, img)) {
      rec.cell = (my - 7) * img + (mx - 7);
      rec.tap0 = (H_OVER * sy + sx) * NV_TAPS;
    }
  }
  pg[t] = rec;
  v2[t] = make_double2(val[t * vs], val[t * vs + 1]);
}

// pos: x, y of point t at pos[t * ps], pos[t * ps + 1]; val: re, im at
// val[t * vs], val[t * vs + 1].  skip_inside: leave out the points whose
// window lies inside the grid (gridded by the tiled route).
__global__ void __launch_bounds__(256)
k_nv_scatter(const double *__restrict__ pos, int ps, const double *__restrict__ val, int vs,
             long long n, const double2 *__restrict__ gcf, double *__restrict__ out,
             double scale, int img, int skip_inside)
{
#pragma clang fp contract(off)
  const long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (t >= n) return;
  const double xs = pos[t * ps] * scale, ys = pos[t * ps + 1] * scale;
  const double fx = floor(xs), fy = floor(ys);
  // no tap of a point outside these bounds lands in the grid (also NaN)
  if (!(fx >= -8.0 && fx <= img + 6.0 && fy >= -8.0 && fy <= img + 6.0)) return;
  int mx, my, sx, sy;
  nv_split(xs, fx, mx, sx);
  nv_split(ys, fy, my, sy);
  if (skip_inside && nv_inside(mx, my, img)) return;
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
  PlanHold plan;
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
    // Points whose window lies inside the grid: binned by tile, FP64 tile
    // sums in LDS added to the zeroed grid.  The rest (edges), or all points
    // where no plan fits: one FP64 atomic per tap.
    const double *vals = shared ? dp + 3 : dv;
    int tiled = 0;
    if (e == hipSuccess && n > 0 && n <= INT_MAX) {
      rcg_geom g{};
      // read only by rcg_tiled_bin's pregrid, which k_nv_records replaces
      g.uvw = reinterpret_cast<const rcg_double3 *>(dp);
      g.chunk_supps = reinterpret_cast<const int32_t *>(dp);
      g.n_chunks = (int)n;
      g.ts_ch = 1;
      g.scale = scale;
      g.wstep = 1.0;
      g.grid_size = img;
      g.grid_pitch = img;
      for (int tile : {32, 64}) {
        const long long nt = (img + tile - 1) / tile;
        if (nt * nt > MAX_PLAN_TILES) continue;
        if (rcg_tiled_plan_create(&g, tile, 1024, NV_GCF, &plan.p) == RCG_OK) break;
        plan.p = nullptr;
      }
      clear_error();
      (void)hipGetLastError();  // a failed plan allocation leaves its error behind
      if (plan.p &&
          hipMalloc(reinterpret_cast<void **>(&plan.v2), (size_t)n * sizeof(double2)) ==
              hipSuccess) {
        hipLaunchKernelGGL(k_nv_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           dp, ps, vals, vs, n, scale, img, plan.p->pg, plan.v2);
        e = hipGetLastError();
        if (e == hipSuccess &&
            (tiled_bin_records(plan.p, st) != RCG_OK ||
             tiled_grid_f64_add(plan.p, dg, (long long)H_OVER * H_OVER * NV_TAPS, plan.v2,
                                reinterpret_cast<double2 *>(dout), st) != RCG_OK))
          e = hipErrorLaunchFailure;
        tiled = 1;
      } else {
        plan.v2 = nullptr;
        (void)hipGetLastError();  // the failed optional allocation: scatter instead
      }
    }
    if (e == hipSuccess && n > 0) {
      const long long blocks = (n * 64 + 255) / 256;
      hipLaunchKernelGGL(k_nv_scatter, dim3((unsigned)blocks), dim3(256), 0, st, dp, ps, vals,
                         vs, n, dg, dout, scale, img, tiled);
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

// ---------------------------------------------------------------------------
// MS6 imaging-glue kernels (Halide generators in MS6/kernel/cpu/gridding:
// init.cpp, image_init.cpp, psf_vis.cpp, rotate.cpp, detile.cpp,
// defacet.cpp, image_sum.cpp).  A Halide function is computed over its
// output buffer's region, dimension i covering [min[i], min[i] + extent[i]);
// element (x0, x1, x2, x3) of a buffer lies at sum_i (x_i - min[i]) *
// stride[i] elements from its start.  The buffers are in host memory; each
// call copies the spans it reads to the device and the output back.

struct G4 {
  int mn[4], ext[4], str[4];
};

G4 g4_of(const buffer_t *b, int dims)
{
  G4 g;
  for (int i = 0; i < 4; i++) {
    g.mn[i] = b->min[i];
    g.ext[i] = i < dims ? b->extent[i] : 1;
    g.str[i] = b->stride[i];
  }
  return g;
}

__device__ __forceinline__ long long g4_off(const G4 &g, int x0, int x1, int x2, int x3)
{
  return (long long)(x0 - g.mn[0]) * g.str[0] + (long long)(x1 - g.mn[1]) * g.str[1] +
         (long long)(x2 - g.mn[2]) * g.str[2] + (long long)(x3 - g.mn[3]) * g.str[3];
}

// region index -> coordinates (x0 fastest)
__device__ __forceinline__ void g4_coord(const G4 &g, long long i, int *x)
{
  for (int d = 0; d < 4; d++) {
    const long long e = g.ext[d];
    x[d] = g.mn[d] + (int)(i % e);
    i /= e;
  }
}

__global__ void k_g_zero(double *__restrict__ p, G4 g, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    int x[4];
    g4_coord(g, i, x);
    p[g4_off(g, x[0], x[1], x[2], x[3])] = 0.0;
  }
}

// psf_vis.cpp:17-20: out(c, t) = c == re ? 1 : c == im ? 0 : vis(c, t)
__global__ void k_g_psf_vis(const double *__restrict__ vis, G4 vg, double *__restrict__ out,
                            G4 og, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    int x[4];
    g4_coord(og, i, x);
    const double v =
        x[0] == 3 ? 1.0 : x[0] == 4 ? 0.0 : vis[g4_off(vg, x[0], x[1], vg.mn[2], vg.mn[3])];
    out[g4_off(og, x[0], x[1], x[2], x[3])] = v;
  }
}

// rotate.cpp:48-68: per point, uvw' = mtx uvw and vis' = vis *
// exp(2 pi i (pos . uvw)), with mtx and pos (= w - R w) precomputed
struct RotPar {
  double m[9], pos[3];
};

__global__ void k_g_rotate(const double *__restrict__ vis, G4 vg, double *__restrict__ out,
                           G4 og, RotPar rp)
{
#pragma clang fp contract(off)
  const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long k = t0; k < og.ext[1]; k += (long long)gridDim.x * blockDim.x) {
    const int t = og.mn[1] + (int)k;
    const double *r = vis + (long long)(t - vg.mn[1]) * vg.str[1] - (long long)vg.mn[0] * vg.str[0];
    const double u = r[0], v = r[vg.str[0]], w = r[2 * vg.str[0]];
    const double re = r[3 * vg.str[0]], im = r[4 * vg.str[0]];
    const double nu = rp.m[0] * u + rp.m[1] * v + rp.m[2] * w;
    const double nv = rp.m[3] * u + rp.m[4] * v + rp.m[5] * w;
    const double nw = rp.m[6] * u + rp.m[7] * v + rp.m[8] * w;
    const double path = rp.pos[0] * u + rp.pos[1] * v + rp.pos[2] * w;
    const double phi = 2.0 * (4.0 * atan(1.0)) * path;
    const double c = cos(phi), s = sin(phi);
    double *o = out + (long long)(t - og.mn[1]) * og.str[1] - (long long)og.mn[0] * og.str[0];
    const double vals[5] = {nu, nv, nw, c * re - s * im, c * im + s * re};
    for (int q = og.mn[0]; q < og.mn[0] + og.ext[0]; q++) o[q * og.str[0]] = vals[q];
  }
}

// detile.cpp:19-30: out(c, x, y) += in(c, x, y) over the overlap of both
// buffers' x and y ranges (c in [0, 2))
__global__ void k_g_detile(const double *__restrict__ in, G4 ig, double *__restrict__ out,
                           G4 og, int x0, int nx, int y0, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i & 1);
    const long long p = i >> 1;
    const int x = x0 + (int)(p % nx), y = y0 + (int)(p / nx);
    out[g4_off(og, c, x, y, og.mn[3])] += in[g4_off(ig, c, x, y, ig.mn[3])];
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

// defacet.cpp:12-20: image(clamp(x + l, left, right), clamp(y + m, left,
// right)) = facet(x, y, l, m), x fastest; the last write in that order
// wins.  Pass 1 records the largest source index per target pixel, pass 2
// lets that source write.  left / right are the image's dimension-0 bounds,
// used for y too, as in the reference.
__global__ void k_g_defacet_claim(G4 fg, G4 og, int left, int right,
                                  unsigned long long *__restrict__ win, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    int x[4];
    g4_coord(fg, i, x);
    const int ix = clampi(x[0] + x[2], left, right), iy = clampi(x[1] + x[3], left, right);
    const long long p = (long long)(iy - og.mn[1]) * og.ext[0] + (ix - og.mn[0]);
    atomicMax(&win[p], (unsigned long long)i + 1);
  }
}

__global__ void k_g_defacet_write(const double *__restrict__ facet, G4 fg, double *__restrict__ img,
                                  G4 og, int left, int right,
                                  const unsigned long long *__restrict__ win, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    int x[4];
    g4_coord(fg, i, x);
    const int ix = clampi(x[0] + x[2], left, right), iy = clampi(x[1] + x[3], left, right);
    const long long p = (long long)(iy - og.mn[1]) * og.ext[0] + (ix - og.mn[0]);
    if (win[p] == (unsigned long long)i + 1)
      img[g4_off(og, ix, iy, og.mn[2], og.mn[3])] = facet[g4_off(fg, x[0], x[1], x[2], x[3])];
  }
}

// image_sum.cpp:12-20: sum(clamp(x, left, right), clamp(y, left, right)) +=
// image(x, y) over the input image's region (left / right: the output's
// dimension-0 bounds, for y too)
__global__ void k_g_image_sum(const double *__restrict__ img, G4 ig, double *__restrict__ sum,
                              G4 og, int left, int right, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    int x[4];
    g4_coord(ig, i, x);
    const int ox = clampi(x[0], left, right), oy = clampi(x[1], left, right);
    unsafeAtomicAdd(&sum[g4_off(og, ox, oy, og.mn[2], og.mn[3])],
                    img[g4_off(ig, x[0], x[1], x[2], x[3])]);
  }
}

long long region_size(const G4 &g)
{
  long long n = 1;
  for (int i = 0; i < 4; i++) n *= g.ext[i];
  return n;
}

// elements from the buffer's start to one past its last element (strides
// positive, offsets measured from min)
long long span_elems(const buffer_t *b, int dims)
{
  long long s = 1;
  for (int i = 0; i < dims; i++)
    if (b->extent[i] > 0) s += (long long)(b->extent[i] - 1) * b->stride[i];
  return s;
}

bool host_buf(const buffer_t *b, int dims, const char *who, const char *name)
{
  bool ok = b && b->host && b->elem_size == 8;
  for (int i = 0; ok && i < dims; i++) ok = b->extent[i] >= 1 && b->stride[i] >= 1;
  if (!ok) set_error(RCG_ERR_ARG, "%s: %s must be a host buffer of doubles with %d dimensions", who,
                     name, dims);
  return ok;
}

// Runs `launch` on device copies of the buffers: ins are uploaded, outs
// uploaded (when `keep`, for updates that read them) and downloaded.
struct Glue {
  const char *who;
  hipStream_t st = nullptr;
  std::vector<double *> dev;
  int rc = 0;
  explicit Glue(const char *w) : who(w)
  {
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) fail(RCG_ERR_HIP);
  }
  void fail(int code)
  {
    if (!rc) set_error(code, "%s: HIP call failed", who);
    rc = -1;
  }
  double *upload(const buffer_t *b, int dims, bool copy)
  {
    if (rc) return nullptr;
    const size_t bytes = (size_t)span_elems(b, dims) * sizeof(double);
    double *p = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&p), bytes) != hipSuccess) {
      fail(RCG_ERR_NOMEM);
      return nullptr;
    }
    dev.push_back(p);
    if (copy && hipMemcpyAsync(p, b->host, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
      fail(RCG_ERR_HIP);
    return p;
  }
  void download(buffer_t *b, int dims, const double *p)
  {
    if (rc) return;
    const size_t bytes = (size_t)span_elems(b, dims) * sizeof(double);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(b->host, p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      fail(RCG_ERR_HIP);
  }
  ~Glue()
  {
    if (st) (void)hipStreamSynchronize(st);
    for (double *p : dev) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
};

unsigned blocks_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 65536); }

int g_zero(buffer_t *b, int dims, const char *who)
{
  clear_error();
  if (!host_buf(b, dims, who, "output")) return -1;
  Glue g(who);
  double *p = g.upload(b, dims, true);
  const G4 o = g4_of(b, dims);
  const long long n = region_size(o);
  if (!g.rc) hipLaunchKernelGGL(k_g_zero, dim3(blocks_for(n)), dim3(256), 0, g.st, p, o, n);
  g.download(b, dims, p);
  return g.rc;
}

// rotate.cpp:33-52 on the host, in FP64
RotPar rotation(double in_lon, double in_lat, double out_lon, double out_lat, int uvproj)
{
  const double pi = 4.0 * std::atan(1.0);
  auto mul = [](const double *a, const double *b, double *c) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  };
  auto rx = [](double p, double *m) {
    const double s = std::sin(p), c = std::cos(p);
    const double r[9] = {1, 0, 0, 0, c, -s, 0, s, c};
    for (int i = 0; i < 9; i++) m[i] = r[i];
  };
  auto rz = [](double p, double *m) {
    const double s = std::sin(p), c = std::cos(p);
    const double r[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    for (int i = 0; i < 9; i++) m[i] = r[i];
  };
  double a[9], b[9], x2u[9], u2x[9], R[9];
  rx(-(pi / 2) + out_lat, a);  // xyz2uvw(out) = rotateX(lat - pi/2) rotateZ(-pi/2 - lon)
  rz(-(pi / 2) - out_lon, b);
  mul(a, b, x2u);
  rz((pi / 2) + in_lon, a);  // uvw2xyz(in) = rotateZ(pi/2 + lon) rotateX(pi/2 - lat)
  rx((pi / 2) - in_lat, b);
  mul(a, b, u2x);
  mul(x2u, u2x, R);
  RotPar rp;
  if (uvproj != 0) {
    // proj = (R as 2x2, inverted, third row/column identity) transposed
    const double det = R[0] * R[4] - R[1] * R[3];
    const double inv[9] = {R[4] / det, -R[1] / det, 0, -R[3] / det, R[0] / det, 0, 0, 0, 1};
    const double proj[9] = {inv[0], inv[3], inv[6], inv[1], inv[4], inv[7], inv[2], inv[5], inv[8]};
    const double Rt[9] = {R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]};
    double pr[9];
    mul(proj, R, pr);
    mul(Rt, pr, rp.m);
  } else {
    const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 9; i++) rp.m[i] = id[i];
  }
  // pos = w - R w, w = (0, 0, 1)
  rp.pos[0] = 0.0 - R[2];
  rp.pos[1] = 0.0 - R[5];
  rp.pos[2] = 1.0 - R[8];
  return rp;
}

// Halide's constraints on the 5-double visibility buffers (psf_vis.cpp:11-13,
// 22-24; rotate.cpp:18-20, 71-73): dimension 0 is min 0, extent 5, stride 1;
// dimension 1 has stride 5
bool vis5(const buffer_t *b, const char *who, const char *name)
{
  if (!host_buf(b, 2, who, name)) return false;
  if (b->min[0] != 0 || b->extent[0] != 5 || b->stride[0] != 1 || b->stride[1] != 5) {
    set_error(RCG_ERR_ARG, "%s: %s must be 5 contiguous doubles per point", who, name);
    return false;
  }
  return true;
}

bool covers_t(const buffer_t *in, const buffer_t *out, const char *who)
{
  if (in->min[1] > out->min[1] ||
      (long long)in->min[1] + in->extent[1] < (long long)out->min[1] + out->extent[1]) {
    set_error(RCG_ERR_ARG, "%s: the input does not cover the output's points", who);
    return false;
  }
  return true;
}

int psf_vis(buffer_t *vis, buffer_t *out)
{
  clear_error();
  const char *who = "kern_psf_vis";
  if (!vis5(vis, who, "vis") || !vis5(out, who, "output") || !covers_t(vis, out, who)) return -1;
  Glue g(who);
  const double *dv = g.upload(vis, 2, true);
  double *dout = g.upload(out, 2, false);
  const G4 vg = g4_of(vis, 2), og = g4_of(out, 2);
  const long long n = region_size(og);
  if (!g.rc)
    hipLaunchKernelGGL(k_g_psf_vis, dim3(blocks_for(n)), dim3(256), 0, g.st, dv, vg, dout, og, n);
  g.download(out, 2, dout);
  return g.rc;
}

int rotate(double in_lon, double in_lat, double out_lon0, double out_lat0, double out_lon_incr,
           double out_lat_incr, int uvproj, buffer_t *vis, buffer_t *rot)
{
  clear_error();
  const char *who = "kern_rotate";
  if (!vis5(vis, who, "vis") || !vis5(rot, who, "output") || !covers_t(vis, rot, who)) return -1;
  // the facet's (l, m) arrive as the output buffer's min[2], min[3]
  const double out_lon = out_lon0 + rot->min[2] * out_lon_incr;
  const double out_lat = out_lat0 + rot->min[3] * out_lat_incr;
  const RotPar rp = rotation(in_lon, in_lat, out_lon, out_lat, uvproj);
  Glue g(who);
  const double *dv = g.upload(vis, 2, true);
  double *dout = g.upload(rot, 2, false);
  const G4 vg = g4_of(vis, 2), og = g4_of(rot, 2);
  if (!g.rc)
    hipLaunchKernelGGL(k_g_rotate, dim3(blocks_for(og.ext[1])), dim3(256), 0, g.st, dv, vg, dout,
                       og, rp);
  g.download(rot, 2, dout);
  return g.rc;
}

// detile.cpp:11-12, 32: complex pairs contiguous (dimension 0 extent 2,
// stride 1; dimension 1 stride 2)
bool cplx3(const buffer_t *b, const char *who, const char *name)
{
  if (!host_buf(b, 3, who, name)) return false;
  if (b->extent[0] != 2 || b->stride[0] != 1 || b->stride[1] != 2) {
    set_error(RCG_ERR_ARG, "%s: %s must hold contiguous complex pairs", who, name);
    return false;
  }
  return true;
}

int detile(buffer_t *in, buffer_t *out)
{
  clear_error();
  const char *who = "kern_detile";
  if (!cplx3(in, who, "in_grid") || !cplx3(out, who, "out_grid")) return -1;
  const int x0 = std::max(out->min[1], in->min[1]);
  const int x1 = (int)std::min((long long)out->min[1] + out->extent[1],
                               (long long)in->min[1] + in->extent[1]);
  const int y0 = std::max(out->min[2], in->min[2]);
  const int y1 = (int)std::min((long long)out->min[2] + out->extent[2],
                               (long long)in->min[2] + in->extent[2]);
  if (x1 <= x0 || y1 <= y0) return 0;
  Glue g(who);
  const double *di = g.upload(in, 3, true);
  double *dout = g.upload(out, 3, true);
  const long long n = 2LL * (x1 - x0) * (y1 - y0);
  if (!g.rc)
    hipLaunchKernelGGL(k_g_detile, dim3(blocks_for(n)), dim3(256), 0, g.st, di, g4_of(in, 3),
                       dout, g4_of(out, 3), x0, x1 - x0, y0, n);
  g.download(out, 3, dout);
  return g.rc;
}

// The y targets clamp(y, left, right) of y in [lo, hi] must lie in the
// output's dimension-1 range (Halide checks the output's bounds).
bool y_in(const buffer_t *o, long long lo, long long hi, const char *who)
{
  const long long left = o->min[0], right = (long long)o->min[0] + o->extent[0] - 1;
  const long long a = std::min(std::max(lo, left), right), b = std::min(std::max(hi, left), right);
  if (a < o->min[1] || b > (long long)o->min[1] + o->extent[1] - 1) {
    set_error(RCG_ERR_ARG, "%s: rows %lld..%lld lie outside the output", who, a, b);
    return false;
  }
  return true;
}

int defacet(buffer_t *facet, buffer_t *image)
{
  clear_error();
  const char *who = "kern_defacet";
  if (!host_buf(facet, 4, who, "facet") || !host_buf(image, 2, who, "image")) return -1;
  const long long ylo = (long long)facet->min[1] + facet->min[3];
  const long long yhi = ylo + facet->extent[1] - 1 + facet->extent[3] - 1;
  if (!y_in(image, ylo, yhi, who)) return -1;
  const int left = image->min[0], right = image->min[0] + image->extent[0] - 1;
  Glue g(who);
  const double *df = g.upload(facet, 4, true);
  double *di = g.upload(image, 2, true);
  const long long pix = (long long)image->extent[0] * image->extent[1];
  unsigned long long *win = nullptr;
  if (!g.rc && hipMalloc(reinterpret_cast<void **>(&win), pix * sizeof(unsigned long long)) !=
                   hipSuccess)
    g.fail(RCG_ERR_NOMEM);
  const G4 fg = g4_of(facet, 4), og = g4_of(image, 2);
  const long long n = region_size(fg);
  if (!g.rc) {
    hipLaunchKernelGGL(k_nv_zero, dim3(blocks_for(pix)), dim3(256), 0, g.st,
                       reinterpret_cast<double *>(win), pix);
    hipLaunchKernelGGL(k_g_defacet_claim, dim3(blocks_for(n)), dim3(256), 0, g.st, fg, og, left,
                       right, win, n);
    hipLaunchKernelGGL(k_g_defacet_write, dim3(blocks_for(n)), dim3(256), 0, g.st, df, fg, di, og,
                       left, right, win, n);
  }
  g.download(image, 2, di);
  if (win) (void)hipFree(win);
  return g.rc;
}

int image_sum(buffer_t *image, buffer_t *sum)
{
  clear_error();
  const char *who = "kern_image_sum";
  if (!host_buf(image, 2, who, "image") || !host_buf(sum, 2, who, "sum")) return -1;
  if (!y_in(sum, image->min[1], (long long)image->min[1] + image->extent[1] - 1, who)) return -1;
  const int left = sum->min[0], right = sum->min[0] + sum->extent[0] - 1;
  Glue g(who);
  const double *di = g.upload(image, 2, true);
  double *ds = g.upload(sum, 2, true);
  const G4 ig = g4_of(image, 2);
  const long long n = region_size(ig);
  if (!g.rc)
    hipLaunchKernelGGL(k_g_image_sum, dim3(blocks_for(n)), dim3(256), 0, g.st, di, ig, ds,
                       g4_of(sum, 2), left, right, n);
  g.download(sum, 2, ds);
  return g.rc;
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

// MS6 imaging glue (Halide generators in MS6/kernel/cpu/gridding)
extern "C" int kern_init(buffer_t *grid) { return rcg::g_zero(grid, 3, "kern_init"); }

extern "C" int kern_image_init(buffer_t *image)
{
  return rcg::g_zero(image, 2, "kern_image_init");
}

extern "C" int kern_psf_vis(buffer_t *vis, buffer_t *out) { return rcg::psf_vis(vis, out); }

extern "C" int kern_rotate(const double in_lon, const double in_lat, const double out_lon0,
                           const double out_lat0, const double out_lon_incr,
                           const double out_lat_incr, const int32_t uvproj, buffer_t *vis,
                           buffer_t *rot)
{
  return rcg::rotate(in_lon, in_lat, out_lon0, out_lat0, out_lon_incr, out_lat_incr, uvproj, vis,
                     rot);
}

extern "C" int kern_detile(buffer_t *in_grid, buffer_t *out_grid)
{
  return rcg::detile(in_grid, out_grid);
}

extern "C" int kern_defacet(buffer_t *facet, buffer_t *image)
{
  return rcg::defacet(facet, image);
}

extern "C" int kern_image_sum(buffer_t *image, buffer_t *sum)
{
  return rcg::image_sum(image, sum);
}
