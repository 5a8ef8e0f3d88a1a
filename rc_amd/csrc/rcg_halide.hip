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
#include <algorithm>
#include <climits>

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

}  // namespace
}  // namespace rcg

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
