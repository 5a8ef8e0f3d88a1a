// rcg_small.hip -- the small kernels next to the gridder: normalisation,
// FP64 phase rotation, nearest-cell gridding, the hermitian fold and the
// even-size fftshift; device-resident entry points plus the reference's
// host-memory entry points over them (uniform reweighting's device kernels
// are in rcg_reweight.hip).
//
// Reference semantics (MS4/kernel/cpu/scatter_grid/
// scatter_gridder_w_dependent_dyn_1p.cpp, MS4/kernel/cpu/herm/
// herm_padded.cpp, MS4/kernel/cpu/fft/fft_dyn_padded.cpp) are cited at each
// kernel.
#include <climits>
#include <cstring>
#include <vector>

#include <algorithm>

#include "rcg_internal.h"

namespace rcg {

// grid *= 1 / (size * pitch) over every cell, pad included (.cpp:293-306)
template <class T>
__global__ void k_normalize(T *__restrict__ g, long long n, T norm)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    g[i] *= norm;
}

// vis *= exp(+2 pi i * w * scale), phase in FP64 (.cpp:355-363, common.h:90-94)
template <class C>
__global__ void k_rotate(const rcg_double3 *__restrict__ uvw, C *__restrict__ vis, int n,
                         double scale)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ph = 2.0 * M_PI * (uvw[i].w * scale);
  double s, c;
  sincos(ph, &s, &c);
  const double vr = vis[i].x, vi = vis[i].y;
  using E = decltype(C{}.x);
  vis[i].x = (E)(vr * c - vi * s);
  vis[i].y = (E)(vr * s + vi * c);
}

// Nearest-cell gridding into a centred grid (.cpp:268-291): cell
// k = round(u*scale)*pitch + round(v*scale) relative to the centre, kept
// when k + trans lies in [0, size*pitch).
__global__ void k_grid0(const rcg_double3 *__restrict__ uvw, const double2 *__restrict__ vis,
                        double2 *__restrict__ centred, double scale, int n, long long last,
                        long long trans, int pitch)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long u = (long long)round(uvw[i].u * scale);
  const long long v = (long long)round(uvw[i].v * scale);
  const long long k = u * pitch + v;
  if (k + trans >= 0 && k + trans < last) {
    double *g = reinterpret_cast<double *>(centred + k);
    unsafeAtomicAdd(g, vis[i].x);
    unsafeAtomicAdd(g + 1, vis[i].y);
  }
}

// A[i][j] += conj(A[size-1-i][size-1-j]) for j < size/2
// (herm_padded.cpp:6-27, default branch); sources are never written.
template <class C>
__global__ void k_herm(C *__restrict__ d, int size, int pitch)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= size / 2 || i >= size) return;
  const C s = d[(long long)(size - 1 - i) * pitch + (size - 1 - j)];
  C &t = d[(long long)i * pitch + j];
  t.x += s.x;
  t.y -= s.y;
}

// out -= in (MS4/kernel/gpu/nvidia_degrid/nvidia_degrid.cu:11-18, the
// major cycle's subtraction of predicted from observed visibilities).
template <class C>
__global__ void k_subtract(C *__restrict__ out, const C *__restrict__ in, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    C o = out[i];
    const C a = in[i];
    o.x -= a.x;
    o.y -= a.y;
    out[i] = o;
  }
}

// The MS6 visibility layout, five doubles (u, v, w, re, im) per point
// (MS6/programs/Kernel/IO.hs:59-71), split into uvw and visibilities, the
// visibility optionally multiplied by exp(2 pi i w scale) in FP64 (rotw,
// MS4/kernel/common/common.h:90-94, as the binner applies it,
// MS4/kernel/cpu/binner/binner.cpp:135).
template <class C>
__global__ void k_unpack5(const double *__restrict__ rec, long long n, int rotate, double scale,
                          rcg_double3 *__restrict__ uvw, C *__restrict__ vis)
{
#pragma clang fp contract(off)
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double *r = rec + 5 * i;
    const double u = r[0], v = r[1], w = r[2];
    double re = r[3], im = r[4];
    if (rotate) {
      double s, c;
      sincos(2.0 * M_PI * (w * scale), &s, &c);
      const double xr = re * c - im * s, xi = re * s + im * c;
      re = xr;
      im = xi;
    }
    uvw[i].u = u;
    uvw[i].v = v;
    uvw[i].w = w;
    C o;
    o.x = re;
    o.y = im;
    vis[i] = o;
  }
}

// Back to five doubles per point.
template <class C>
__global__ void k_pack5(const rcg_double3 *__restrict__ uvw, const C *__restrict__ vis,
                        long long n, double *__restrict__ rec)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double *r = rec + 5 * i;
    r[0] = uvw[i].u;
    r[1] = uvw[i].v;
    r[2] = uvw[i].w;
    r[3] = (double)vis[i].x;
    r[4] = (double)vis[i].y;
  }
}

// vis = value (constVis, MS4/dna-programs/Data.hs:242-247: the PSF's
// visibilities are all 1).
template <class C, class E>
__global__ void k_const(C *__restrict__ v, long long n, E re, E im)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    C c;
    c.x = re;
    c.y = im;
    v[i] = c;
  }
}

// Quadrant swap of an even size x size array, row pitch in elements
// (fft_dyn_padded.cpp:26-48): (i, j) <-> (i + size/2, (j + size/2) % size)
// for i < size/2.
template <int B>
__global__ void k_fftshift(char *__restrict__ d, int size, int pitch)
{
  struct alignas(B < 16 ? B : 16) E { char b[B]; };
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  const int h = size / 2;
  if (j >= size || i >= h) return;
  const int jj = j < h ? j + h : j - h;
  E *a = reinterpret_cast<E *>(d + ((long long)i * pitch + j) * B);
  E *b = reinterpret_cast<E *>(d + ((long long)(i + h) * pitch + jj) * B);
  const E t = *a;
  *a = *b;
  *b = t;
}

int blocks_for(long long n, int threads = 256)
{
  long long b = (n + threads - 1) / threads;
  return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

rcg_status launch_fftshift(void *data, int elem_bytes, int size, int pitch, hipStream_t st)
{
  if (size <= 0 || (size & 1) || pitch < size) {
    set_error(RCG_ERR_ARG, "fftshift needs an even size <= pitch (size %d, pitch %d)", size,
              pitch);
    return RCG_ERR_ARG;
  }
  dim3 grid((size + 255) / 256, size / 2);
  char *d = static_cast<char *>(data);
  switch (elem_bytes) {
    case 4: hipLaunchKernelGGL(k_fftshift<4>, grid, dim3(256), 0, st, d, size, pitch); break;
    case 8: hipLaunchKernelGGL(k_fftshift<8>, grid, dim3(256), 0, st, d, size, pitch); break;
    case 16: hipLaunchKernelGGL(k_fftshift<16>, grid, dim3(256), 0, st, d, size, pitch); break;
    default:
      set_error(RCG_ERR_ARG, "fftshift element size %d not 4, 8 or 16", elem_bytes);
      return RCG_ERR_ARG;
  }
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
};

bool grid_dims_ok(int pitch, int size)
{
  if (size <= 0 || pitch < size || (long long)size * pitch > 0x7fffffffLL) {
    set_error(RCG_ERR_ARG, "bad grid dimensions (size %d, pitch %d)", size, pitch);
    return false;
  }
  return true;
}

}  // namespace
}  // namespace rcg

using namespace rcg;

// ------------------------------------------------------------ device API --

extern "C" rcg_status rcg_normalize(void *grid, int f64, int grid_pitch, int grid_size,
                                    void *stream)
{
  clear_error();
  if (!grid || !grid_dims_ok(grid_pitch, grid_size)) {
    if (!grid) set_error(RCG_ERR_ARG, "rcg_normalize: null grid");
    return RCG_ERR_ARG;
  }
  const long long n = 2LL * grid_size * grid_pitch;
  hipStream_t st = (hipStream_t)stream;
  if (f64) {
    const double norm = 1.0 / (double)(grid_size * grid_pitch);
    hipLaunchKernelGGL(k_normalize<double>, dim3(blocks_for(n)), dim3(256), 0, st,
                       static_cast<double *>(grid), n, norm);
  } else {
    const float norm = (float)(1.0 / (double)(grid_size * grid_pitch));
    hipLaunchKernelGGL(k_normalize<float>, dim3(blocks_for(n)), dim3(256), 0, st,
                       static_cast<float *>(grid), n, norm);
  }
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_rotate(const rcg_double3 *uvw, void *vis, int f64, int n,
                                 double scale, void *stream)
{
  clear_error();
  if (n < 0 || (n > 0 && (!uvw || !vis))) {
    set_error(RCG_ERR_ARG, "rcg_rotate: bad arguments");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  hipStream_t st = (hipStream_t)stream;
  if (f64)
    hipLaunchKernelGGL(k_rotate<double2>, dim3((n + 255) / 256), dim3(256), 0, st, uvw,
                       static_cast<double2 *>(vis), n, scale);
  else
    hipLaunchKernelGGL(k_rotate<float2>, dim3((n + 255) / 256), dim3(256), 0, st, uvw,
                       static_cast<float2 *>(vis), n, scale);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_grid0(const rcg_double3 *uvw, const rcg_c128 *vis,
                                rcg_c128 *grid_centred, double scale, int n, int grid_pitch,
                                int grid_size, void *stream)
{
  clear_error();
  if (!grid_dims_ok(grid_pitch, grid_size)) return RCG_ERR_ARG;
  if (n < 0 || (n > 0 && (!uvw || !vis || !grid_centred))) {
    set_error(RCG_ERR_ARG, "rcg_grid0: bad arguments");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  const long long last = (long long)grid_size * grid_pitch;
  const long long trans = (long long)(grid_size / 2) * (grid_pitch + 1);
  hipLaunchKernelGGL(k_grid0, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, uvw,
                     reinterpret_cast<const double2 *>(vis),
                     reinterpret_cast<double2 *>(grid_centred), scale, n, last, trans,
                     grid_pitch);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_herm_padded(rcg_c128 *data, int size, int pitch, void *stream)
{
  clear_error();
  if (!data || (size & 1) || !grid_dims_ok(pitch, size)) {
    if (!data || (size & 1)) set_error(RCG_ERR_ARG, "rcg_herm_padded: null data or odd size");
    return RCG_ERR_ARG;
  }
  hipLaunchKernelGGL(k_herm<double2>, dim3((size / 2 + 255) / 256, size), dim3(256), 0,
                     (hipStream_t)stream, reinterpret_cast<double2 *>(data), size, pitch);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_herm(void *data, int f64, int size, int pitch, void *stream)
{
  clear_error();
  if (!data || size <= 0 || pitch < size) {
    set_error(RCG_ERR_ARG, "rcg_herm: bad arguments");
    return RCG_ERR_ARG;
  }
  const dim3 g((size / 2 + 255) / 256, size);
  if (f64)
    hipLaunchKernelGGL(k_herm<double2>, g, dim3(256), 0, (hipStream_t)stream,
                       static_cast<double2 *>(data), size, pitch);
  else
    hipLaunchKernelGGL(k_herm<float2>, g, dim3(256), 0, (hipStream_t)stream,
                       static_cast<float2 *>(data), size, pitch);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_subtract_vis(void *out, const void *in, int f64, long long n,
                                       void *stream)
{
  clear_error();
  if ((n > 0 && (!out || !in)) || n < 0) {
    set_error(RCG_ERR_ARG, "rcg_subtract_vis: bad arguments");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>((n + 255) / 256, 4096);
  if (f64)
    hipLaunchKernelGGL(k_subtract<double2>, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       static_cast<double2 *>(out), static_cast<const double2 *>(in), n);
  else
    hipLaunchKernelGGL(k_subtract<float2>, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       static_cast<float2 *>(out), static_cast<const float2 *>(in), n);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_unpack_vis5(const double *rec, long long n, int rotate, double scale,
                                      rcg_double3 *uvw, void *vis, int f64, void *stream)
{
  clear_error();
  if (n < 0 || (n > 0 && (!rec || !uvw || !vis))) {
    set_error(RCG_ERR_ARG, "rcg_unpack_vis5: bad arguments");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>((n + 255) / 256, 4096);
  if (f64)
    hipLaunchKernelGGL(k_unpack5<double2>, dim3(nb), dim3(256), 0, (hipStream_t)stream, rec, n,
                       rotate, scale, uvw, static_cast<double2 *>(vis));
  else
    hipLaunchKernelGGL(k_unpack5<float2>, dim3(nb), dim3(256), 0, (hipStream_t)stream, rec, n,
                       rotate, scale, uvw, static_cast<float2 *>(vis));
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_pack_vis5(const rcg_double3 *uvw, const void *vis, int f64,
                                    long long n, double *rec, void *stream)
{
  clear_error();
  if (n < 0 || (n > 0 && (!rec || !uvw || !vis))) {
    set_error(RCG_ERR_ARG, "rcg_pack_vis5: bad arguments");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>((n + 255) / 256, 4096);
  if (f64)
    hipLaunchKernelGGL(k_pack5<double2>, dim3(nb), dim3(256), 0, (hipStream_t)stream, uvw,
                       static_cast<const double2 *>(vis), n, rec);
  else
    hipLaunchKernelGGL(k_pack5<float2>, dim3(nb), dim3(256), 0, (hipStream_t)stream, uvw,
                       static_cast<const float2 *>(vis), n, rec);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_const_vis(void *vis, int f64, long long n, double re, double im,
                                    void *stream)
{
  clear_error();
  if ((n > 0 && !vis) || n < 0) {
    set_error(RCG_ERR_ARG, "rcg_const_vis: bad arguments");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>((n + 255) / 256, 4096);
  if (f64)
    hipLaunchKernelGGL((k_const<double2, double>), dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       static_cast<double2 *>(vis), n, re, im);
  else
    hipLaunchKernelGGL((k_const<float2, float>), dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       static_cast<float2 *>(vis), n, (float)re, (float)im);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_fftshift_even(void *data, int elem_bytes, int size, int pitch,
                                        void *stream)
{
  clear_error();
  if (!data) {
    set_error(RCG_ERR_ARG, "rcg_fftshift_even: null data");
    return RCG_ERR_ARG;
  }
  return launch_fftshift(data, elem_bytes, size, pitch, (hipStream_t)stream);
}

// ------------------------------------------------ reference host entry points --
// Each copies its buffers to the device, runs the kernel and copies back.

namespace {

template <class F>
void host_call(const char *what, F &&body)
{
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    set_error(RCG_ERR_HIP, "%s: stream creation failed", what);
    return;
  }
  body(st);
  (void)hipStreamDestroy(st);
}

}  // namespace

extern "C" void normalizeCPU(rcg_c128 src[], int grid_pitch, int grid_size)
{
  clear_error();
  if (!src || !grid_dims_ok(grid_pitch, grid_size)) {
    if (!src) set_error(RCG_ERR_ARG, "normalizeCPU: null grid");
    return;
  }
  host_call("normalizeCPU", [&](hipStream_t st) {
    const size_t bytes = (size_t)grid_size * grid_pitch * sizeof(rcg_c128);
    DevBuf d;
    if (d.alloc(bytes) != hipSuccess) { set_error(RCG_ERR_NOMEM, "normalizeCPU: device allocation"); return; }
    if (hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess) { set_error(RCG_ERR_HIP, "normalizeCPU: copy in"); return; }
    if (rcg_normalize(d.p, 1, grid_pitch, grid_size, st) != RCG_OK) return;
    if (hipMemcpyAsync(src, d.p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "normalizeCPU: copy out");
  });
}

extern "C" void rotateCPU(const rcg_double3 uvw[], rcg_c128 vis[], int n, double scale)
{
  clear_error();
  if (n <= 0) return;
  if (!uvw || !vis) { set_error(RCG_ERR_ARG, "rotateCPU: null argument"); return; }
  host_call("rotateCPU", [&](hipStream_t st) {
    DevBuf du, dv;
    if (du.alloc((size_t)n * sizeof(rcg_double3)) != hipSuccess ||
        dv.alloc((size_t)n * sizeof(rcg_c128)) != hipSuccess) {
      set_error(RCG_ERR_NOMEM, "rotateCPU: device allocation");
      return;
    }
    if (hipMemcpyAsync(du.p, uvw, (size_t)n * sizeof(rcg_double3), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dv.p, vis, (size_t)n * sizeof(rcg_c128), hipMemcpyHostToDevice, st) != hipSuccess) {
      set_error(RCG_ERR_HIP, "rotateCPU: copy in");
      return;
    }
    if (rcg_rotate(static_cast<const rcg_double3 *>(du.p), dv.p, 1, n, scale, st) != RCG_OK) return;
    if (hipMemcpyAsync(vis, dv.p, (size_t)n * sizeof(rcg_c128), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "rotateCPU: copy out");
  });
}

extern "C" void grid0(const rcg_double3 uvw[], const rcg_c128 vis[], rcg_c128 grid[],
                      double scale, int n, int grid_pitch, int grid_size)
{
  clear_error();
  if (!grid_dims_ok(grid_pitch, grid_size)) return;
  if (n <= 0) return;
  if (!uvw || !vis || !grid) { set_error(RCG_ERR_ARG, "grid0: null argument"); return; }
  // grid is centred by the caller: the buffer starts trans cells before it
  const long long trans = (long long)(grid_size / 2) * (grid_pitch + 1);
  const size_t gbytes = (size_t)grid_size * grid_pitch * sizeof(rcg_c128);
  rcg_c128 *base = grid - trans;
  host_call("grid0", [&](hipStream_t st) {
    DevBuf du, dv, dg;
    if (du.alloc((size_t)n * sizeof(rcg_double3)) != hipSuccess ||
        dv.alloc((size_t)n * sizeof(rcg_c128)) != hipSuccess || dg.alloc(gbytes) != hipSuccess) {
      set_error(RCG_ERR_NOMEM, "grid0: device allocation");
      return;
    }
    if (hipMemcpyAsync(du.p, uvw, (size_t)n * sizeof(rcg_double3), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dv.p, vis, (size_t)n * sizeof(rcg_c128), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dg.p, base, gbytes, hipMemcpyHostToDevice, st) != hipSuccess) {
      set_error(RCG_ERR_HIP, "grid0: copy in");
      return;
    }
    rcg_c128 *centred = static_cast<rcg_c128 *>(dg.p) + trans;
    if (rcg_grid0(static_cast<const rcg_double3 *>(du.p), static_cast<const rcg_c128 *>(dv.p),
                  centred, scale, n, grid_pitch, grid_size, st) != RCG_OK)
      return;
    if (hipMemcpyAsync(base, dg.p, gbytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "grid0: copy out");
  });
}

extern "C" void reweight(const rcg_double3 uvw[], rcg_c128 vis[], double scale, int n,
                         int grid_size)
{
  clear_error();
  if (n <= 0) return;
  if (!uvw || !vis || grid_size <= 0) { set_error(RCG_ERR_ARG, "reweight: bad argument"); return; }
  host_call("reweight", [&](hipStream_t st) {
    DevBuf du, dv, dc;
    if (du.alloc((size_t)n * sizeof(rcg_double3)) != hipSuccess ||
        dv.alloc((size_t)n * sizeof(rcg_c128)) != hipSuccess ||
        dc.alloc((size_t)grid_size * grid_size * sizeof(int)) != hipSuccess) {
      set_error(RCG_ERR_NOMEM, "reweight: device allocation");
      return;
    }
    if (hipMemcpyAsync(du.p, uvw, (size_t)n * sizeof(rcg_double3), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dv.p, vis, (size_t)n * sizeof(rcg_c128), hipMemcpyHostToDevice, st) != hipSuccess) {
      set_error(RCG_ERR_HIP, "reweight: copy in");
      return;
    }
    if (rcg_reweight(static_cast<const rcg_double3 *>(du.p), static_cast<rcg_c128 *>(dv.p),
                     scale, n, grid_size, static_cast<int *>(dc.p), st) != RCG_OK)
      return;
    if (hipMemcpyAsync(vis, dv.p, (size_t)n * sizeof(rcg_c128), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "reweight: copy out");
  });
}

extern "C" void herm_padded_inplace(rcg_c128 *data, int size, int pitch)
{
  clear_error();
  if (!data || (size & 1) || !grid_dims_ok(pitch, size)) {
    if (!data || (size & 1)) set_error(RCG_ERR_ARG, "herm_padded_inplace: null data or odd size");
    return;
  }
  host_call("herm_padded_inplace", [&](hipStream_t st) {
    const size_t bytes = (size_t)size * pitch * sizeof(rcg_c128);
    DevBuf d;
    if (d.alloc(bytes) != hipSuccess) { set_error(RCG_ERR_NOMEM, "herm_padded_inplace: device allocation"); return; }
    if (hipMemcpyAsync(d.p, data, bytes, hipMemcpyHostToDevice, st) != hipSuccess) { set_error(RCG_ERR_HIP, "herm_padded_inplace: copy in"); return; }
    if (rcg_herm_padded(static_cast<rcg_c128 *>(d.p), size, pitch, st) != RCG_OK) return;
    if (hipMemcpyAsync(data, d.p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "herm_padded_inplace: copy out");
  });
}
