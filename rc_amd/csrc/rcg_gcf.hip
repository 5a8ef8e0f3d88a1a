// rcg_gcf.hip -- w-projection GCF (gridding convolution function) taps on
// the GPU.
//
// Restates the reference's generator, MS4/kernel/cpu/gcf/GCF.cpp:16-219,
// and the per-plane loop of its caller, MS4/dna-programs/Kernel/CPU/GCF.hs:
// 36-89:
//
//  calcAccums   per w-plane sum of w and point count, plane = round(w/wstep)
//               (C rounding, half away from zero), planes centred at
//               n_planes / 2 (GCF.cpp:148-219);
//  mkGCFLayer   one plane's 64 oversampled layers: the w screen
//               exp(-2 pi i w (1 - sqrt(1 - x^2 - y^2))) on a max_support^2
//               block at the centre of an (8 max_support)^2 arena (row
//               pitch 8 max_support + src_pad), zero elsewhere; the backward
//               transform with fftshifts (fft_inplace_even, sign +1); the
//               sum of real parts over each of the 8 x 8 oversampling
//               sub-lattices; the support^2 window of every sub-lattice
//               divided by its sum -> dst[over_u][over_v][support][support]
//               (GCF.cpp:16-120).
//
// The reference's sums walk the padded arena as if it were contiguous
// (GCF.cpp:28-32), which reads the wrong elements when src_pad != 0 (the MS4
// Haskell caller uses src_pad = 4).  The default here is the pad-correct
// sum; `pad_quirk` selects the reference's walk.
//
// GPU structure: one kernel writes the screen and zeroes the rest of the
// arena (FP64 phase and sincos, contraction off); rocFFT does the
// transform (rcg_fft plan, FP64, the reference's shift convention); 64
// workgroups reduce the sub-lattice sums; one kernel divides and scatters
// the window into the [8][8][S][S] layout (FP64 or complex64 output).
// calcAccums uses per-workgroup LDS histograms and one FP64 atomic per
// touched plane per workgroup.
#include <algorithm>
#include <climits>
#include <cmath>

#include "rcg_internal.h"

constexpr int GCF_OVER = 8;  // MS4/kernel/common/metrix.h:5
constexpr int ACC_MAX_PLANES = 8192;

struct rcg_gcf_gen {
  int max_support, src_pad, L, pitch;
  double2 *arena = nullptr;  // L rows x pitch complex
  double *sums = nullptr;    // 64 sub-lattice sums
  rcg_fft_plan *fft = nullptr;
};

namespace rcg {
namespace {

__global__ void __launch_bounds__(256)
k_wscreen(double2 *__restrict__ arena, int L, int pitch, int ms, double t2, double w)
{
#pragma clang fp contract(off)
  const int radius = ms / 2;
  const double normer = t2 / (double)radius;
  const int r0 = L / 2 - radius;
  const long long n = (long long)L * pitch;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(k / pitch), c = (int)(k - (long long)r * pitch);
    double2 v = make_double2(0.0, 0.0);
    const int a = r - r0, b = c - r0;
    if (a >= 0 && a < ms && b >= 0 && b < ms) {
      const double x = (double)(a - radius) * normer;
      const double y = (double)(b - radius) * normer;
      const double ph = w * (1.0 - sqrt(1.0 - x * x - y * y));
      double s, co;
      sincos(2.0 * M_PI * ph, &s, &co);
      v = make_double2(co, -s);
    }
    arena[k] = v;
  }
}

// Sum of real parts over sub-lattice (over_u, over_v) = blockIdx.x: arena
// cells (su*8 + over_u, sv*8 + over_v), su, sv < ms; with pad_quirk the
// element at contiguous offset ((su*8 + over_u)*ms + sv)*8 + over_v.
__global__ void __launch_bounds__(256)
k_gcf_sums(const double2 *__restrict__ arena, int pitch, int ms, int pad_quirk,
           double *__restrict__ sums)
{
  __shared__ double sh[4];
  const int ou = blockIdx.x / GCF_OVER, ov = blockIdx.x % GCF_OVER;
  double acc = 0.0;
  for (int e = threadIdx.x; e < ms * ms; e += 256) {
    const int su = e / ms, sv = e - (e / ms) * ms;
    const long long k = pad_quirk
        ? ((long long)(su * GCF_OVER + ou) * ms + sv) * GCF_OVER + ov
        : (long long)(su * GCF_OVER + ou) * pitch + sv * GCF_OVER + ov;
    acc += arena[k].x;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

template <class T>
__global__ void __launch_bounds__(256)
k_gcf_extract(const double2 *__restrict__ arena, int pitch, int ms, int support,
              const double *__restrict__ sums, T *__restrict__ dst)
{
  const int start = (ms / 2 - support / 2) * GCF_OVER;
  const long long per = (long long)support * support;
  const long long n = 64 * per;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const int layer = (int)(k / per);
    const int e = (int)(k - layer * per);
    const int ou = layer / GCF_OVER, ov = layer % GCF_OVER;
    const int su = e / support, sv = e - (e / support) * support;
    const double2 a = arena[(long long)(start + ou + su * GCF_OVER) * pitch + start + ov +
                            sv * GCF_OVER];
    const double s = sums[layer];
    T o;
    o.x = a.x / s;
    o.y = a.y / s;
    dst[k] = o;
  }
}

// C round() of w / wstep + centre; LDS histogram per workgroup.
__global__ void __launch_bounds__(256)
k_accums(const rcg_double3 *__restrict__ uvw, long long n, double wstep, int n_planes,
         double *__restrict__ sums, int *__restrict__ npts, int *__restrict__ err)
{
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  double *hs = lds;
  int *hn = reinterpret_cast<int *>(lds + n_planes);
  for (int p = threadIdx.x; p < n_planes; p += 256) {
    hs[p] = 0.0;
    hn[p] = 0;
  }
  __syncthreads();
  const int centre = n_planes / 2;
  int bad = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double w = uvw[i].w;
    const double q = round(w / wstep);
    if (!(q >= -centre && q < n_planes - centre)) {
      bad = 1;
      continue;
    }
    const int p = (int)q + centre;
    atomicAdd(&hs[p], w);
    atomicAdd(&hn[p], 1);
  }
  if (bad) atomicOr(err, ERR_RANGE);
  __syncthreads();
  for (int p = threadIdx.x; p < n_planes; p += 256) {
    if (hn[p]) {
      unsafeAtomicAdd(&sums[p], hs[p]);
      atomicAdd(&npts[p], hn[p]);
    }
  }
}

__global__ void k_zero_acc(double *__restrict__ sums, int *__restrict__ npts, int *err, int n)
{
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    sums[p] = 0.0;
    npts[p] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && err) *err = 0;
}

}  // namespace
}  // namespace rcg

using namespace rcg;

extern "C" rcg_status rcg_gcf_gen_destroy(rcg_gcf_gen *g)
{
  if (!g) return RCG_OK;
  if (g->fft) rcg_fft_plan_destroy(g->fft);
  if (g->arena) (void)hipFree(g->arena);
  if (g->sums) (void)hipFree(g->sums);
  delete g;
  return RCG_OK;
}

extern "C" rcg_status rcg_gcf_gen_create(int max_support, int src_pad, rcg_gcf_gen **out)
{
  clear_error();
  if (!out || max_support < 2 || max_support > MAX_SUPPORT || src_pad < 0 || src_pad > 4096) {
    set_error(RCG_ERR_ARG, "rcg_gcf_gen_create: max_support %d / src_pad %d out of range",
              max_support, src_pad);
    return RCG_ERR_ARG;
  }
  *out = nullptr;
  rcg_gcf_gen *g = new (std::nothrow) rcg_gcf_gen();
  if (!g) {
    set_error(RCG_ERR_NOMEM, "host allocation failed");
    return RCG_ERR_NOMEM;
  }
  g->max_support = max_support;
  g->src_pad = src_pad;
  g->L = max_support * GCF_OVER;
  g->pitch = g->L + src_pad;
  const size_t bytes = (size_t)g->L * g->pitch * sizeof(double2);
  if (hipMalloc(reinterpret_cast<void **>(&g->arena), bytes) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&g->sums), 64 * sizeof(double)) != hipSuccess) {
    g->arena = nullptr;
    rcg_gcf_gen_destroy(g);
    set_error(RCG_ERR_NOMEM, "rcg_gcf_gen_create: device arena of %zu bytes", bytes);
    return RCG_ERR_NOMEM;
  }
  rcg_status s = rcg_fft_plan_create(1, 1, g->L, g->pitch, &g->fft);
  if (s != RCG_OK) {
    g->fft = nullptr;
    rcg_gcf_gen_destroy(g);
    return s;
  }
  *out = g;
  return RCG_OK;
}

extern "C" rcg_status rcg_gcf_gen_layer(rcg_gcf_gen *g, void *dst, int f64, int support,
                                        double t2, double w, int pad_quirk, void *stream)
{
  clear_error();
  if (!g || !dst || support < 1 || support > g->max_support) {
    set_error(RCG_ERR_ARG, "rcg_gcf_gen_layer: support %d outside [1, %d]", support,
              g ? g->max_support : 0);
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const long long cells = (long long)g->L * g->pitch;
  const int nb = (int)std::min<long long>((cells + 255) / 256, 2048);
  hipLaunchKernelGGL(k_wscreen, dim3(nb), dim3(256), 0, st, g->arena, g->L, g->pitch,
                     g->max_support, t2, w);
  RCG_HIP_TRY(hipGetLastError());
  rcg_status s = rcg_fft_execute(g->fft, g->arena, stream);
  if (s != RCG_OK) return s;
  hipLaunchKernelGGL(k_gcf_sums, dim3(64), dim3(256), 0, st, g->arena, g->pitch,
                     g->max_support, pad_quirk ? 1 : 0, g->sums);
  const long long n = 64LL * support * support;
  const int nb2 = (int)std::min<long long>((n + 255) / 256, 1024);
  if (f64)
    hipLaunchKernelGGL(k_gcf_extract<double2>, dim3(nb2), dim3(256), 0, st, g->arena,
                       g->pitch, g->max_support, support, g->sums,
                       static_cast<double2 *>(dst));
  else
    hipLaunchKernelGGL(k_gcf_extract<float2>, dim3(nb2), dim3(256), 0, st, g->arena,
                       g->pitch, g->max_support, support, g->sums,
                       static_cast<float2 *>(dst));
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_calc_accums(const rcg_double3 *uvw, long long n, double wstep,
                                      int n_planes, double *sums, int *npts, int *err_dev,
                                      void *stream)
{
  clear_error();
  if ((n > 0 && !uvw) || !sums || !npts || !err_dev || n < 0 || n_planes < 1 ||
      n_planes > ACC_MAX_PLANES || !(wstep > 0.0)) {
    set_error(RCG_ERR_ARG, "rcg_calc_accums: bad argument (n %lld, planes %d, wstep %g)", n,
              n_planes, wstep);
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_acc, dim3(std::max(1, (n_planes + 255) / 256)), dim3(256), 0, st,
                     sums, npts, err_dev, n_planes);
  RCG_HIP_TRY(hipGetLastError());
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>((n + 255) / 256, 1024);
  const size_t lds = (size_t)n_planes * (sizeof(double) + sizeof(int));
  if (lds > 64 * 1024) {
    RCG_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_accums),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  }
  hipLaunchKernelGGL(k_accums, dim3(nb), dim3(256), lds, st, uvw, n, wstep, n_planes, sums,
                     npts, err_dev);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

// ---------------------------------------------------------------------------
// The reference's host entry points (host memory; GCF.cpp:123-146, 148-152)

namespace {
thread_local bool t_pad_quirk = false;
}

extern "C" void rcg_gcf_pad_quirk(int enable) { t_pad_quirk = enable != 0; }

extern "C" void *mkGCFLayer(void *p, rcg_c128 dst[], rcg_c128 *table[], rcg_c128 arena[],
                            int support, int max_support, int src_pad, double t2, double w)
{
  (void)arena;  // the work buffer lives on the device
  clear_error();
  if (!dst || !table) {
    set_error(RCG_ERR_ARG, "mkGCFLayer: null argument");
    return p;
  }
  rcg_gcf_gen *g = static_cast<rcg_gcf_gen *>(p);
  if (g && (g->max_support != max_support || g->src_pad != src_pad)) {
    set_error(RCG_ERR_ARG, "mkGCFLayer: generator made for max_support %d / pad %d",
              g->max_support, g->src_pad);
    return p;
  }
  if (!g && rcg_gcf_gen_create(max_support, src_pad, &g) != RCG_OK) return nullptr;
  const size_t bytes = 64ull * support * support * sizeof(rcg_c128);
  hipStream_t st = nullptr;
  void *d = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&d, bytes) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "mkGCFLayer: device buffers");
  } else if (rcg_gcf_gen_layer(g, d, 1, support, t2, w, t_pad_quirk, st) == RCG_OK) {
    if (hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "mkGCFLayer: copy out");
    for (int i = 0; i < 64; i++) table[i] = dst + (size_t)i * support * support;
  }
  if (st) (void)hipStreamSynchronize(st);
  if (d) (void)hipFree(d);
  if (st) (void)hipStreamDestroy(st);
  return g;
}

extern "C" void calcAccums(const rcg_double3 uvw[], double sums[], int npts[], double wstep,
                           int numDataPoints, int numOfWPlanes)
{
  clear_error();
  if (!uvw || !sums || !npts || numDataPoints < 0 || numOfWPlanes < 1) {
    set_error(RCG_ERR_ARG, "calcAccums: bad argument");
    return;
  }
  hipStream_t st = nullptr;
  void *du = nullptr, *ds = nullptr, *dn = nullptr, *de = nullptr;
  const size_t ub = (size_t)numDataPoints * sizeof(rcg_double3);
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&du, ub ? ub : 1) != hipSuccess ||
      hipMalloc(&ds, numOfWPlanes * sizeof(double)) != hipSuccess ||
      hipMalloc(&dn, numOfWPlanes * sizeof(int)) != hipSuccess ||
      hipMalloc(&de, sizeof(int)) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "calcAccums: device buffers");
  } else if ((ub && hipMemcpyAsync(du, uvw, ub, hipMemcpyHostToDevice, st) != hipSuccess)) {
    set_error(RCG_ERR_HIP, "calcAccums: copy in");
  } else if (rcg_calc_accums(static_cast<rcg_double3 *>(du), numDataPoints, wstep, numOfWPlanes,
                             static_cast<double *>(ds), static_cast<int *>(dn),
                             static_cast<int *>(de), st) == RCG_OK) {
    int e = 0;
    if (hipMemcpyAsync(sums, ds, numOfWPlanes * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(npts, dn, numOfWPlanes * sizeof(int), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(&e, de, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "calcAccums: copy out");
    else if (e)
      set_error(RCG_ERR_RANGE, "calcAccums: points with a w-plane outside the %d planes "
                "(skipped)", numOfWPlanes);
  }
  if (st) (void)hipStreamSynchronize(st);
  void *bufs[] = {du, ds, dn, de};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (st) (void)hipStreamDestroy(st);
}
