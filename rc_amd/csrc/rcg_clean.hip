// rcg_clean.hip -- Hogbom CLEAN minor cycle on the GPU.
//
// Interface of the reference's deconvolve (MS4/kernel/common/hogbom.h:7-16,
// body MS4/kernel/cpu/hogbom/hogbom.cpp:94-121) with textbook semantics:
//   model = 0; p = argmax |psf|, pv = psf[p]
//   repeat niters: r = argmax |res| (first index on ties), v = res[r];
//     stop if |v| < threshold; a = gain * v / pv;
//     res[x] -= a * psf[x - r + p] where x - r + p lies in the image;
//     model[r] += a.
// (The reference's own findPeak reads past the array and returns |x|, so its
// results are not reproduced; see the oracle.)
//
// GPU structure: every iteration is one launch that subtracts the previous
// peak's shifted PSF and forms per-block maxima of the updated residual
// (wavefront reductions on (|value|, -index)), and one single-block launch
// that reduces the block maxima, applies the threshold and updates the
// model.  Peak, PSF peak, a stop flag and the iteration count live in device
// memory, so no host synchronisation is needed between iterations.
#include <cfloat>

#include "rcg_internal.h"

namespace rcg {
namespace {

constexpr int CB = 512;   // blocks of the partial-maximum pass
constexpr int CT = 256;   // threads per block

struct CleanState {
  long long pos;    // residual peak (row*pitch + col)
  long long ppos;   // PSF peak
  double val;       // signed residual peak value
  double pval;      // signed PSF peak value
  int done;
  int iters;
};

struct Best {
  double a;        // |value|
  long long idx;   // row-major index row*siz + col
};

__device__ __forceinline__ Best better(Best x, Best y)
{
  if (y.a > x.a || (y.a == x.a && y.idx < x.idx)) return y;
  return x;
}

__device__ __forceinline__ Best wave_best(Best b)
{
  for (int off = 32; off > 0; off >>= 1) {
    Best o;
    o.a = __shfl_xor(b.a, off);
    o.idx = __shfl_xor(b.idx, off);
    b = better(b, o);
  }
  return b;
}

__device__ __forceinline__ Best block_best(Best b, Best *sh)
{
  b = wave_best(b);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) b = better(b, sh[w]);
    sh[0] = b;
  }
  __syncthreads();
  return sh[0];
}

// Optional subtraction of the current peak, then partial maxima.
template <class T>
__global__ void __launch_bounds__(CT)
k_sub_argmax(T *__restrict__ img, const T *__restrict__ psf, int siz, int pitch,
             CleanState *__restrict__ st, double gain, int subtract, int find,
             double *__restrict__ pa, long long *__restrict__ pi)
{
#pragma clang fp contract(off)
  __shared__ Best sh[CT / 64];
  const bool act = !st->done;
  long long dr = 0, dc = 0;
  T a = 0;
  if (subtract && act) {
    const long long rr = st->pos / pitch, rc = st->pos % pitch;
    const long long pr = st->ppos / pitch, pc = st->ppos % pitch;
    dr = pr - rr;
    dc = pc - rc;
    a = (T)(gain * st->val / st->pval);
  }
  Best b{-1.0, 0x7fffffffffffffffLL};
  for (int r = blockIdx.x; r < siz; r += gridDim.x) {
    T *row = img + (long long)r * pitch;
    const long long sr = r + dr;
    const bool in_r = sr >= 0 && sr < siz;
    for (int c = threadIdx.x; c < siz; c += CT) {
      T x = row[c];
      if (subtract && act && in_r) {
        const long long sc = c + dc;
        if (sc >= 0 && sc < siz) {
          x -= a * psf[sr * pitch + sc];
          row[c] = x;
        }
      }
      if (find) b = better(b, Best{(double)fabs(x), (long long)r * siz + c});
    }
  }
  if (find) {
    b = block_best(b, sh);
    if (threadIdx.x == 0) {
      pa[blockIdx.x] = b.a;
      pi[blockIdx.x] = b.idx;
    }
  }
}

// mode 0: PSF peak; mode 1: residual peak of iteration `it`
template <class T>
__global__ void __launch_bounds__(CT)
k_clean_final(const T *__restrict__ img, T *__restrict__ model, int siz, int pitch,
              CleanState *__restrict__ st, const double *__restrict__ pa,
              const long long *__restrict__ pi, int nb, int mode, int it, double gain,
              double threshold)
{
#pragma clang fp contract(off)
  __shared__ Best sh[CT / 64];
  if (mode == 1 && st->done) return;
  Best b{-1.0, 0x7fffffffffffffffLL};
  for (int k = threadIdx.x; k < nb; k += CT) b = better(b, Best{pa[k], pi[k]});
  b = block_best(b, sh);
  if (threadIdx.x != 0) return;
  const long long r = b.idx / siz, c = b.idx % siz;
  const long long pos = r * pitch + c;
  const double v = (double)img[pos];
  if (mode == 0) {
    st->ppos = pos;
    st->pval = v;
    st->done = 0;
    st->iters = 0;
    return;
  }
  st->pos = pos;
  st->val = v;
  if (fabs(v) < threshold || st->pval == 0.0) {
    st->done = 1;
    return;
  }
  model[pos] += (T)(gain * v / st->pval);
  st->iters = it + 1;
}

template <class T>
__global__ void k_zero(T *__restrict__ p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0;
}

template <class T>
rcg_status hogbom(T *model, T *res, const T *psf, int siz, int pitch, unsigned niters,
                  double gain, double threshold, void *scratch, hipStream_t st)
{
  CleanState *cs = static_cast<CleanState *>(scratch);
  double *pa = reinterpret_cast<double *>(cs + 1);
  long long *pi = reinterpret_cast<long long *>(pa + CB);
  const long long n = (long long)siz * pitch;
  hipLaunchKernelGGL(k_zero<T>, dim3(1024), dim3(256), 0, st, model, n);
  // PSF peak
  hipLaunchKernelGGL(k_sub_argmax<T>, dim3(CB), dim3(CT), 0, st, const_cast<T *>(psf), psf,
                     siz, pitch, cs, gain, 0, 1, pa, pi);
  hipLaunchKernelGGL(k_clean_final<T>, dim3(1), dim3(CT), 0, st, psf, model, siz, pitch, cs,
                     pa, pi, CB, 0, 0, gain, threshold);
  if (niters == 0) {
    RCG_HIP_TRY(hipGetLastError());
    return RCG_OK;
  }
  // first residual peak
  hipLaunchKernelGGL(k_sub_argmax<T>, dim3(CB), dim3(CT), 0, st, res, psf, siz, pitch, cs,
                     gain, 0, 1, pa, pi);
  hipLaunchKernelGGL(k_clean_final<T>, dim3(1), dim3(CT), 0, st, res, model, siz, pitch, cs,
                     pa, pi, CB, 1, 0, gain, threshold);
  for (unsigned it = 1; it < niters; it++) {
    hipLaunchKernelGGL(k_sub_argmax<T>, dim3(CB), dim3(CT), 0, st, res, psf, siz, pitch, cs,
                       gain, 1, 1, pa, pi);
    hipLaunchKernelGGL(k_clean_final<T>, dim3(1), dim3(CT), 0, st, res, model, siz, pitch, cs,
                       pa, pi, CB, 1, (int)it, gain, threshold);
  }
  // subtraction of the last accepted peak
  hipLaunchKernelGGL(k_sub_argmax<T>, dim3(CB), dim3(CT), 0, st, res, psf, siz, pitch, cs,
                     gain, 1, 0, pa, pi);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

}  // namespace
}  // namespace rcg

using namespace rcg;

extern "C" size_t rcg_hogbom_scratch_bytes(void)
{
  return sizeof(CleanState) + CB * (sizeof(double) + sizeof(long long));
}

extern "C" rcg_status rcg_hogbom(void *model, void *res, const void *psf, int f64, int siz,
                                 int pitch, unsigned niters, double gain, double threshold,
                                 void *scratch, void *stream)
{
  clear_error();
  if (!model || !res || !psf || !scratch || siz <= 0 || pitch < siz ||
      (long long)siz * siz > 0x7fffffffffffLL) {
    set_error(RCG_ERR_ARG, "rcg_hogbom: bad arguments");
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  if (f64)
    return hogbom<double>(static_cast<double *>(model), static_cast<double *>(res),
                          static_cast<const double *>(psf), siz, pitch, niters, gain,
                          threshold, scratch, st);
  return hogbom<float>(static_cast<float *>(model), static_cast<float *>(res),
                       static_cast<const float *>(psf), siz, pitch, niters, gain, threshold,
                       scratch, st);
}

// Iterations the last rcg_hogbom on this scratch performed (synchronises).
extern "C" rcg_status rcg_hogbom_iterations(const void *scratch, unsigned *iters, void *stream)
{
  clear_error();
  if (!scratch || !iters) {
    set_error(RCG_ERR_ARG, "rcg_hogbom_iterations: null argument");
    return RCG_ERR_ARG;
  }
  CleanState cs;
  RCG_HIP_TRY(hipMemcpyAsync(&cs, scratch, sizeof(cs), hipMemcpyDeviceToHost,
                             (hipStream_t)stream));
  RCG_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  *iters = (unsigned)cs.iters;
  return RCG_OK;
}

extern "C" void deconvolve(double *mod_p, double *res_p, double *psf_p, int siz, int pitch,
                           unsigned int niters, double gain, double threshold)
{
  clear_error();
  if (!mod_p || !res_p || !psf_p || siz <= 0 || pitch < siz) {
    set_error(RCG_ERR_ARG, "deconvolve: bad arguments");
    return;
  }
  const size_t bytes = (size_t)siz * pitch * sizeof(double);
  hipStream_t st = nullptr;
  void *dm = nullptr, *dr = nullptr, *dp = nullptr, *ds = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&dm, bytes) != hipSuccess || hipMalloc(&dr, bytes) != hipSuccess ||
      hipMalloc(&dp, bytes) != hipSuccess ||
      hipMalloc(&ds, rcg_hogbom_scratch_bytes()) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "deconvolve: device buffers");
  } else if (hipMemcpyAsync(dr, res_p, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(dp, psf_p, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error(RCG_ERR_HIP, "deconvolve: copy in");
  } else if (rcg_hogbom(dm, dr, dp, 1, siz, pitch, niters, gain, threshold, ds, st) == RCG_OK) {
    if (hipMemcpyAsync(mod_p, dm, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(res_p, dr, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "deconvolve: copy out");
  }
  if (st) (void)hipStreamSynchronize(st);
  void *bufs[] = {dm, dr, dp, ds};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (st) (void)hipStreamDestroy(st);
}
