// rcg_fft.hip -- grid <-> image transforms over rocFFT with the reference's
// conventions (MS4/kernel/cpu/fft/fft_dyn_padded.cpp:26-127):
//
//   sign -1: complex forward (exp(-i)), sign +1: complex backward (exp(+i)),
//   sign  0: real -> half complex, sign 2: half complex -> real;
//   every transform is unnormalised and runs between two even-size fftshifts
//   (quadrant swaps) of the size x size region of a pitch-strided array.
//
// In-place layout of the reference's guru plans ({size, pitch, pitch},
// {size, 1, 1} in units of each side's element type): complex rows are
// `pitch` complex apart, real rows `pitch` reals apart, both starting at the
// buffer's first byte.  rocFFT's in-place real transforms need a different
// padding, so the real transforms go through a scratch buffer: R2C writes
// the size/2+1 spectrum columns of each row (the others keep their bytes),
// C2R writes the size x size real image.  C2C transforms run in place.
#include <map>
#include <mutex>
#include <new>

#include <rocfft/rocfft.h>

#include "rcg_internal.h"

namespace rcg {
rcg_status launch_fftshift(void *data, int elem_bytes, int size, int pitch, hipStream_t st);
}

struct rcg_fft_plan {
  int sign, f64, size, pitch;
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  void *work = nullptr;
  size_t work_bytes = 0;
  void *scratch = nullptr;  // real transforms only
  bool counted = false;     // holds a reference on the rocFFT library state
};

namespace rcg {
namespace {

// rocFFT library state: set up by the first plan, released by rcg_shutdown
// once no plan is alive (rocfft_setup / rocfft_cleanup pair).
std::mutex g_fft_mu;
bool g_fft_ready = false;
int g_fft_plans = 0;

rcg_status fft_acquire()
{
  std::lock_guard<std::mutex> lk(g_fft_mu);
  if (!g_fft_ready) {
    const rocfft_status rs = rocfft_setup();
    if (rs != rocfft_status_success) {
      set_error(RCG_ERR_FFT, "rocfft_setup failed (%d)", (int)rs);
      return RCG_ERR_FFT;
    }
    g_fft_ready = true;
  }
  g_fft_plans++;
  return RCG_OK;
}

void fft_release()
{
  std::lock_guard<std::mutex> lk(g_fft_mu);
  g_fft_plans--;
}

// real <-> scratch copies for the reference's in-place strides
template <class R>
__global__ void k_real_rows(R *__restrict__ dst, int dpitch, const R *__restrict__ src,
                            int spitch, int cols, int rows)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c < cols && r < rows) dst[(long long)r * dpitch + c] = src[(long long)r * spitch + c];
}

template <class R>
rcg_status copy_rows(void *dst, int dpitch, const void *src, int spitch, int cols, int rows,
                     hipStream_t st)
{
  hipLaunchKernelGGL(k_real_rows<R>, dim3((cols + 255) / 256, rows), dim3(256), 0, st,
                     static_cast<R *>(dst), dpitch, static_cast<const R *>(src), spitch, cols,
                     rows);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

struct c64e { float x, y; };
struct c128e { double x, y; };

}  // namespace
}  // namespace rcg

using namespace rcg;

extern "C" rcg_status rcg_fft_plan_destroy(rcg_fft_plan *p)
{
  if (!p) return RCG_OK;
  if (p->info) rocfft_execution_info_destroy(p->info);
  if (p->plan) rocfft_plan_destroy(p->plan);
  if (p->work) (void)hipFree(p->work);
  if (p->scratch) (void)hipFree(p->scratch);
  if (p->counted) fft_release();
  delete p;
  return RCG_OK;
}

extern "C" rcg_status rcg_fft_plan_create(int sign, int f64, int size, int pitch,
                                          rcg_fft_plan **out)
{
  clear_error();
  if (!out) {
    set_error(RCG_ERR_ARG, "rcg_fft_plan_create: null output");
    return RCG_ERR_ARG;
  }
  *out = nullptr;
  if (sign != -1 && sign != 0 && sign != 1 && sign != 2) {
    set_error(RCG_ERR_ARG, "sign %d not one of -1, 0, 1, 2", sign);
    return RCG_ERR_ARG;
  }
  if (size <= 0 || (size & 1) || pitch < size || (long long)size * pitch > 0x7fffffffLL) {
    set_error(RCG_ERR_ARG, "fft needs an even size <= pitch (size %d, pitch %d)", size, pitch);
    return RCG_ERR_ARG;
  }
  rcg_fft_plan *p = new (std::nothrow) rcg_fft_plan();
  if (!p) {
    set_error(RCG_ERR_NOMEM, "host allocation failed");
    return RCG_ERR_NOMEM;
  }
  if (fft_acquire() != RCG_OK) {
    delete p;
    return RCG_ERR_FFT;
  }
  p->counted = true;
  p->sign = sign;
  p->f64 = f64 ? 1 : 0;
  p->size = size;
  p->pitch = pitch;
  const rocfft_precision prec = f64 ? rocfft_precision_double : rocfft_precision_single;
  const size_t n = (size_t)size;
  rocfft_plan_description desc = nullptr;
  rocfft_status rs = rocfft_plan_description_create(&desc);
  rocfft_transform_type tt;
  rocfft_result_placement place;
  if (rs == rocfft_status_success) {
    if (sign == -1 || sign == 1) {
      tt = sign == -1 ? rocfft_transform_type_complex_forward
                      : rocfft_transform_type_complex_inverse;
      place = rocfft_placement_inplace;
      const size_t strides[2] = {1, (size_t)pitch};
      const size_t dist = (size_t)pitch * n;
      rs = rocfft_plan_description_set_data_layout(
          desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved,
          nullptr, nullptr, 2, strides, dist, 2, strides, dist);
    } else {
      // scratch is a dense [size][size/2+1] complex array
      place = rocfft_placement_notinplace;
      const size_t rstr[2] = {1, n};
      const size_t cstr[2] = {1, n / 2 + 1};
      if (sign == 0) {
        tt = rocfft_transform_type_real_forward;
        rs = rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved, nullptr,
            nullptr, 2, rstr, n * n, 2, cstr, n * (n / 2 + 1));
      } else {
        tt = rocfft_transform_type_real_inverse;
        rs = rocfft_plan_description_set_data_layout(
            desc, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, nullptr,
            nullptr, 2, cstr, n * (n / 2 + 1), 2, rstr, n * n);
      }
    }
  }
  const size_t lengths[2] = {n, n};
  if (rs == rocfft_status_success)
    rs = rocfft_plan_create(&p->plan, place, tt, prec, 2, lengths, 1, desc);
  if (desc) rocfft_plan_description_destroy(desc);
  if (rs == rocfft_status_success) rs = rocfft_execution_info_create(&p->info);
  if (rs == rocfft_status_success) rs = rocfft_plan_get_work_buffer_size(p->plan, &p->work_bytes);
  if (rs != rocfft_status_success) {
    rcg_fft_plan_destroy(p);
    set_error(RCG_ERR_FFT, "rocFFT plan creation failed (%d)", (int)rs);
    return RCG_ERR_FFT;
  }
  hipError_t e = hipSuccess;
  if (p->work_bytes) {
    e = hipMalloc(&p->work, p->work_bytes);
    if (e == hipSuccess)
      rs = rocfft_execution_info_set_work_buffer(p->info, p->work, p->work_bytes);
  }
  if (e == hipSuccess && rs == rocfft_status_success && (sign == 0 || sign == 2)) {
    // real scratch [size][size] and complex scratch [size][size/2+1]
    const size_t elem = f64 ? 8 : 4;
    const size_t bytes = n * n * elem + n * (n / 2 + 1) * 2 * elem;
    e = hipMalloc(&p->scratch, bytes);
  }
  if (e != hipSuccess || rs != rocfft_status_success) {
    rcg_fft_plan_destroy(p);
    set_error(e != hipSuccess ? RCG_ERR_NOMEM : RCG_ERR_FFT, "rocFFT buffers: %s",
              e != hipSuccess ? hipGetErrorString(e) : "work buffer");
    return e != hipSuccess ? RCG_ERR_NOMEM : RCG_ERR_FFT;
  }
  *out = p;
  return RCG_OK;
}

// Executes the plan in place on device memory `data` (the reference's
// layout, see above), on `stream`.
extern "C" rcg_status rcg_fft_execute(rcg_fft_plan *p, void *data, void *stream)
{
  clear_error();
  if (!p || !data) {
    set_error(RCG_ERR_ARG, "rcg_fft_execute: null argument");
    return RCG_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int n = p->size, pitch = p->pitch;
  const int re = p->f64 ? 8 : 4, ce = 2 * re;
  rocfft_status rs = rocfft_execution_info_set_stream(p->info, st);
  if (rs != rocfft_status_success) {
    set_error(RCG_ERR_FFT, "rocfft_execution_info_set_stream (%d)", (int)rs);
    return RCG_ERR_FFT;
  }
  rcg_status s;
  if (p->sign == -1 || p->sign == 1) {
    if ((s = launch_fftshift(data, ce, n, pitch, st)) != RCG_OK) return s;
    void *bufs[1] = {data};
    rs = rocfft_execute(p->plan, bufs, nullptr, p->info);
    if (rs != rocfft_status_success) {
      set_error(RCG_ERR_FFT, "rocfft_execute (%d)", (int)rs);
      return RCG_ERR_FFT;
    }
    return launch_fftshift(data, ce, n, pitch, st);
  }
  char *rscr = static_cast<char *>(p->scratch);
  char *cscr = rscr + (size_t)n * n * re;
  const int hc = n / 2 + 1;
  if (p->sign == 0) {
    // real rows of `pitch` reals -> dense real scratch -> half spectrum
    if ((s = launch_fftshift(data, re, n, pitch, st)) != RCG_OK) return s;
    s = p->f64 ? copy_rows<double>(rscr, n, data, pitch, n, n, st)
               : copy_rows<float>(rscr, n, data, pitch, n, n, st);
    if (s != RCG_OK) return s;
    void *in[1] = {rscr}, *outb[1] = {cscr};
    rs = rocfft_execute(p->plan, in, outb, p->info);
    if (rs != rocfft_status_success) {
      set_error(RCG_ERR_FFT, "rocfft_execute (%d)", (int)rs);
      return RCG_ERR_FFT;
    }
    s = p->f64 ? copy_rows<c128e>(data, pitch, cscr, hc, hc, n, st)
               : copy_rows<c64e>(data, pitch, cscr, hc, hc, n, st);
    if (s != RCG_OK) return s;
    return launch_fftshift(data, ce, n, pitch, st);
  }
  // sign 2: complex rows of `pitch` -> half spectrum scratch -> real image
  if ((s = launch_fftshift(data, ce, n, pitch, st)) != RCG_OK) return s;
  s = p->f64 ? copy_rows<c128e>(cscr, hc, data, pitch, hc, n, st)
             : copy_rows<c64e>(cscr, hc, data, pitch, hc, n, st);
  if (s != RCG_OK) return s;
  void *in[1] = {cscr}, *outb[1] = {rscr};
  rs = rocfft_execute(p->plan, in, outb, p->info);
  if (rs != rocfft_status_success) {
    set_error(RCG_ERR_FFT, "rocfft_execute (%d)", (int)rs);
    return RCG_ERR_FFT;
  }
  s = p->f64 ? copy_rows<double>(data, pitch, rscr, n, n, n, st)
             : copy_rows<float>(data, pitch, rscr, n, n, n, st);
  if (s != RCG_OK) return s;
  return launch_fftshift(data, re, n, pitch, st);
}

extern "C" rcg_status rcg_shutdown(void)
{
  clear_error();
  std::lock_guard<std::mutex> lk(g_fft_mu);
  if (g_fft_plans > 0) {
    set_error(RCG_ERR_ARG, "rcg_shutdown: %d FFT plans still alive", g_fft_plans);
    return RCG_ERR_ARG;
  }
  if (g_fft_ready) {
    (void)rocfft_cleanup();
    g_fft_ready = false;
  }
  return RCG_OK;
}

// ------------------------------------------------ reference host entry points --

extern "C" void fftInitThreading(void) {}

extern "C" void rcg_fft_destroy_plan(void *plan)
{
  rcg_fft_plan_destroy(static_cast<rcg_fft_plan *>(plan));
}

extern "C" void *fft_inplace_even(void *plan, int sign, void *data, int size, int pitch)
{
  clear_error();
  if (!data) {
    set_error(RCG_ERR_ARG, "fft_inplace_even: null data");
    return plan;
  }
  rcg_fft_plan *p = static_cast<rcg_fft_plan *>(plan);
  if (p && (p->sign != sign || p->size != size || p->pitch != pitch || !p->f64)) {
    set_error(RCG_ERR_ARG, "fft_inplace_even: plan made for another transform");
    return plan;
  }
  if (!p && rcg_fft_plan_create(sign, 1, size, pitch, &p) != RCG_OK) return nullptr;
  // the caller's buffer holds size rows of pitch complex (MS4 UV grids)
  const size_t bytes = (size_t)size * pitch * sizeof(rcg_c128);
  hipStream_t st = nullptr;
  void *d = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&d, bytes) != hipSuccess) {
    set_error(RCG_ERR_NOMEM, "fft_inplace_even: device buffers");
    if (st) (void)hipStreamDestroy(st);
    return p;
  }
  if (hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, st) == hipSuccess &&
      rcg_fft_execute(p, d, st) == RCG_OK) {
    if (hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      set_error(RCG_ERR_HIP, "fft_inplace_even: copy out");
  } else if (!rcg_last_error()) {
    set_error(RCG_ERR_HIP, "fft_inplace_even: copy in");
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(d);
  (void)hipStreamDestroy(st);
  return p;
}
