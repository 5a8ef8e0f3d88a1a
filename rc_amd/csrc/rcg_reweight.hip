// rcg_reweight.hip -- uniform weighting of visibilities on the device: the
// number of points per rounded grid cell is counted, then every in-grid
// visibility is divided by its cell's count (reference
// MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p.cpp:311-353).
//
// Two entry points share the kernels:
//  * rcg_reweight counts into a buffer of grid_size^2 ints that the caller
//    passes;
//  * rcg_reweight_static counts into `scratch`, an array of this file that
//    every device holds for the life of the process, so that the call
//    allocates nothing (for callers that cannot: code under graph capture,
//    serving loops).  Grids up to RCG_REWEIGHT_STATIC_MAX_SIZE square.  The
//    counts are cleared by a kernel at every call; two calls that overlap on
//    one device share the array and race.
#include <algorithm>
#include <mutex>

#include "rcg_internal.h"

namespace rcg {
namespace {

constexpr long long STATIC_CELLS =
    (long long)RCG_REWEIGHT_STATIC_MAX_SIZE * RCG_REWEIGHT_STATIC_MAX_SIZE;

// Count buffer of rcg_reweight_static (64 MiB per device).  In host code the
// name stands for a host-side handle, not the device address: kernels get
// the address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) int scratch[STATIC_CELLS];

__global__ void k_zero_counts(int *__restrict__ p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Cell of a point relative to the grid's first cell: round(u scale) rows
// and round(v scale) columns from the centre (size / 2, size / 2).
__device__ __forceinline__ long long reweight_cell(const rcg_double3 &p, double scale,
                                                   int size, long long trans)
{
#pragma clang fp contract(off)
  const long long u = (long long)round(p.u * scale);
  const long long v = (long long)round(p.v * scale);
  return u * size + v + trans;
}

__global__ void k_reweight_count(const rcg_double3 *__restrict__ uvw, int n, double scale,
                                 int size, long long last, long long trans,
                                 int *__restrict__ count)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long kg = reweight_cell(uvw[i], scale, size, trans);
  if (kg >= 0 && kg < last) atomicAdd(count + kg, 1);
}

__global__ void k_reweight_apply(const rcg_double3 *__restrict__ uvw, double2 *__restrict__ vis,
                                 int n, double scale, int size, long long last,
                                 long long trans, const int *__restrict__ count)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long kg = reweight_cell(uvw[i], scale, size, trans);
  if (kg >= 0 && kg < last) {
    const double c = (double)count[kg];
    vis[i].x /= c;
    vis[i].y /= c;
  }
}

rcg_status check_args(const char *who, const rcg_double3 *uvw, const rcg_c128 *vis, int n,
                      int grid_size)
{
  if (grid_size <= 0 || (long long)grid_size * grid_size > 0x7fffffffLL || n < 0 ||
      (n > 0 && (!uvw || !vis))) {
    set_error(RCG_ERR_ARG, "%s: bad arguments", who);
    return RCG_ERR_ARG;
  }
  return RCG_OK;
}

rcg_status launch(const rcg_double3 *uvw, rcg_c128 *vis, double scale, int n, int grid_size,
                  int *count, hipStream_t st)
{
  const long long last = (long long)grid_size * grid_size;
  const long long trans = (long long)(grid_size / 2) * (grid_size + 1);
  const long long zb = std::min((last + 255) / 256, 65536LL);
  hipLaunchKernelGGL(k_zero_counts, dim3((unsigned)zb), dim3(256), 0, st, count, last);
  hipLaunchKernelGGL(k_reweight_count, dim3((n + 255) / 256), dim3(256), 0, st, uvw, n, scale,
                     grid_size, last, trans, count);
  hipLaunchKernelGGL(k_reweight_apply, dim3((n + 255) / 256), dim3(256), 0, st, uvw,
                     reinterpret_cast<double2 *>(vis), n, scale, grid_size, last, trans, count);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

// Device address of `scratch` on each device, looked up at the first call
// on that device (with it current) and kept.
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
int *g_scratch[MAX_DEVICES] = {};

rcg_status static_counts(int **out)
{
  int dev = 0;
  RCG_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) {
    set_error(RCG_ERR_ARG, "rcg_reweight_static: device %d beyond %d", dev, MAX_DEVICES);
    return RCG_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(RCG_ERR_HIP, "rcg_reweight_static: hipGetSymbolAddress of the count buffer "
                "failed on device %d: %s", dev, hipGetErrorString(e));
      return RCG_ERR_HIP;
    }
    g_scratch[dev] = static_cast<int *>(p);
  }
  *out = g_scratch[dev];
  return RCG_OK;
}

}  // namespace
}  // namespace rcg

using namespace rcg;

extern "C" rcg_status rcg_reweight(const rcg_double3 *uvw, rcg_c128 *vis, double scale, int n,
                                   int grid_size, int *count_scratch, void *stream)
{
  clear_error();
  rcg_status s = check_args("rcg_reweight", uvw, vis, n, grid_size);
  if (s != RCG_OK) return s;
  if (n > 0 && !count_scratch) {
    set_error(RCG_ERR_ARG, "rcg_reweight: null count buffer");
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  return launch(uvw, vis, scale, n, grid_size, count_scratch, (hipStream_t)stream);
}

extern "C" rcg_status rcg_reweight_static(const rcg_double3 *uvw, rcg_c128 *vis, double scale,
                                          int n, int grid_size, void *stream)
{
  clear_error();
  rcg_status s = check_args("rcg_reweight_static", uvw, vis, n, grid_size);
  if (s != RCG_OK) return s;
  if (grid_size > RCG_REWEIGHT_STATIC_MAX_SIZE) {
    set_error(RCG_ERR_ARG, "rcg_reweight_static: grid_size %d above the %d of its count buffer "
              "(use rcg_reweight with a buffer of grid_size^2 ints)", grid_size,
              RCG_REWEIGHT_STATIC_MAX_SIZE);
    return RCG_ERR_ARG;
  }
  if (n == 0) return RCG_OK;
  int *count = nullptr;
  s = static_counts(&count);
  if (s != RCG_OK) return s;
  return launch(uvw, vis, scale, n, grid_size, count, (hipStream_t)stream);
}
