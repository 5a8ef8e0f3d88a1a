// rcg_plain.hip -- the FP64 "plain" gridder and degridder.
//
// One launch covers every point; one wavefront handles one point at a time
// and its lanes cover the point's S x S window (rows of min(S,64) lanes);
// gridding adds every tap to HBM with a native FP64 atomic
// (global_atomic_add_f64), degridding reduces the window across the
// wavefront with cross-lane shuffles.  Arithmetic follows the reference's
// complex product (scatter_gridder_w_dependent_dyn_1p.cpp:169 grid,
// :171 degrid) in FP64 with contraction switched off.  This is the
// straightforward GPU gridder and the denominator the tiled FP32 path is
// measured against.
#include "rcg_internal.h"

namespace rcg {

__device__ __forceinline__ void window_lanes(int S, int lane, int *r0,
                                             int *c0, int *rows, int *cols)
{
  const int c = S < 64 ? (S > 0 ? S : 1) : 64;
  *cols = c;
  *rows = 64 / c;
  *r0 = lane / c;
  *c0 = lane - (lane / c) * c;
}

__global__ void __launch_bounds__(256)
k_grid_plain_f64(const rcg_pg *__restrict__ pg,
                 const double2 *__restrict__ vis,
                 const double2 *__restrict__ taps, double2 *__restrict__ grid,
                 int pitch, long long n)
{
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long i = wave; i < n; i += nwaves) {
    const rcg_pg r = pg[i];
    if (r.cell < 0) continue;
    const double2 v = vis[i];
    const int S = r.supp, st = r.stride;
    int r0, c0, rows, cols;
    window_lanes(S, lane, &r0, &c0, &rows, &cols);
    if (r0 >= rows) continue;
    const double sgn = (r.flags & PG_CONJ) ? -1.0 : 1.0;
    for (int su = r0; su < S; su += rows) {
      const double2 *trow = taps + r.tap0 + su * st;
      double *grow = reinterpret_cast<double *>(grid + r.cell + (long long)su * pitch);
      for (int sv = c0; sv < S; sv += cols) {
        double2 t = trow[sv];
        t.y *= sgn;
        const double pr = v.x * t.x - v.y * t.y;
        const double pi = v.x * t.y + v.y * t.x;
        unsafeAtomicAdd(grow + 2 * sv, pr);
        unsafeAtomicAdd(grow + 2 * sv + 1, pi);
      }
    }
  }
}

__global__ void __launch_bounds__(256)
k_degrid_plain_f64(const rcg_pg *__restrict__ pg,
                   const double2 *__restrict__ grid,
                   const double2 *__restrict__ taps,
                   double2 *__restrict__ vis, int pitch, long long n)
{
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long i = wave; i < n; i += nwaves) {
    const rcg_pg r = pg[i];
    double ar = 0.0, ai = 0.0;
    if (r.cell >= 0) {
      const int S = r.supp, st = r.stride;
      int r0, c0, rows, cols;
      window_lanes(S, lane, &r0, &c0, &rows, &cols);
      const double sgn = (r.flags & PG_CONJ) ? -1.0 : 1.0;
      if (r0 < rows) {
        for (int su = r0; su < S; su += rows) {
          const double2 *trow = taps + r.tap0 + su * st;
          const double2 *grow = grid + r.cell + (long long)su * pitch;
          for (int sv = c0; sv < S; sv += cols) {
            double2 t = trow[sv];
            t.y *= sgn;
            const double2 g = grow[sv];
            const double pr = g.x * t.x - g.y * t.y;
            const double pi = g.x * t.y + g.y * t.x;
            ar += pr;
            ai += pi;
          }
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        ar += __shfl_xor(ar, off);
        ai += __shfl_xor(ai, off);
      }
    }
    if (lane == 0) vis[i] = make_double2(ar, ai);
  }
}

__global__ void k_zero_f64(double *__restrict__ p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

static int plain_blocks(long long n)
{
  // four wavefronts per block, grid-stride beyond 64k blocks
  long long b = (n + 3) / 4;
  if (b > 65536) b = 65536;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace rcg

extern "C" rcg_status rcg_grid_plain_f64(const rcg_geom *geom,
                                         const rcg_gcf *gcf,
                                         const rcg_pg *pg,
                                         const rcg_c128 *vis, rcg_c128 *grid,
                                         void *stream)
{
  using namespace rcg;
  clear_error();
  if (!geom || !gcf || !pg || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_grid_plain_f64: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_geom(geom);
  if (s != RCG_OK) return s;
  hipStream_t st = (hipStream_t)stream;
  // zeroed by a kernel rather than a memset node, so the call can be
  // captured into a graph
  const long long nd = 2LL * geom->grid_size * geom->grid_pitch;
  hipLaunchKernelGGL(k_zero_f64, dim3(2048), dim3(256), 0, st,
                     reinterpret_cast<double *>(grid), nd);
  const long long n = (long long)geom->n_chunks * geom->ts_ch;
  if (n == 0) return RCG_OK;
  hipLaunchKernelGGL(k_grid_plain_f64, dim3(plain_blocks(n)), dim3(256), 0, st,
                     pg, reinterpret_cast<const double2 *>(vis),
                     reinterpret_cast<const double2 *>(gcf->data),
                     reinterpret_cast<double2 *>(grid), geom->grid_pitch, n);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

extern "C" rcg_status rcg_degrid_plain_f64(const rcg_geom *geom,
                                           const rcg_gcf *gcf,
                                           const rcg_pg *pg,
                                           const rcg_c128 *grid,
                                           rcg_c128 *vis, void *stream)
{
  using namespace rcg;
  clear_error();
  if (!geom || !gcf || !pg || !vis || !grid) {
    set_error(RCG_ERR_ARG, "rcg_degrid_plain_f64: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_geom(geom);
  if (s != RCG_OK) return s;
  const long long n = (long long)geom->n_chunks * geom->ts_ch;
  if (n == 0) return RCG_OK;
  hipLaunchKernelGGL(k_degrid_plain_f64, dim3(plain_blocks(n)), dim3(256), 0,
                     (hipStream_t)stream, pg,
                     reinterpret_cast<const double2 *>(grid),
                     reinterpret_cast<const double2 *>(gcf->data),
                     reinterpret_cast<double2 *>(vis), geom->grid_pitch, n);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}
