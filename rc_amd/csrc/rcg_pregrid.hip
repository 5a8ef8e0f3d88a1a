// rcg_pregrid.hip -- FP64 uvw -> (grid cell, GCF window) pregrid on gfx950.
//
// Restates pregridPoint<8, mirror> (MS4/kernel/common/common.h:115-170) and
// the per-point GCF lookup, recentring and in-bound test of the reference
// gridder (MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p.cpp
// :127-144).  The integer outputs must equal the reference's bit for bit, so
// every step is a single IEEE FP64 operation with contraction switched off:
// us = u*scale (one rounding), floor(us), us - floor(us) (exact), 8*(...)
// (exact), floor; w/wstep (correctly rounded division) and round-half-away.
#include <algorithm>
#include <climits>

#include "rcg_internal.h"
#include "rcg_pregrid.h"

namespace rcg {

// One point: its record, plus the in-bound taps and error bits it adds.
__device__ __forceinline__ void pregrid_record(const rcg_geom &g, const rcg_gcf &t, int i,
                                               rcg_pg *__restrict__ out,
                                               unsigned long long &taps, int &err)
{
  // one point per chunk (a support per point) needs no division
  const int S = g.chunk_supps[g.ts_ch == 1 ? (unsigned)i : (unsigned)i / (unsigned)g.ts_ch];
  PregridPoint p = pregrid_point(g.uvw[i], g.scale, g.wstep, S, g.grid_size, t.half != 0);
  rcg_pg rec;
  rec.cell = -1;
  rec.tap0 = 0;
  rec.stride = 0;
  rec.supp = (int16_t)S;
  rec.flags = 0;
  if (p.inbound) {
    int tap0 = 0, stride = 0;
    bool ok = resolve_window(t, p, S, &tap0, &stride);
    if (ok) {
      rec.cell = p.u * g.grid_pitch + p.v;
      rec.tap0 = tap0;
      rec.stride = (int16_t)stride;
      rec.flags = p.conj ? PG_CONJ : 0;
      taps += (unsigned long long)S * (unsigned long long)S;
    } else {
      rec.flags = PG_RANGE_ERR;
      err = ERR_RANGE;
    }
  }
  *out = rec;
}

// Per-workgroup reduction of the tap count and error bits, then one atomic
// each (a single counter hit by every wavefront serialises at the L2).
__device__ __forceinline__ void pregrid_flush(unsigned long long my_taps, int my_err,
                                              unsigned long long *__restrict__ taps_dev,
                                              int *__restrict__ err_dev)
{
  __shared__ unsigned long long s_taps[4];
  __shared__ int s_err;
  if (threadIdx.x == 0) s_err = 0;
  for (int off = 32; off > 0; off >>= 1) my_taps += __shfl_xor(my_taps, off);
  if ((threadIdx.x & 63) == 0) s_taps[threadIdx.x >> 6] = my_taps;
  __syncthreads();
  if (my_err) atomicOr(&s_err, my_err);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long tot = s_taps[0] + s_taps[1] + s_taps[2] + s_taps[3];
    if (tot && taps_dev) atomicAdd(taps_dev, tot);
    if (s_err && err_dev) atomicOr(err_dev, s_err);
  }
}

// Grid-stride over the points of one geometry (point counts are below 2^31,
// check_geom, so 32-bit index arithmetic).
__global__ void __launch_bounds__(256)
k_pregrid(rcg_geom g, rcg_gcf t, rcg_pg *__restrict__ pg,
          unsigned long long *__restrict__ taps_dev, int *__restrict__ err_dev)
{
  const int n = g.n_chunks * g.ts_ch;
  unsigned long long my_taps = 0;
  int my_err = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    pregrid_record(g, t, i, pg + i, my_taps, my_err);
  pregrid_flush(my_taps, my_err, taps_dev, err_dev);
}

// Grid-stride over the concatenated points of n_inst geometries; point i
// belongs to the instance k with pt_off[k] <= i < pt_off[k + 1].
__global__ void __launch_bounds__(256)
k_pregrid_batch(const rcg_geom *__restrict__ geoms, const int *__restrict__ pt_off,
                int n_inst, rcg_gcf t, rcg_pg *__restrict__ pg,
                unsigned long long *__restrict__ taps_dev, int *__restrict__ err_dev)
{
  const int n = pt_off[n_inst];
  unsigned long long my_taps = 0;
  int my_err = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pt_off[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    // The points of a wavefront nearly always belong to one instance; its
    // geometry then comes through scalar loads rather than a vector load of
    // the same record in every lane.
    const int lo0 = __builtin_amdgcn_readfirstlane(lo);
    if (__all(lo == lo0)) {
      const rcg_geom g = geoms[lo0];
      pregrid_record(g, t, i - pt_off[lo0], pg + i, my_taps, my_err);
    } else {
      const rcg_geom g = geoms[lo];
      pregrid_record(g, t, i - pt_off[lo], pg + i, my_taps, my_err);
    }
  }
  pregrid_flush(my_taps, my_err, taps_dev, err_dev);
}

rcg_status launch_pregrid_batch(const rcg_geom *geoms_dev, const int *pt_off_dev, int n_inst,
                                long long n, const rcg_gcf *gcf, rcg_pg *pg,
                                unsigned long long *taps_dev, int *err_dev,
                                hipStream_t stream)
{
  if (n == 0) return RCG_OK;
  // grid-stride: more workgroups add same-address tap-count atomics (measured
  // on 10 M points: 2048 workgroups 0.01 ms faster than 4096, 8192 and
  // 16384 slower)
  const int nb = (int)std::min<long long>(ceil_div(n, 256), 2048);
  hipLaunchKernelGGL(k_pregrid_batch, dim3(nb), dim3(256), 0, stream, geoms_dev, pt_off_dev,
                     n_inst, *gcf, pg, taps_dev, err_dev);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

// Ranges over the in-bound points: out[0..1] = min/max w-plane, out[2..3] =
// min/max "layer plane" (the plane whose 64 layers hold the point's layer:
// floor(index/64) for a full table, |index|/64 for a half one).  The two
// differ when an over-sampling offset rounds up to 8 (a scaled coordinate
// one ulp below an integer) and the layer index runs into the next plane.
__global__ void __launch_bounds__(256)
k_plane_range(rcg_geom g, int half, int *__restrict__ out)
{
  __shared__ int s_r[4][4];
  const int n = g.n_chunks * g.ts_ch;
  int lo = INT_MAX, hi = INT_MIN, llo = INT_MAX, lhi = INT_MIN;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int S = g.chunk_supps[(unsigned)i / (unsigned)g.ts_ch];
    PregridPoint p = pregrid_point(g.uvw[i], g.scale, g.wstep, S, g.grid_size,
                                   half != 0);
    if (p.inbound) {
      lo = min(lo, p.w_plane);
      hi = max(hi, p.w_plane);
      const int idx = half ? (p.index < 0 ? -p.index : p.index) : p.index;
      const int lp = idx >= 0 ? idx / 64 : -((-idx + 63) / 64);
      llo = min(llo, lp);
      lhi = max(lhi, lp);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
    llo = min(llo, __shfl_xor(llo, off));
    lhi = max(lhi, __shfl_xor(lhi, off));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_r[w][0] = lo;
    s_r[w][1] = hi;
    s_r[w][2] = llo;
    s_r[w][3] = lhi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      lo = min(lo, s_r[k][0]);
      hi = max(hi, s_r[k][1]);
      llo = min(llo, s_r[k][2]);
      lhi = max(lhi, s_r[k][3]);
    }
    if (lo <= hi) {
      atomicMin(out, lo);
      atomicMax(out + 1, hi);
      atomicMin(out + 2, llo);
      atomicMax(out + 3, lhi);
    }
  }
}

rcg_status launch_plane_range(const rcg_geom *g, int half, int *out_dev,
                              hipStream_t stream)
{
  const long long n = (long long)g->n_chunks * g->ts_ch;
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>(ceil_div(n, 256), 2048);
  hipLaunchKernelGGL(k_plane_range, dim3(nb), dim3(256), 0, stream, *g, half, out_dev);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}

}  // namespace rcg

extern "C" rcg_status rcg_pregrid(const rcg_geom *geom, const rcg_gcf *gcf,
                                  rcg_pg *pg, unsigned long long *taps_dev,
                                  int *err_dev, void *stream)
{
  using namespace rcg;
  clear_error();
  if (!geom || !gcf || !pg) {
    set_error(RCG_ERR_ARG, "rcg_pregrid: null argument");
    return RCG_ERR_ARG;
  }
  rcg_status s = check_geom(geom);
  if (s != RCG_OK) return s;
  s = check_gcf(gcf);
  if (s != RCG_OK) return s;
  const long long n = (long long)geom->n_chunks * geom->ts_ch;
  if (n == 0) return RCG_OK;
  const int nb = (int)std::min<long long>(ceil_div(n, 256), 2048);
  hipLaunchKernelGGL(k_pregrid, dim3(nb), dim3(256), 0, (hipStream_t)stream, *geom, *gcf,
                     pg, taps_dev, err_dev);
  RCG_HIP_TRY(hipGetLastError());
  return RCG_OK;
}
