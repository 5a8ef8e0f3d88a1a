// rcg_legacy.hip -- the reference's MS4 gridder/degridder entry points over
// HIP, with every buffer in host memory (the contract of
// MS4/kernel/common/scatter_gridder_w_dependent_dyn_1p.h:16-91 as called
// from MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:64-111 and
// MS4/misc/cppcycle.cpp:210-265).
//
// Each call: gather the per-chunk uvw/vis pointers into one contiguous
// buffer (or copy straight from the caller's memory when its chunks are
// consecutive, as the reference's callers lay them out), find the w-plane
// range of the in-bound points on the GPU, upload only the GCF layers of
// those planes, pregrid (FP64, exact), grid (tiled FP64 gridder over binned
// pairs, or the plain one where no plan fits) or degrid (plain FP64
// degridder), copy the result back.  No state survives the call, as in the reference; calls from
// different host threads use their own streams and buffers.
#include <algorithm>
#include <climits>
#include <cstring>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>

#include "rcg_internal.h"
#include "rcg_tiled.h"

namespace rcg {
namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct PlanGuard {
  rcg_tiled_plan *p = nullptr;
  ~PlanGuard() { if (p) (void)rcg_tiled_plan_destroy(p); }
};

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() { if (s) (void)hipStreamDestroy(s); }
};

#define LEG_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e__ = (expr);                                                 \
    if (e__ != hipSuccess) {                                                 \
      set_error(RCG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e__)); \
      return fail();                                                         \
    }                                                                        \
  } while (0)

struct LegacyArgs {
  bool degrid, half;
  double scale, wstep;
  int baselines;
  const int *bl_supps;
  rcg_c128 *grid_out;            // gridding
  const rcg_c128 *grid_in;       // degridding
  const rcg_c128 *const *gcf;    // centred layer table
  const rcg_double3 *const *uvw;
  const rcg_c128 *const *vis_in; // gridding
  rcg_c128 *const *vis_out;      // degridding
  int ts_ch, pitch, size;
  const int *gcf_supps;          // centred
};

// Gridder the calling thread's last gridKernelCPU*GCF call ran
// (rcg_legacy_grid_path).
thread_local int t_last_path = RCG_PATH_NONE;

ull legacy_call(const LegacyArgs &a)
{
  clear_error();
  t_last_path = RCG_PATH_NONE;
  const size_t gcells = (a.size > 0 && a.pitch > 0) ? (size_t)a.size * a.pitch : 0;
  auto zero_outputs = [&]() {
    if (!a.degrid) {
      if (a.grid_out && gcells) memset(a.grid_out, 0, gcells * sizeof(rcg_c128));
    } else if (a.vis_out && a.ts_ch > 0) {
      for (int bl = 0; bl < a.baselines; bl++)
        if (a.vis_out[bl]) memset(a.vis_out[bl], 0, (size_t)a.ts_ch * sizeof(rcg_c128));
    }
  };
  auto fail = [&]() -> ull { zero_outputs(); return 0ULL; };
  // host staging without value-initialisation (the copies fill it)
  auto host_array = [](size_t count, auto *tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    return std::unique_ptr<T[]>(new (std::nothrow) T[count ? count : 1]);
  };

  if (a.baselines < 0 || a.ts_ch < 0 || a.size <= 0 || a.size > 32767 ||
      a.pitch < a.size || (long long)a.size * a.pitch > 0x7fffffffLL) {
    set_error(RCG_ERR_ARG, "bad sizes: baselines %d, ts_ch %d, size %d, pitch %d",
              a.baselines, a.ts_ch, a.size, a.pitch);
    return 0;
  }
  if ((a.degrid && (!a.grid_in || (!a.vis_out && a.baselines))) ||
      (!a.degrid && (!a.grid_out || (!a.vis_in && a.baselines)))) {
    set_error(RCG_ERR_ARG, "null grid or visibility table");
    return 0;
  }
  const long long n = (long long)a.baselines * a.ts_ch;
  if (n == 0) {
    zero_outputs();
    return 0;
  }
  if (n > 0x7fffffffLL) {
    set_error(RCG_ERR_ARG, "%lld points exceed 2^31 - 1", n);
    return fail();
  }
  if (!a.bl_supps || !a.uvw || !a.gcf || !a.gcf_supps) {
    set_error(RCG_ERR_ARG, "null support, uvw or GCF table");
    return fail();
  }
  for (int bl = 0; bl < a.baselines; bl++) {
    if (a.bl_supps[bl] < 0 || a.bl_supps[bl] > MAX_SUPPORT) {
      set_error(RCG_ERR_ARG, "chunk %d support %d outside [0, %d]", bl,
                a.bl_supps[bl], MAX_SUPPORT);
      return fail();
    }
    if (!a.uvw[bl] || (a.degrid ? !a.vis_out[bl] : !a.vis_in[bl])) {
      set_error(RCG_ERR_ARG, "chunk %d: null uvw or vis pointer", bl);
      return fail();
    }
  }

  // ---- gather host inputs --------------------------------------------------
  // Chunks laid out one after the other in one array are copied from there
  // directly; otherwise they are gathered into a staging buffer.
  auto consecutive = [&](const void *const *tab, size_t elem) {
    const char *b0 = static_cast<const char *>(tab[0]);
    for (int bl = 1; bl < a.baselines; bl++)
      if (static_cast<const char *>(tab[bl]) != b0 + (size_t)bl * a.ts_ch * elem) return false;
    return true;
  };
  const bool uvw_flat = consecutive(reinterpret_cast<const void *const *>(a.uvw),
                                    sizeof(rcg_double3));
  std::unique_ptr<rcg_double3[]> huvw_buf;
  const rcg_double3 *huvw = a.uvw[0];
  if (!uvw_flat) {
    huvw_buf = host_array((size_t)n, (rcg_double3 *)nullptr);
    if (!huvw_buf) {
      set_error(RCG_ERR_NOMEM, "no host memory for %lld points", n);
      return fail();
    }
    for (int bl = 0; bl < a.baselines; bl++)
      memcpy(&huvw_buf[(size_t)bl * a.ts_ch], a.uvw[bl], (size_t)a.ts_ch * sizeof(rcg_double3));
    huvw = huvw_buf.get();
  }

  StreamGuard sg;
  LEG_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  hipStream_t st = sg.s;

  DevBuf d_uvw, d_supps, d_range, d_pg, d_taps, d_err;
  LEG_TRY(d_uvw.alloc((size_t)n * sizeof(rcg_double3)));
  LEG_TRY(d_supps.alloc((size_t)a.baselines * sizeof(int)));
  LEG_TRY(d_range.alloc(4 * sizeof(int)));
  LEG_TRY(d_pg.alloc((size_t)n * sizeof(rcg_pg)));
  LEG_TRY(d_taps.alloc(sizeof(unsigned long long)));
  LEG_TRY(d_err.alloc(sizeof(int)));
  LEG_TRY(hipMemcpyAsync(d_uvw.p, huvw, (size_t)n * sizeof(rcg_double3),
                         hipMemcpyHostToDevice, st));
  LEG_TRY(hipMemcpyAsync(d_supps.p, a.bl_supps, (size_t)a.baselines * sizeof(int),
                         hipMemcpyHostToDevice, st));
  int hrange[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  LEG_TRY(hipMemcpyAsync(d_range.p, hrange, sizeof(hrange), hipMemcpyHostToDevice, st));
  LEG_TRY(hipMemsetAsync(d_taps.p, 0, sizeof(unsigned long long), st));
  LEG_TRY(hipMemsetAsync(d_err.p, 0, sizeof(int), st));

  rcg_geom g;
  g.uvw = d_uvw.as<rcg_double3>();
  g.chunk_supps = d_supps.as<int32_t>();
  g.n_chunks = a.baselines;
  g.ts_ch = a.ts_ch;
  g.scale = a.scale;
  g.wstep = a.wstep;
  g.grid_size = a.size;
  g.grid_pitch = a.pitch;
  if (launch_plane_range(&g, a.half ? 1 : 0, d_range.as<int>(), st) != RCG_OK)
    return fail();
  LEG_TRY(hipMemcpyAsync(hrange, d_range.p, sizeof(hrange), hipMemcpyDeviceToHost, st));
  LEG_TRY(hipStreamSynchronize(st));
  if (hrange[0] > hrange[1]) {  // no in-bound point: all zero
    zero_outputs();
    return 0;
  }

  // ---- upload the GCF layers in use ----------------------------------------
  // Supports are needed for the w-planes of the in-bound points, layers for
  // the planes their layer indices fall in (the same planes except where an
  // over-sampling offset of 8 runs into the next plane's first layers).
  // Both tables are assumed to extend that far, as in the reference, which
  // reads them through the same pointers.
  const int pmin = hrange[0], pmax = hrange[1];
  const int lmin = hrange[2], lmax = hrange[3];
  const int M = std::max(std::abs(pmin), std::abs(pmax));
  int n_slots, center;
  if (!a.half) {
    n_slots = (lmax - lmin + 1) * 64;
    center = -lmin * 64;
  } else {
    n_slots = (lmax + 1) * 64;
    center = 0;
  }
  std::vector<int64_t> hoff((size_t)n_slots, 0);
  std::vector<int32_t> hlen((size_t)n_slots, 0);
  std::vector<int32_t> hsupp((size_t)(2 * M + 1), 0);
  for (int p = pmin; p <= pmax; p++) hsupp[(size_t)(p + M)] = a.gcf_supps[p];
  // (slot, source layer pointer, length) for every layer to upload; a
  // layer of plane q holds gcf_supps[q]^2 taps
  struct Src { int slot; const rcg_c128 *ptr; int len; };
  std::vector<Src> srcs;
  for (int q = lmin; q <= lmax; q++) {
    const int gs = a.gcf_supps[q];
    for (int k = 0; k < 64; k++) {
      const int idx = q * 64 + k;
      srcs.push_back({idx + center, a.gcf[idx], gs > 0 ? gs * gs : 0});
    }
  }
  long long total = 0;
  for (const Src &s : srcs) {
    if (!s.ptr) { hlen[(size_t)s.slot] = 0; continue; }
    hoff[(size_t)s.slot] = total;
    hlen[(size_t)s.slot] = s.len;
    total += s.len;
  }
  const size_t n_data = (size_t)(total > 0 ? total : 1);
  std::unique_ptr<rcg_c128[]> hdata = host_array(n_data, (rcg_c128 *)nullptr);
  if (!hdata) {
    set_error(RCG_ERR_NOMEM, "no host memory for %lld GCF taps", total);
    return fail();
  }
  if (total == 0) hdata[0] = rcg_c128{0.0, 0.0};
  for (const Src &s : srcs)
    if (s.ptr && s.len)
      memcpy(&hdata[(size_t)hoff[(size_t)s.slot]], s.ptr, (size_t)s.len * sizeof(rcg_c128));

  DevBuf d_data, d_off, d_len, d_tsupp;
  LEG_TRY(d_data.alloc(n_data * sizeof(rcg_c128)));
  LEG_TRY(d_off.alloc(hoff.size() * sizeof(int64_t)));
  LEG_TRY(d_len.alloc(hlen.size() * sizeof(int32_t)));
  LEG_TRY(d_tsupp.alloc(hsupp.size() * sizeof(int32_t)));
  LEG_TRY(hipMemcpyAsync(d_data.p, hdata.get(), n_data * sizeof(rcg_c128),
                         hipMemcpyHostToDevice, st));
  LEG_TRY(hipMemcpyAsync(d_off.p, hoff.data(), hoff.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, st));
  LEG_TRY(hipMemcpyAsync(d_len.p, hlen.data(), hlen.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, st));
  LEG_TRY(hipMemcpyAsync(d_tsupp.p, hsupp.data(), hsupp.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, st));
  rcg_gcf t;
  t.data = d_data.p;
  t.layer_off = d_off.as<int64_t>();
  t.layer_len = d_len.as<int32_t>();
  t.n_slots = n_slots;
  t.center = center;
  t.supps = d_tsupp.as<int32_t>();
  t.max_w_plane = M;
  t.half = a.half ? 1 : 0;
  t.n_taps = (int64_t)n_data;

  if (rcg_pregrid(&g, &t, d_pg.as<rcg_pg>(), d_taps.as<unsigned long long>(),
                  d_err.as<int>(), st) != RCG_OK)
    return fail();

  DevBuf d_vis, d_grid;
  LEG_TRY(d_vis.alloc((size_t)n * sizeof(rcg_c128)));
  LEG_TRY(d_grid.alloc(gcells * sizeof(rcg_c128)));
  // Gridding: the tiled FP64 gridder over binned pairs where a plan fits
  // the grid (at most MAX_PLAN_TILES tiles: every grid of the int16
  // coordinate range at tile 32) and the supports (at most 64 tiles per
  // window); otherwise, or when the plan's memory cannot be had, the plain
  // FP64 gridder.  Degridding: the plain FP64 degridder, a gather without
  // atomics that runs faster than the tiled one (3.8 ms against 5.2 ms on
  // the 4096^2 w-projection points in 32640 chunks of 306).
  PlanGuard plan;
  int max_s = 1;
  for (int bl = 0; bl < a.baselines; bl++) max_s = std::max(max_s, a.bl_supps[bl]);
  for (int tile : {32, 64}) {
    if (a.degrid) break;
    const long long nt = (a.size + tile - 1) / tile;
    if (nt * nt > MAX_PLAN_TILES) continue;
    if (rcg_tiled_plan_create(&g, tile, 1024, max_s, &plan.p) == RCG_OK) break;
    plan.p = nullptr;
  }
  clear_error();
  (void)hipGetLastError();  // a failed plan allocation leaves its error behind
  if (plan.p) {
    if (rcg_tiled_bin(plan.p, &t, st) != RCG_OK) return fail();
    int perr = 0;
    if (rcg_tiled_stats(plan.p, nullptr, nullptr, nullptr, &perr, st) != RCG_OK) return fail();
    if (perr & ERR_CAPACITY) {
      set_error(RCG_ERR_HIP, "tiled plan: pair buffer overflow");
      return fail();
    }
  }
  t_last_path = plan.p ? RCG_PATH_TILED : RCG_PATH_PLAIN;
  if (!a.degrid) {
    std::unique_ptr<rcg_c128[]> hvis_buf;
    const rcg_c128 *hvis = a.vis_in[0];
    if (!consecutive(reinterpret_cast<const void *const *>(a.vis_in), sizeof(rcg_c128))) {
      hvis_buf = host_array((size_t)n, (rcg_c128 *)nullptr);
      if (!hvis_buf) {
        set_error(RCG_ERR_NOMEM, "no host memory for %lld visibilities", n);
        return fail();
      }
      for (int bl = 0; bl < a.baselines; bl++)
        memcpy(&hvis_buf[(size_t)bl * a.ts_ch], a.vis_in[bl],
               (size_t)a.ts_ch * sizeof(rcg_c128));
      hvis = hvis_buf.get();
    }
    LEG_TRY(hipMemcpyAsync(d_vis.p, hvis, (size_t)n * sizeof(rcg_c128),
                           hipMemcpyHostToDevice, st));
    if (plan.p) {
      if (rcg_grid_tiled_f64(plan.p, &t, d_vis.as<rcg_c128>(), d_grid.as<rcg_c128>(), st) !=
          RCG_OK)
        return fail();
    } else if (rcg_grid_plain_f64(&g, &t, d_pg.as<rcg_pg>(), d_vis.as<rcg_c128>(),
                                  d_grid.as<rcg_c128>(), st) != RCG_OK) {
      return fail();
    }
    LEG_TRY(hipMemcpyAsync(a.grid_out, d_grid.p, gcells * sizeof(rcg_c128),
                           hipMemcpyDeviceToHost, st));
    LEG_TRY(hipStreamSynchronize(st));
  } else {
    LEG_TRY(hipMemcpyAsync(d_grid.p, a.grid_in, gcells * sizeof(rcg_c128),
                           hipMemcpyHostToDevice, st));
    if (rcg_degrid_plain_f64(&g, &t, d_pg.as<rcg_pg>(), d_grid.as<rcg_c128>(),
                             d_vis.as<rcg_c128>(), st) != RCG_OK)
      return fail();
    if (consecutive(reinterpret_cast<const void *const *>(a.vis_out), sizeof(rcg_c128))) {
      LEG_TRY(hipMemcpyAsync(a.vis_out[0], d_vis.p, (size_t)n * sizeof(rcg_c128),
                             hipMemcpyDeviceToHost, st));
      LEG_TRY(hipStreamSynchronize(st));
    } else {
      std::unique_ptr<rcg_c128[]> hvis = host_array((size_t)n, (rcg_c128 *)nullptr);
      if (!hvis) {
        set_error(RCG_ERR_NOMEM, "no host memory for %lld visibilities", n);
        return fail();
      }
      LEG_TRY(hipMemcpyAsync(hvis.get(), d_vis.p, (size_t)n * sizeof(rcg_c128),
                             hipMemcpyDeviceToHost, st));
      LEG_TRY(hipStreamSynchronize(st));
      for (int bl = 0; bl < a.baselines; bl++)
        memcpy(a.vis_out[bl], &hvis[(size_t)bl * a.ts_ch], (size_t)a.ts_ch * sizeof(rcg_c128));
    }
  }
  unsigned long long taps = 0;
  int err = 0;
  LEG_TRY(hipMemcpy(&taps, d_taps.p, sizeof(taps), hipMemcpyDeviceToHost));
  LEG_TRY(hipMemcpy(&err, d_err.p, sizeof(err), hipMemcpyDeviceToHost));
  if (err & ERR_RANGE) {
    set_error(RCG_ERR_RANGE,
              "an in-bound point referenced a GCF layer outside the table or a "
              "window larger than its layer (chunk support above plane support)");
    return fail();
  }
  return taps;
}

}  // namespace
}  // namespace rcg

#define RCG_LEGACY_GRID(name, is_half)                                         \
  extern "C" ull name(double scale, double wstep, int baselines,              \
                      const int bl_supps[], rcg_c128 grid[],                  \
                      const rcg_c128 *gcf[], const rcg_double3 *uvw[],        \
                      const rcg_c128 *vis[], int ts_ch, int grid_pitch,       \
                      int grid_size, int gcf_supps[])                         \
  {                                                                           \
    rcg::LegacyArgs a;                                                        \
    a.degrid = false; a.half = is_half; a.scale = scale; a.wstep = wstep;     \
    a.baselines = baselines; a.bl_supps = bl_supps; a.grid_out = grid;        \
    a.grid_in = nullptr; a.gcf = gcf; a.uvw = uvw; a.vis_in = vis;            \
    a.vis_out = nullptr; a.ts_ch = ts_ch; a.pitch = grid_pitch;               \
    a.size = grid_size; a.gcf_supps = gcf_supps;                              \
    return rcg::legacy_call(a);                                               \
  }

#define RCG_LEGACY_DEGRID(name, is_half)                                       \
  extern "C" ull name(double scale, double wstep, int baselines,              \
                      const int bl_supps[], const rcg_c128 grid[],            \
                      const rcg_c128 *gcf[], const rcg_double3 *uvw[],        \
                      rcg_c128 *vis[], int ts_ch, int grid_pitch,             \
                      int grid_size, int gcf_supps[])                         \
  {                                                                           \
    rcg::LegacyArgs a;                                                        \
    a.degrid = true; a.half = is_half; a.scale = scale; a.wstep = wstep;      \
    a.baselines = baselines; a.bl_supps = bl_supps; a.grid_out = nullptr;     \
    a.grid_in = grid; a.gcf = gcf; a.uvw = uvw; a.vis_in = nullptr;           \
    a.vis_out = vis; a.ts_ch = ts_ch; a.pitch = grid_pitch;                   \
    a.size = grid_size; a.gcf_supps = gcf_supps;                              \
    return rcg::legacy_call(a);                                               \
  }

RCG_LEGACY_GRID(gridKernelCPUFullGCF, false)
RCG_LEGACY_GRID(gridKernelCPUHalfGCF, true)
RCG_LEGACY_DEGRID(deGridKernelCPUFullGCF, false)
RCG_LEGACY_DEGRID(deGridKernelCPUHalfGCF, true)

extern "C" int rcg_legacy_path(void) { return rcg::t_last_path; }
