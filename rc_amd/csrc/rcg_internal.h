// rcg_internal.h -- shared declarations of the rcgrid library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/rcgrid.h"

namespace rcg {

// ---------------------------------------------------------------- errors --
void set_error(int code, const char *fmt, ...);
void clear_error();

#define RCG_HIP_TRY(expr)                                                    \
  do {                                                                       \
    hipError_t e__ = (expr);                                                 \
    if (e__ != hipSuccess) {                                                 \
      ::rcg::set_error(RCG_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                       hipGetErrorString(e__), __FILE__, __LINE__);          \
      return RCG_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

// Flags of rcg_pg.flags
constexpr int PG_CONJ = 1;
constexpr int PG_RANGE_ERR = 2;

// Device error word bits (err_dev)
constexpr int ERR_RANGE = 1;

// Largest support handled by any kernel (int16 record field, LDS tiles).
constexpr int MAX_SUPPORT = 256;

// Launch helpers ------------------------------------------------------------
inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Validates geometry sizes shared by every entry point.
rcg_status check_geom(const rcg_geom *g);
rcg_status check_gcf(const rcg_gcf *t);

// Kernels launched from the legacy host wrappers ----------------------------
// Plane ranges over in-bound points: out[0..1] = min/max w-plane,
// out[2..3] = min/max plane of the referenced layers (out must be preset to
// {INT_MAX, INT_MIN, INT_MAX, INT_MIN}).
rcg_status launch_plane_range(const rcg_geom *g, int half, int *out_dev,
                              hipStream_t stream);

// Pregrid of n_inst geometries (device array) whose points are concatenated:
// instance k owns points [pt_off[k], pt_off[k + 1]) of pg, n = pt_off[n_inst].
rcg_status launch_pregrid_batch(const rcg_geom *geoms_dev, const int *pt_off_dev, int n_inst,
                                long long n, const rcg_gcf *gcf, rcg_pg *pg,
                                unsigned long long *taps_dev, int *err_dev,
                                hipStream_t stream);

}  // namespace rcg
