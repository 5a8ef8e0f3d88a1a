// rcg_pregrid.h -- exact FP64 pregrid shared by every gridding kernel.
#pragma once

#include "rcg_internal.h"

namespace rcg {

struct PregridPoint {
  int u, v;          // window top-left cell (already shifted by N/2 - S/2)
  int over_u, over_v;
  int w_plane;
  int index;         // signed GCF layer index
  bool conj;         // HalfGCF and w_plane < 0
  bool inbound;      // representable and window strictly inside the grid
};

// pregridPoint<8, mirror>, MS4/kernel/common/common.h:115-170, plus the
// in-bound test of scatter_gridder_w_dependent_dyn_1p.cpp:142-144.  The
// reference narrows every coordinate to int16; values outside that range are
// undefined behaviour there and are reported here as not in-bound.
__device__ __forceinline__ PregridPoint pregrid_point(rcg_double3 p,
                                                      double scale,
                                                      double wstep, int S,
                                                      int N, bool mirror)
{
#pragma clang fp contract(off)
  PregridPoint r;
  const double us = p.u * scale;              // common.h:123
  const double vs = p.v * scale;              // common.h:124
  const double wr = round(p.w / wstep);       // common.h:128 (w unscaled)
  const double fu = floor(us);                // common.h:129
  const double fv = floor(vs);                // common.h:130
  const bool repr = fu >= -32768.0 && fu <= 32767.0 && fv >= -32768.0 &&
                    fv <= 32767.0 && wr >= -32768.0 && wr <= 32767.0;
  if (!repr) {
    r.u = r.v = r.over_u = r.over_v = r.w_plane = r.index = 0;
    r.conj = false;
    r.inbound = false;
    return r;
  }
  int u = (int)fu, v = (int)fv;
  const int wp = (int)wr;
  // common.h:131-132: over * (uvw.u - u); us - fu is exact
  const int ou = (int)floor(8.0 * (us - fu));
  const int ov = (int)floor(8.0 * (vs - fv));
  // common.h:136-137
  u += N / 2 - S / 2;
  v += N / 2 - S / 2;
  int idx;
  bool conj = false;
  if (mirror && wp < 0) {                     // common.h:158-163
    idx = -((-wp * 8 + ou) * 8 + ov);
    conj = true;
  } else {
    idx = (wp * 8 + ou) * 8 + ov;
  }
  r.u = u;
  r.v = v;
  r.over_u = ou;
  r.over_v = ov;
  r.w_plane = wp;
  r.index = idx;
  r.conj = conj;
  r.inbound = idx >= -32768 && idx <= 32767 &&
              !(u < 0 || u >= N - S || v < 0 || v >= N - S);
  return r;
}

// Layer lookup and recentring, scatter_gridder_w_dependent_dyn_1p.cpp:
// 128-140: slot = |index| (half) or index + center (full); plane support gs
// from the signed plane; window start (gs - S)/2*(gs + 1) with C truncation.
// Returns false when the slot or plane is outside the table or the S x S
// window does not lie inside the layer (undefined behaviour in the
// reference).
__device__ __forceinline__ bool resolve_window(const rcg_gcf &t,
                                               const PregridPoint &p, int S,
                                               int *tap0, int *stride)
{
  const int slot = t.half ? (p.index < 0 ? -p.index : p.index)
                          : p.index + t.center;
  if (slot < 0 || slot >= t.n_slots) return false;
  if (p.w_plane < -t.max_w_plane || p.w_plane > t.max_w_plane) return false;
  const int gs = t.supps[p.w_plane + t.max_w_plane];
  if (gs <= 0) return false;
  const int off = (gs - S) / 2 * (gs + 1);
  if (off < 0) return false;
  if (S > 0) {
    const long long last = (long long)off + (long long)(S - 1) * (gs + 1);
    if (last >= (long long)t.layer_len[slot]) return false;
  }
  const long long start = t.layer_off[slot] + off;
  if (start < 0 || start > 0x7fffffffLL) return false;
  *tap0 = (int)start;
  *stride = gs;
  return true;
}

}  // namespace rcg
