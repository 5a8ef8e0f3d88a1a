"""CPU restatement of the MS6 dna-flow visibility binning kernels
(TEST INFRASTRUCTURE ONLY: tests/ use it as the checker; the product
library never imports it).

Reference: MS6/programs/Kernel/Binning.hs
  binSizer  (:34-97)   w range over the points inside the uv bounding box of
                       the regions, then per region and w bin
                       (start, end, count) triples
  binner    (:99-181)  every point appended, in input order, to the output
                       vector of its (u, v, w) box, then zero padding

Regions and bins are given in uv units (the caller applies
gridXY2UV, MS6/programs/Kernel/Config.hs:211-212).  Haskell's Data.Map
lookups are restated literally: lookupLE k = the entry with the greatest
key <= k.

Pinning (no GHC here, so the Haskell cannot run): hand-worked KATs of the
bin edges, the inflated upper bound, the lookupLE chains (a point in a gap
between regions is counted by binSizer, which checks no upper bound, and
dropped by binner, which does), the all-zero-w case; invariants: the
counts add up to the points inside the bounding box, and binner's output
is the input's kept rows in input order per box.
"""
from __future__ import annotations

import math

import numpy as np


def _lookup_le(keys, k):
    """Index of the greatest key <= k in the sorted list `keys`, or -1."""
    best = -1
    for j, x in enumerate(keys):
        if x <= k:
            best = j
    return best


def _hs_floor_int(x: float) -> int:
    """GHC's floor :: Double -> Int on x86-64: floorDouble, then the
    truncating conversion, which maps NaN (and out-of-range values) to
    minBound."""
    if math.isnan(x) or math.isinf(x):
        return -(2 ** 63)
    return int(math.floor(x))


def bin_sizes(vis: np.ndarray, regions: np.ndarray, bins: int):
    """binSizer (Binning.hs:34-97).  vis (n, width >= 3) doubles, regions
    (R, 4) = (u_low, u_high, v_low, v_high) in uv units.  Returns (low,
    high, out) with out (R, bins, 3): start, end (floats) and count."""
    umin = float(np.min(regions[:, 0]))
    vmin = float(np.min(regions[:, 2]))
    umax = float(np.max(regions[:, 1]))
    vmax = float(np.max(regions[:, 3]))
    low, high = 0.0, 0.0                       # foldM from (0, 0), :51
    for u, v, w in vis[:, :3]:
        if u >= umin and u < umax and v >= vmin and v < vmax:
            if math.isnan(w):
                # Haskell's max makes the fold's result depend on where the
                # NaN sits; the library rejects such input
                raise ValueError("NaN w inside the regions")
            low = w if w <= low else low       # Prelude min / max on Double
            high = high if w <= high else w
    high = high + (high - low) * 0.0001        # :57

    def bin_start(b):                          # :61
        return low + float(b) * (high - low) / float(bins)

    out = np.zeros((len(regions), bins, 3))
    for r in range(len(regions)):
        for b in range(bins):
            out[r, b, 0] = bin_start(b)
            out[r, b, 1] = bin_start(b + 1)
    # binVecMap: u_low -> (v_low -> region); Map.fromListWith Map.union
    ulows = sorted(set(float(x) for x in regions[:, 0]))
    vmaps = {ul: {} for ul in ulows}
    for r, (ul, _, vl, _) in enumerate(regions):
        vmaps[float(ul)][float(vl)] = r        # later duplicates win (:79)
    for u, v, w in vis[:, :3]:
        if not (u >= umin and u < umax and v >= vmin and v < vmax and w >= low and w <= high):
            continue
        iu = _lookup_le(ulows, u)
        if iu < 0:
            continue
        vm = vmaps[ulows[iu]]
        vkeys = sorted(vm)
        iv = _lookup_le(vkeys, v)
        if iv < 0:
            continue
        r = vm[vkeys[iv]]
        with np.errstate(invalid="ignore", divide="ignore"):
            b = _hs_floor_int(float(np.float64(w - low) / np.float64(high - low)) * float(bins))
        b = max(0, min(bins - 1, b))
        out[r, b, 2] += 1
    return low, high, out


def binner(vis: np.ndarray, boxes: np.ndarray, box_bins: list[list[tuple[float, float, int]]],
           fill: float = 0.0):
    """binner (Binning.hs:99-181).  boxes (B, 4) = (u_low, u_high, v_low,
    v_high); box_bins[b] = [(w_low, w_high, size), ...] of box b.  Box b's
    output holds sum(size) rows of 5 doubles, initially `fill` (the
    reference's vectors are not cleared).  Returns (outs, filled):
    outs[b] (rows, 5) after the zero padding, filled[b] rows appended.
    Raises if a box receives more rows than it holds (the reference writes
    past its vector)."""
    if vis.shape[1] != 5:
        raise ValueError("wBinner: Unexpected data width!")
    # outPtrMap: w_low -> v_low -> u_low -> (ranges, box)
    m: dict = {}
    for b, (ul, uh, vl, vh) in enumerate(boxes):
        for (wl, wh, _) in box_bins[b]:
            inner = m.setdefault(float(wl), {}).setdefault(float(vl), {})
            inner.setdefault(float(ul), (float(ul), float(uh), float(vl), float(vh),
                                         float(wl), float(wh), b))  # first one wins
    cap = [sum(s for _, _, s in bb) for bb in box_bins]
    outs = [np.full((c, 5), fill) for c in cap]
    filled = [0] * len(boxes)
    wkeys = sorted(m)
    for row in vis:
        u, v, w = row[0], row[1], row[2]
        iw = _lookup_le(wkeys, w)
        if iw < 0:
            continue
        mv = m[wkeys[iw]]
        vkeys = sorted(mv)
        iv = _lookup_le(vkeys, v)
        if iv < 0:
            continue
        mu = mv[vkeys[iv]]
        ukeys = sorted(mu)
        iu = _lookup_le(ukeys, u)
        if iu < 0:
            continue
        ul, uh, vl, vh, wl, wh, b = mu[ukeys[iu]]
        if not (ul <= u < uh and vl <= v < vh and wl <= w < wh):
            continue
        if filled[b] >= cap[b]:
            raise OverflowError(f"box {b} receives more than {cap[b]} rows")
        outs[b][filled[b]] = row
        filled[b] += 1
    # padding (:156-178): every bin of the box pads from the box's fill to
    # the bin's own size
    for b in range(len(boxes)):
        for (_, _, size) in box_bins[b]:
            if size > filled[b]:
                outs[b][filled[b]:size] = 0.0
    return outs, filled
