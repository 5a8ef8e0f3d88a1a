"""Python calls of the MS6 visibility binning entry points
(MS6/programs/Kernel/Binning.hs binSizer / binner), which run on the GPU
through librcgrid.so.  Regions, boxes and bins are in uv units."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _check(status: int, what: str) -> None:
    if status != 0:
        raise _lib.RcgError(f"{what}: {_lib.last_error()[1]}")


def bin_sizes(vis: np.ndarray, regions: np.ndarray, bins: int):
    """binSizer: returns (low, high, start_end (R, bins, 2) float64,
    counts (R, bins) int64) for vis (n, width >= 3) and regions (R, 4) =
    (u_low, u_high, v_low, v_high)."""
    vis = np.ascontiguousarray(vis, np.float64)
    reg = np.ascontiguousarray(regions, np.float64)
    out = np.zeros((len(reg), bins, 3), np.int64)
    lh = np.zeros(2)
    n, width = vis.shape
    _check(_lib.lib().rcg_flow_bin_sizes(vis.ctypes.data, n, width, reg.ctypes.data, len(reg),
                                         bins, out.ctypes.data, lh.ctypes.data),
           "rcg_flow_bin_sizes")
    se = out[:, :, :2].copy().view(np.float64)
    return float(lh[0]), float(lh[1]), se, out[:, :, 2].copy()


def binner(vis: np.ndarray, boxes: np.ndarray, box_bins, fill: float = np.nan):
    """binner: box_bins[b] = [(w_low, w_high, size), ...].  Returns (outs,
    filled) with outs[b] (sum of sizes, 5) float64, first filled with
    `fill` (the reference does not clear its vectors)."""
    vis = np.ascontiguousarray(vis, np.float64)
    bx = np.ascontiguousarray(boxes, np.float64)
    off = np.zeros(len(bx) + 1, np.int32)
    ranges, sizes = [], []
    for b, bb in enumerate(box_bins):
        off[b + 1] = off[b] + len(bb)
        for wl, wh, sz in bb:
            ranges += [wl, wh]
            sizes.append(sz)
    rng = np.ascontiguousarray(ranges if ranges else [0.0], np.float64)
    szs = np.ascontiguousarray(sizes if sizes else [0], np.int64)
    outs = [np.full((int(sum(s for _, _, s in bb)), 5), fill) for bb in box_bins]
    ptrs = (C.c_void_p * len(outs))(*[o.ctypes.data if o.size else None for o in outs])
    filled = np.zeros(len(bx), np.int64)
    n, width = vis.shape
    _check(_lib.lib().rcg_flow_binner(vis.ctypes.data, n, width, bx.ctypes.data, len(bx),
                                      off.ctypes.data, rng.ctypes.data, szs.ctypes.data,
                                      ptrs, filled.ctypes.data),
           "rcg_flow_binner")
    return outs, filled
