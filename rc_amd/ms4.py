"""Python calls of the reference's MS4 kernel entry points (host memory).

Each function takes numpy arrays, builds the same pointer tables the
reference's callers build (MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:64-85,
MS4/misc/cppcycle.cpp:190-223) and calls the same-named symbol of
librcgrid.so, which runs on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .tables import GCFTable, VisChunks


def _legacy(name: str, vc: VisChunks, gcf: GCFTable, grid: np.ndarray,
            vis: np.ndarray, grid_size: int, grid_pitch: int) -> int:
    h = _lib.lib()
    fn = getattr(h, name)
    uvw = np.ascontiguousarray(vc.uvw, np.float64)
    supps = np.ascontiguousarray(vc.supps, np.int32)
    uvw_p = vc.chunk_pointers(uvw)
    vis_p = vc.chunk_pointers(vis)
    gcf_c, sup_c = gcf.pointer_tables()
    taps = fn(vc.scale, vc.wstep, vc.n_chunks, supps.ctypes.data,
              grid.ctypes.data, gcf_c, uvw_p, vis_p, vc.ts_ch, grid_pitch,
              grid_size, sup_c)
    code, msg = _lib.last_error()
    if code:
        raise _lib.RcgError(f"{name}: {msg}")
    return int(taps)


def grid(vc: VisChunks, gcf: GCFTable, grid_size: int,
         grid_pitch: int | None = None) -> tuple[np.ndarray, int]:
    """gridKernelCPU{Full,Half}GCF: returns (grid[size, pitch] complex128,
    in-bound tap count)."""
    pitch = grid_pitch or grid_size
    if gcf.data.dtype != np.complex128:
        raise TypeError("the MS4 entry points take complex128 GCF taps")
    g = np.empty((grid_size, pitch), np.complex128)
    vis = np.ascontiguousarray(vc.vis, np.complex128)
    name = "gridKernelCPUHalfGCF" if gcf.half else "gridKernelCPUFullGCF"
    taps = _legacy(name, vc, gcf, g, vis, grid_size, pitch)
    return g, taps


def degrid(vc: VisChunks, gcf: GCFTable, grid_arr: np.ndarray,
           grid_size: int | None = None) -> tuple[np.ndarray, int]:
    """deGridKernelCPU{Full,Half}GCF: returns (vis[n_chunks, ts_ch]
    complex128, in-bound tap count)."""
    g = np.ascontiguousarray(grid_arr, np.complex128)
    size = grid_size or g.shape[0]
    pitch = g.shape[1]
    if gcf.data.dtype != np.complex128:
        raise TypeError("the MS4 entry points take complex128 GCF taps")
    out = np.zeros((vc.n_chunks, vc.ts_ch), np.complex128)
    name = "deGridKernelCPUHalfGCF" if gcf.half else "deGridKernelCPUFullGCF"
    taps = _legacy(name, vc, gcf, g, out, size, pitch)
    return out, taps
