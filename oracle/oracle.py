"""ctypes access to the checker libraries (TEST INFRASTRUCTURE ONLY).

    liboracle.so          our C restatement (rc_oracle.c)
    _ref/libms4ref.so     the reference's own CPU gridder, compiled from its
                          sources by oracle/Makefile (no fast-math)
    _ref/libms4ref_fast.so  the same with the reference's release flags

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker.  Data arguments use rc_amd.tables (the host data
formats), never the GPU library.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from rc_amd.tables import GCFTable, VisChunks

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libms4ref.so"
REF_FAST_LIB = HERE / "_ref" / "libms4ref_fast.so"

vp = C.c_void_p
_GRID_ARGS = [C.c_double, C.c_double, C.c_int, vp, vp, vp, vp, vp, C.c_int,
              C.c_int, C.c_int, vp]

_cache: dict[str, C.CDLL] = {}


def _load(path: Path) -> C.CDLL:
    key = str(path)
    if key not in _cache:
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        _cache[key] = C.CDLL(key)
    return _cache[key]


def oracle_lib() -> C.CDLL:
    h = _load(ORACLE_LIB)
    h.orc_grid.restype = C.c_ulonglong
    h.orc_grid.argtypes = [C.c_int, *_GRID_ARGS, C.c_int]
    h.orc_degrid.restype = C.c_ulonglong
    h.orc_degrid.argtypes = [C.c_int, *_GRID_ARGS, C.c_int]
    h.orc_pregrid_many.restype = None
    h.orc_pregrid_many.argtypes = [C.c_double, C.c_double, vp, C.c_int, C.c_int,
                                   C.c_int, C.c_int, vp]
    h.orc_normalize.argtypes = [vp, C.c_int, C.c_int]
    h.orc_rotate.argtypes = [vp, vp, C.c_int, C.c_double]
    h.orc_grid0.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int]
    h.orc_reweight.argtypes = [vp, vp, C.c_double, C.c_int, C.c_int]
    h.orc_herm_padded.argtypes = [vp, C.c_int, C.c_int]
    h.orc_fftshift_even.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    h.orc_hogbom.restype = C.c_uint
    h.orc_hogbom.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_uint, C.c_double,
                             C.c_double]
    return h


def ref_available(fast: bool = False) -> bool:
    return (REF_FAST_LIB if fast else REF_LIB).exists()


def ref_lib(fast: bool = False) -> C.CDLL:
    h = _load(REF_FAST_LIB if fast else REF_LIB)
    for name in ("gridKernelCPUFullGCF", "gridKernelCPUHalfGCF",
                 "deGridKernelCPUFullGCF", "deGridKernelCPUHalfGCF"):
        fn = getattr(h, name)
        fn.restype = C.c_ulonglong
        fn.argtypes = _GRID_ARGS
    h.normalizeCPU.argtypes = [vp, C.c_int, C.c_int]
    h.rotateCPU.argtypes = [vp, vp, C.c_int, C.c_double]
    h.grid0.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int]
    h.reweight.argtypes = [vp, vp, C.c_double, C.c_int, C.c_int]
    return h


def _args(vc: VisChunks, gcf: GCFTable, uvw, vis, grid, size, pitch):
    supps = np.ascontiguousarray(vc.supps, np.int32)
    gcf_c, sup_c = gcf.pointer_tables()
    keep = (supps, vc.chunk_pointers(uvw), vc.chunk_pointers(vis))
    return keep, (vc.scale, vc.wstep, vc.n_chunks, supps.ctypes.data,
                  grid.ctypes.data, gcf_c, keep[1], keep[2], vc.ts_ch, pitch,
                  size, sup_c)


def aligned_zeros(shape, dtype, align: int = 64) -> np.ndarray:
    """Zeroed array whose data is `align`-byte aligned: the reference's
    addGrids writes the caller's grid with aligned 32-byte AVX stores
    (scatter_gridder_w_dependent_dyn_1p.cpp:33-42)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape))
    raw = np.zeros(n * dtype.itemsize + align, np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + n * dtype.itemsize].view(dtype).reshape(shape)


def _check128(gcf: GCFTable):
    if gcf.data.dtype != np.complex128:
        raise TypeError("checker takes complex128 GCF taps")


def grid(vc: VisChunks, gcf: GCFTable, size: int, pitch: int | None = None,
         impl: str = "oracle", nthreads: int = 1):
    """Grid with the restatement (impl='oracle') or the compiled reference
    (impl='ref' / 'ref_fast'; nthreads via OMP_NUM_THREADS of the process).
    Returns (grid[size, pitch] complex128, taps)."""
    _check128(gcf)
    pitch = pitch or size
    g = aligned_zeros((size, pitch), np.complex128)
    uvw = np.ascontiguousarray(vc.uvw, np.float64)
    vis = np.ascontiguousarray(vc.vis, np.complex128)
    keep, args = _args(vc, gcf, uvw, vis, g, size, pitch)
    if impl == "oracle":
        taps = oracle_lib().orc_grid(int(gcf.half), *args, nthreads)
    else:
        h = ref_lib(fast=(impl == "ref_fast"))
        fn = h.gridKernelCPUHalfGCF if gcf.half else h.gridKernelCPUFullGCF
        taps = fn(*args)
    del keep
    return g, int(taps)


def degrid(vc: VisChunks, gcf: GCFTable, grid_arr: np.ndarray,
           size: int | None = None, impl: str = "oracle", nthreads: int = 1):
    """Degrid; returns (vis[n_chunks, ts_ch] complex128, taps)."""
    _check128(gcf)
    g = np.ascontiguousarray(grid_arr, np.complex128)
    size = size or g.shape[0]
    pitch = g.shape[1]
    out = np.zeros((vc.n_chunks, vc.ts_ch), np.complex128)
    uvw = np.ascontiguousarray(vc.uvw, np.float64)
    keep, args = _args(vc, gcf, uvw, out, g, size, pitch)
    if impl == "oracle":
        taps = oracle_lib().orc_degrid(int(gcf.half), *args, nthreads)
    else:
        h = ref_lib(fast=(impl == "ref_fast"))
        fn = h.deGridKernelCPUHalfGCF if gcf.half else h.deGridKernelCPUFullGCF
        taps = fn(*args)
    del keep
    return out, int(taps)


def pregrid(vc: VisChunks, grid_size: int, half: bool) -> np.ndarray:
    """Per-point (u, v, over_u, over_v, w_plane, layer_index) int32 [n, 6];
    layer_index = INT32_MIN marks a point outside the int16 range."""
    h = oracle_lib()
    uvw = np.ascontiguousarray(vc.uvw, np.float64)
    out = np.zeros((vc.n_chunks, vc.ts_ch, 6), np.int32)
    for b in range(vc.n_chunks):
        h.orc_pregrid_many(vc.scale, vc.wstep, uvw[b].ctypes.data, vc.ts_ch,
                           int(vc.supps[b]), grid_size, int(half),
                           out[b].ctypes.data)
    return out.reshape(-1, 6)


def default_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def bin_points(uvw: np.ndarray, vis: np.ndarray, scale: float, wstep: float,
               gcf_supps: np.ndarray, max_w_plane: int, grid_size: int,
               mirror: bool = False):
    """orc_bin (the MS4 binner restated): returns (counts[32, 32], pre
    (total, 4) int16, vis (total,) complex128, src (total,) point index) in
    bin order, or None where the reference would fail its assert."""
    h = oracle_lib()
    h.orc_bin.restype = C.c_longlong
    h.orc_bin.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, vp, vp, vp, C.c_int,
                          vp, vp, vp, vp, C.c_longlong]
    uvw = np.ascontiguousarray(uvw, np.float64)
    vis = np.ascontiguousarray(vis, np.complex128)
    sup = np.ascontiguousarray(gcf_supps, np.int32)
    centre = sup.ctypes.data + 4 * max_w_plane
    counts = np.zeros(32 * 32, np.int32)
    args = (int(mirror), len(vis), scale, wstep, vis.ctypes.data, uvw.ctypes.data, centre,
            grid_size, counts.ctypes.data)
    total = h.orc_bin(*args, None, None, None, 0)
    if total < 0:
        return None
    pre = np.zeros((total, 4), np.int16)
    out = np.zeros(total, np.complex128)
    src = np.zeros(total, np.int32)
    h.orc_bin(*args, pre.ctypes.data, out.ctypes.data, src.ctypes.data, total)
    return counts.reshape(32, 32), pre, out, src
