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
         grid_pitch: int | None = None,
         out: np.ndarray | None = None) -> tuple[np.ndarray, int]:
    """gridKernelCPU{Full,Half}GCF: returns (grid[size, pitch] complex128,
    in-bound tap count).  out: a C-contiguous complex128 [size, pitch]
    array to grid into (overwritten), as a caller reuses its grid."""
    pitch = grid_pitch or grid_size
    if gcf.data.dtype != np.complex128:
        raise TypeError("the MS4 entry points take complex128 GCF taps")
    if out is None:
        g = np.empty((grid_size, pitch), np.complex128)
    elif out.dtype != np.complex128 or out.shape != (grid_size, pitch) or \
            not out.flags.c_contiguous:
        raise ValueError(f"out must be a C-contiguous complex128 ({grid_size}, {pitch}) array")
    else:
        g = out
    vis = np.ascontiguousarray(vc.vis, np.complex128)
    name = "gridKernelCPUHalfGCF" if gcf.half else "gridKernelCPUFullGCF"
    taps = _legacy(name, vc, gcf, g, vis, grid_size, pitch)
    return g, taps


def degrid(vc: VisChunks, gcf: GCFTable, grid_arr: np.ndarray,
           grid_size: int | None = None,
           out: np.ndarray | None = None) -> tuple[np.ndarray, int]:
    """deGridKernelCPU{Full,Half}GCF: returns (vis[n_chunks, ts_ch]
    complex128, in-bound tap count).  out: a C-contiguous complex128
    [n_chunks, ts_ch] array to write (overwritten)."""
    g = np.ascontiguousarray(grid_arr, np.complex128)
    size = grid_size or g.shape[0]
    pitch = g.shape[1]
    if gcf.data.dtype != np.complex128:
        raise TypeError("the MS4 entry points take complex128 GCF taps")
    if out is None:
        out = np.zeros((vc.n_chunks, vc.ts_ch), np.complex128)
    elif out.dtype != np.complex128 or out.shape != (vc.n_chunks, vc.ts_ch) or \
            not out.flags.c_contiguous:
        raise ValueError(f"out must be a C-contiguous complex128 ({vc.n_chunks}, {vc.ts_ch}) array")
    name = "deGridKernelCPUHalfGCF" if gcf.half else "deGridKernelCPUFullGCF"
    taps = _legacy(name, vc, gcf, g, out, size, pitch)
    return out, taps


def _raise_on_error(name: str) -> None:
    code, msg = _lib.last_error()
    if code:
        raise _lib.RcgError(f"{name}: {msg}")


def normalize(grid_arr: np.ndarray) -> None:
    """normalizeCPU in place on a (size, pitch) complex128 grid."""
    if grid_arr.dtype != np.complex128 or not grid_arr.flags.c_contiguous:
        raise TypeError("grid must be C-contiguous complex128")
    size, pitch = grid_arr.shape
    _lib.lib().normalizeCPU(grid_arr.ctypes.data, pitch, size)
    _raise_on_error("normalizeCPU")


def rotate(uvw: np.ndarray, vis: np.ndarray, scale: float) -> None:
    """rotateCPU in place: vis *= exp(2 pi i w scale)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    if vis.dtype != np.complex128 or not vis.flags.c_contiguous:
        raise TypeError("vis must be C-contiguous complex128")
    _lib.lib().rotateCPU(uvw.ctypes.data, vis.ctypes.data, vis.size, scale)
    _raise_on_error("rotateCPU")


def grid0(uvw: np.ndarray, vis: np.ndarray, grid_arr: np.ndarray, scale: float) -> None:
    """grid0: nearest-cell gridding added into grid_arr (size, pitch); the
    centring is done here as the reference's caller does
    (MS4/misc/cppcycle.cpp:93)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    vis = np.ascontiguousarray(vis, np.complex128)
    size, pitch = grid_arr.shape
    centre = grid_arr.ctypes.data + (size // 2 * (pitch + 1)) * 16
    _lib.lib().grid0(uvw.ctypes.data, vis.ctypes.data, centre, scale, vis.size, pitch, size)
    _raise_on_error("grid0")


def reweight(uvw: np.ndarray, vis: np.ndarray, scale: float, grid_size: int) -> None:
    """reweight in place (uniform weighting)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    if vis.dtype != np.complex128 or not vis.flags.c_contiguous:
        raise TypeError("vis must be C-contiguous complex128")
    _lib.lib().reweight(uvw.ctypes.data, vis.ctypes.data, scale, vis.size, grid_size)
    _raise_on_error("reweight")


def herm_padded_inplace(data: np.ndarray) -> None:
    size, pitch = data.shape
    _lib.lib().herm_padded_inplace(data.ctypes.data, size, pitch)
    _raise_on_error("herm_padded_inplace")


def fft_inplace_even(sign: int, data: np.ndarray, size: int, pitch: int) -> None:
    """fft_inplace_even on a host buffer of size*pitch complex128 (plan
    created and released per call)."""
    if data.dtype != np.complex128 or data.size < size * pitch:
        raise TypeError("data must hold size*pitch complex128")
    h = _lib.lib()
    plan = h.fft_inplace_even(None, sign, data.ctypes.data, size, pitch)
    code, msg = _lib.last_error()
    if plan:
        h.rcg_fft_destroy_plan(plan)
    if code:
        raise _lib.RcgError(f"fft_inplace_even: {msg}")


def deconvolve(res: np.ndarray, psf: np.ndarray, niters: int, gain: float,
               threshold: float) -> np.ndarray:
    """deconvolve (Hogbom CLEAN): updates res in place, returns the model."""
    if res.dtype != np.float64 or psf.dtype != np.float64:
        raise TypeError("res and psf must be float64")
    siz, pitch = res.shape
    model = np.zeros_like(res)
    _lib.lib().deconvolve(model.ctypes.data, res.ctypes.data, psf.ctypes.data, siz, pitch,
                          niters, gain, threshold)
    _raise_on_error("deconvolve")
    return model


class BinnedVis:
    """The block bin / binm return, copied out of its pinned host memory.

    counts[qu, qv] copies per bin; vis_off / pre_off[qu, qv] byte offsets
    from the end of the table (the reference's binTable,
    MS4/kernel/cpu/binner/binner.cpp:18-28); data_size the whole block's
    bytes; vis (total,) complex128 and pre (total, 4) int16 records
    (u, v, layer index, support) in bin order."""

    def __init__(self, table: np.ndarray, data: bytes):
        n = _BIN_DIVS * _BIN_DIVS
        self.data_size = int(table[0])
        self.data_off = int(table[1])
        self.vis_off = table[2:2 + n].reshape(_BIN_DIVS, _BIN_DIVS).copy()
        self.pre_off = table[2 + n:2 + 2 * n].reshape(_BIN_DIVS, _BIN_DIVS).copy()
        self.counts = table[2 + 2 * n:2 + 3 * n].reshape(_BIN_DIVS, _BIN_DIVS).copy()
        total = int(self.counts.sum())
        self.raw = data
        self.vis = np.frombuffer(data, np.complex128, count=total).copy()
        self.pre = np.frombuffer(data, np.int16, count=4 * total,
                                 offset=16 * total).reshape(total, 4).copy()

    def bin(self, qu: int, qv: int) -> tuple[np.ndarray, np.ndarray]:
        """(vis, pre) of one bin, read through the table's byte offsets."""
        n = int(self.counts[qu, qv])
        vo, po = int(self.vis_off[qu, qv]), int(self.pre_off[qu, qv])
        vis = np.frombuffer(self.raw, np.complex128, count=n, offset=vo)
        pre = np.frombuffer(self.raw, np.int16, count=4 * n, offset=po).reshape(n, 4)
        return vis, pre


_BIN_DIVS = 32
BIN_TABLE_BYTES = 12320


def bin_points(uvw: np.ndarray, vis: np.ndarray, scale: float, wstep: float,
               gcf_supps: np.ndarray, max_w_plane: int, grid_size: int,
               mirror: bool = False) -> BinnedVis:
    """bin (mirror False) / binm (True) over host arrays: uvw (n, 3) float64,
    vis (n,) complex128, gcf_supps the 2*max_w_plane + 1 supports of planes
    -max_w_plane..max_w_plane (the callee indexes its centre)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    vis = np.ascontiguousarray(vis, np.complex128)
    sup = np.ascontiguousarray(gcf_supps, np.int32)
    if uvw.shape != (len(vis), 3) or len(sup) != 2 * max_w_plane + 1:
        raise ValueError("uvw must be (n, 3), gcf_supps 2*max_w_plane + 1 long")
    h = _lib.lib()
    fn = h.binm if mirror else h.bin
    centre = sup.ctypes.data + 4 * max_w_plane
    p = fn(len(vis), scale, wstep, vis.ctypes.data, uvw.ctypes.data, centre, grid_size)
    if not p:
        code, msg = _lib.last_error()
        raise _lib.RcgError(f"{'binm' if mirror else 'bin'}: {msg}")
    try:
        base = p - BIN_TABLE_BYTES
        table = np.frombuffer(C.string_at(base, BIN_TABLE_BYTES), np.int32)
        size = int(table[0])
        data = C.string_at(p, size - BIN_TABLE_BYTES)
    finally:
        h.rcg_bin_free(p)
    return BinnedVis(table, data)
