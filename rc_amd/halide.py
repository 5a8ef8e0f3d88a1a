"""ctypes view of the MS6 Halide buffer_t ABI (include/rcgrid_halide.h) and
calls of kern_scatter / kern_degrid / kern_scatter_gpu1 / kern_degrid_gpu1 /
nvGridder, and of the imaging-glue kernels kern_init / kern_image_init /
kern_psf_vis / kern_rotate / kern_detile / kern_defacet / kern_image_sum.

Array conventions (the MS6 generators' buffer layouts):
  vis  float64 [n, 5]      (u, v, w, re, im)
  gcf  float64 [64, S, S, 2]  [over][y][x][re|im], over = ou + 8*ov
  uvg  float64 [Y, P, 2]   [y][x][re|im], P >= X columns; buffer min = (x0, y0)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class BufferT(C.Structure):
    _fields_ = [("dev", C.c_uint64), ("host", C.c_void_p), ("extent", C.c_int32 * 4),
                ("stride", C.c_int32 * 4), ("min", C.c_int32 * 4), ("elem_size", C.c_int32),
                ("host_dirty", C.c_bool), ("dev_dirty", C.c_bool),
                ("_padding", C.c_uint8 * (10 - C.sizeof(C.c_void_p)))]


assert C.sizeof(BufferT) == 72


def _buf(arr: np.ndarray | None, extent, stride, mins=(0, 0, 0, 0), dev: int = 0,
         host_dirty: bool = False) -> BufferT:
    b = BufferT()
    b.dev = dev
    b.host = arr.ctypes.data if arr is not None else None
    for i in range(4):
        b.extent[i] = extent[i] if i < len(extent) else 0
        b.stride[i] = stride[i] if i < len(stride) else 0
        b.min[i] = mins[i] if i < len(mins) else 0
    b.elem_size = 8
    b.host_dirty = host_dirty
    return b


def vis_buffer(vis: np.ndarray, first: int = 0, dev: int = 0, host_dirty=False) -> BufferT:
    assert vis.dtype == np.float64 and vis.ndim == 2 and vis.shape[1] == 5
    assert vis.flags.c_contiguous
    return _buf(vis, (5, vis.shape[0]), (1, 5), (0, first), dev, host_dirty)


def gcf_buffer(gcf: np.ndarray, dev: int = 0, host_dirty=False) -> BufferT:
    assert gcf.dtype == np.float64 and gcf.ndim == 4 and gcf.shape[0] == 64
    assert gcf.shape[3] == 2 and gcf.shape[1] == gcf.shape[2] and gcf.flags.c_contiguous
    S = gcf.shape[1]
    return _buf(gcf, (2, S, S, 64), (1, 2, 2 * S, 2 * S * S), (0, 0, 0, 0), dev, host_dirty)


def uvg_buffer(uvg: np.ndarray, width: int | None = None, origin=(0, 0), dev: int = 0,
               host_dirty=False) -> BufferT:
    assert uvg.dtype == np.float64 and uvg.ndim == 3 and uvg.shape[2] == 2
    assert uvg.flags.c_contiguous
    Y, P = uvg.shape[:2]
    X = width or P
    return _buf(uvg, (2, X, Y), (1, 2, 2 * P), (0, origin[0], origin[1]), dev, host_dirty)


def scatter(scale, grid_size, margin, vis_b, gcf_b, uvg_b, gpu1: bool = False) -> int:
    h = _lib.lib()
    fn = h.kern_scatter_gpu1 if gpu1 else h.kern_scatter
    return int(fn(scale, grid_size, margin, C.addressof(vis_b), C.addressof(gcf_b),
                  C.addressof(uvg_b)))


def degrid(scale, grid_size, margin, gcf_b, uvg_b, vis_b, out_b, gpu1: bool = False) -> int:
    h = _lib.lib()
    fn = h.kern_degrid_gpu1 if gpu1 else h.kern_degrid
    return int(fn(scale, grid_size, margin, C.addressof(gcf_b), C.addressof(uvg_b),
                  C.addressof(vis_b), C.addressof(out_b)))


def nv_gcf_buffer(gcf: np.ndarray) -> BufferT:
    """nvGridder's GCF: float64 [64, 16, 16, 2], [8*sub_y + sub_x][b][a]."""
    assert gcf.dtype == np.float64 and gcf.shape == (64, 16, 16, 2) and gcf.flags.c_contiguous
    return _buf(gcf, (2, 16, 16, 64), (1, 2, 32, 512))


def nv_gridder(scale, grid_size, vis_b, gcf_b, uvg_b) -> int:
    """nvGridder: sorts the vis buffer's host array in place and overwrites
    the first grid_size^2 complex cells of the uvg buffer's host array."""
    return int(_lib.lib().nvGridder(scale, grid_size, C.addressof(vis_b), C.addressof(gcf_b),
                                    C.addressof(uvg_b)))


def bin_sort(recs: np.ndarray) -> None:
    """bin_sort3 / bin_sort5: nvGridder's in-place sort of float64 [n, 3] or
    [n, 5] records."""
    assert recs.dtype == np.float64 and recs.ndim == 2 and recs.shape[1] in (3, 5)
    assert recs.flags.c_contiguous
    fn = _lib.lib().bin_sort3 if recs.shape[1] == 3 else _lib.lib().bin_sort5
    fn(recs.ctypes.data, recs.shape[0])


def grid_gpu(scale: float, gcf: np.ndarray, img_dim: int, combined: np.ndarray | None = None,
             pos: np.ndarray | None = None, vals: np.ndarray | None = None) -> np.ndarray:
    """gridGPUc (combined [n, 5]) or gridGPUs (pos [n, 3], vals [n, 2]):
    returns the img_dim x img_dim x 2 grid (x fast)."""
    assert gcf.dtype == np.float64 and gcf.shape == (64, 16, 16, 2) and gcf.flags.c_contiguous
    out = np.zeros((img_dim, img_dim, 2))
    h = _lib.lib()
    if combined is not None:
        assert combined.dtype == np.float64 and combined.shape[1] == 5
        assert combined.flags.c_contiguous
        h.gridGPUc(scale, out.ctypes.data, combined.ctypes.data, combined.shape[0], img_dim,
                   gcf.ctypes.data)
    else:
        assert pos.dtype == np.float64 and pos.shape[1] == 3 and pos.flags.c_contiguous
        assert vals.dtype == np.float64 and vals.shape == (pos.shape[0], 2)
        assert vals.flags.c_contiguous
        h.gridGPUs(scale, out.ctypes.data, pos.ctypes.data, vals.ctypes.data, pos.shape[0],
                   img_dim, gcf.ctypes.data)
    code, msg = _lib.last_error()
    if code:
        raise _lib.RcgError(f"gridGPU failed: {msg}")
    return out


def buffer_of(arr: np.ndarray, mins=(0, 0, 0, 0)) -> BufferT:
    """A buffer_t over a float64 array whose last index is Halide dimension 0
    (shape [..., x1, x0]); mins in Halide order (x0 first)."""
    assert arr.dtype == np.float64 and arr.ndim <= 4
    ext = arr.shape[::-1]
    st = tuple(s // 8 for s in arr.strides[::-1])
    return _buf(arr, ext, st, tuple(mins) + (0,) * (4 - len(mins)))


def _call(name: str, *args) -> int:
    conv = [C.addressof(a) if isinstance(a, BufferT) else a for a in args]
    return int(getattr(_lib.lib(), name)(*conv))


def init(b: BufferT) -> int:
    return _call("kern_init", b)


def image_init(b: BufferT) -> int:
    return _call("kern_image_init", b)


def psf_vis(vis_b: BufferT, out_b: BufferT) -> int:
    return _call("kern_psf_vis", vis_b, out_b)


def rotate(in_lon, in_lat, out_lon0, out_lat0, lon_incr, lat_incr, uvproj, vis_b: BufferT,
           out_b: BufferT) -> int:
    return _call("kern_rotate", in_lon, in_lat, out_lon0, out_lat0, lon_incr, lat_incr,
                 int(uvproj), vis_b, out_b)


def detile(in_b: BufferT, out_b: BufferT) -> int:
    return _call("kern_detile", in_b, out_b)


def defacet(facet_b: BufferT, image_b: BufferT) -> int:
    return _call("kern_defacet", facet_b, image_b)


def image_sum(image_b: BufferT, sum_b: BufferT) -> int:
    return _call("kern_image_sum", image_b, sum_b)
