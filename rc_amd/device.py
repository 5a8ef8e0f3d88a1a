"""Device-resident data (torch tensors on the HIP device) and the rcg_* calls.

PyTorch provides the device memory and the stream; every computation is a
librcgrid.so kernel.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .tables import GCFTable, VisChunks

PG_DTYPE = np.dtype([("cell", "<i4"), ("tap0", "<i4"), ("stride", "<i2"),
                     ("supp", "<i2"), ("flags", "<i4")])
PG_CONJ = 1
PG_RANGE_ERR = 2


def _stream(stream: torch.cuda.Stream | None) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _complex_tensor(a: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


@dataclass
class DeviceGCF:
    """GCF table on the device (struct rcg_gcf)."""
    data: torch.Tensor
    layer_off: torch.Tensor
    layer_len: torch.Tensor
    supps: torch.Tensor
    center: int
    max_w_plane: int
    half: bool

    @staticmethod
    def from_table(t: GCFTable, device="cuda", dtype=None) -> "DeviceGCF":
        data = t.data if dtype is None else t.data.astype(dtype)
        return DeviceGCF(
            data=_complex_tensor(data, device),
            layer_off=torch.from_numpy(t.layer_off.astype(np.int64)).to(device),
            layer_len=torch.from_numpy(t.layer_len.astype(np.int32)).to(device),
            supps=torch.from_numpy(t.supps.astype(np.int32)).to(device),
            center=t.center, max_w_plane=t.max_w_plane, half=t.half)

    def struct(self) -> _lib.Gcf:
        return _lib.Gcf(self.data.data_ptr(), self.layer_off.data_ptr(),
                        self.layer_len.data_ptr(), int(self.layer_off.numel()),
                        self.center, self.supps.data_ptr(), self.max_w_plane,
                        int(self.half), int(self.data.numel()))


@dataclass
class DeviceGeom:
    """Visibility geometry on the device (struct rcg_geom)."""
    uvw: torch.Tensor          # float64 [n, 3]
    chunk_supps: torch.Tensor  # int32 [n_chunks]
    n_chunks: int
    ts_ch: int
    scale: float
    wstep: float
    grid_size: int
    grid_pitch: int

    @property
    def n(self) -> int:
        return self.n_chunks * self.ts_ch

    @staticmethod
    def from_chunks(vc: VisChunks, grid_size: int, grid_pitch: int | None = None,
                    device="cuda") -> "DeviceGeom":
        return DeviceGeom(
            uvw=torch.from_numpy(np.ascontiguousarray(vc.uvw.reshape(-1, 3), np.float64)).to(device),
            chunk_supps=torch.from_numpy(np.ascontiguousarray(vc.supps, np.int32)).to(device),
            n_chunks=vc.n_chunks, ts_ch=vc.ts_ch, scale=float(vc.scale),
            wstep=float(vc.wstep), grid_size=grid_size,
            grid_pitch=grid_pitch or grid_size)

    def struct(self) -> _lib.Geom:
        return _lib.Geom(self.uvw.data_ptr(), self.chunk_supps.data_ptr(),
                         self.n_chunks, self.ts_ch, self.scale, self.wstep,
                         self.grid_size, self.grid_pitch)


def pregrid(geom: DeviceGeom, gcf: DeviceGCF, stream=None):
    """rcg_pregrid: returns (records uint8 tensor [n*16], taps tensor u64[1]
    as int64, error word int32[1])."""
    dev = geom.uvw.device
    pg = torch.empty(geom.n * 16, dtype=torch.uint8, device=dev)
    taps = torch.zeros(1, dtype=torch.int64, device=dev)
    err = torch.zeros(1, dtype=torch.int32, device=dev)
    g, t = geom.struct(), gcf.struct()
    _lib.check(_lib.lib().rcg_pregrid(C.byref(g), C.byref(t), pg.data_ptr(),
                                      taps.data_ptr(), err.data_ptr(),
                                      _stream(stream)), "rcg_pregrid")
    return pg, taps, err


def records_to_numpy(pg: torch.Tensor) -> np.ndarray:
    return pg.cpu().numpy().view(PG_DTYPE)


def grid_plain_f64(geom: DeviceGeom, gcf: DeviceGCF, pg: torch.Tensor,
                   vis: torch.Tensor, grid: torch.Tensor | None = None,
                   stream=None) -> torch.Tensor:
    if vis.dtype != torch.complex128 or gcf.data.dtype != torch.complex128:
        raise TypeError("FP64 path takes complex128 visibilities and taps")
    if grid is None:
        grid = torch.empty((geom.grid_size, geom.grid_pitch), dtype=torch.complex128,
                           device=vis.device)
    g, t = geom.struct(), gcf.struct()
    _lib.check(_lib.lib().rcg_grid_plain_f64(C.byref(g), C.byref(t), pg.data_ptr(),
                                             vis.data_ptr(), grid.data_ptr(),
                                             _stream(stream)), "rcg_grid_plain_f64")
    return grid


def degrid_plain_f64(geom: DeviceGeom, gcf: DeviceGCF, pg: torch.Tensor,
                     grid: torch.Tensor, vis: torch.Tensor | None = None,
                     stream=None) -> torch.Tensor:
    if grid.dtype != torch.complex128 or gcf.data.dtype != torch.complex128:
        raise TypeError("FP64 path takes a complex128 grid and taps")
    if vis is None:
        vis = torch.empty(geom.n, dtype=torch.complex128, device=grid.device)
    g, t = geom.struct(), gcf.struct()
    _lib.check(_lib.lib().rcg_degrid_plain_f64(C.byref(g), C.byref(t), pg.data_ptr(),
                                               grid.data_ptr(), vis.data_ptr(),
                                               _stream(stream)), "rcg_degrid_plain_f64")
    return vis


STAGE_BIN, STAGE_GRID, STAGE_DEGRID = 1, 2, 4


class TiledPlan(_lib.Handle):
    """A tiled FP32 gridding plan (rcg_tiled_plan): pregrid + uv-tile bins for
    one geometry or a batch of instances (a list of geometries with one grid
    shape), reused by grid and degrid.  For a batch, visibilities are the
    instances' points concatenated in order and the grid tensor is
    [n_inst, grid_size, grid_pitch].  Device scratch is owned by the library
    and released by close()."""

    LAYOUTS = {"pairs": 0, "halo": 1, "halo16": 2}

    def __init__(self, geom, tile: int = 32, chunk: int = 512,
                 max_support: int | None = None, layout: str = "pairs", waves: int = 0,
                 degrid_waves: int = 0):
        geoms = list(geom) if isinstance(geom, (list, tuple)) else [geom]
        self.batch = isinstance(geom, (list, tuple))
        self.geoms = geoms
        self.geom = geoms[0]
        self.n = sum(g.n for g in geoms)
        self.n_inst = len(geoms)
        if max_support is None:
            max_support = max((int(g.chunk_supps.max().item()) for g in geoms if g.n_chunks),
                              default=1)
        self.tile, self.chunk, self.max_support = tile, chunk, max(1, max_support)
        if layout not in self.LAYOUTS:
            raise ValueError(f"layout must be one of {sorted(self.LAYOUTS)}")
        self.layout = layout
        h = C.c_void_p()
        arr = (_lib.Geom * len(geoms))(*[g.struct() for g in geoms])
        _lib.check(_lib.lib().rcg_tiled_plan_create_ex(arr, len(geoms), tile, chunk,
                                                       self.max_support, self.LAYOUTS[layout],
                                                       C.byref(h)),
                   "rcg_tiled_plan_create")
        self._own(h)
        self.waves = 0
        self.degrid_waves = 0
        if waves:
            self.set_grid_waves(waves)
        if degrid_waves:
            self.set_degrid_waves(degrid_waves)

    def set_degrid_waves(self, waves: int) -> None:
        """Wavefronts per degridding workgroup, 2, 4 or 8; 0 restores the
        default 4 (rcg_tiled_set_degrid_waves)."""
        _lib.check(_lib.lib().rcg_tiled_set_degrid_waves(self._h, int(waves)),
                   "rcg_tiled_set_degrid_waves")
        self.degrid_waves = int(waves)

    def set_grid_waves(self, waves: int) -> None:
        """Wavefronts (private tile copies) per gridding workgroup, 1..4;
        0 restores the default (rcg_tiled_set_grid_waves)."""
        _lib.check(_lib.lib().rcg_tiled_set_grid_waves(self._h, int(waves)),
                   "rcg_tiled_set_grid_waves")
        self.waves = int(waves)

    @property
    def grid_shape(self) -> tuple:
        g = (self.geom.grid_size, self.geom.grid_pitch)
        return (self.n_inst,) + g if self.batch else g

    DESTROY = "rcg_tiled_plan_destroy"

    def bin(self, gcf: DeviceGCF, stream=None) -> None:
        t = gcf.struct()
        _lib.check(_lib.lib().rcg_tiled_bin(self._h, C.byref(t), _stream(stream)),
                   "rcg_tiled_bin")

    def _check_vis(self, vis: torch.Tensor) -> None:
        if vis.dtype != torch.complex64:
            raise TypeError("FP32 path takes complex64 visibilities")
        if vis.numel() != self.n or not vis.is_contiguous():
            raise ValueError("vis must be a contiguous tensor of n points")

    def _check_grid(self, grid: torch.Tensor) -> None:
        if grid.dtype != torch.complex64:
            raise TypeError("FP32 path takes a complex64 grid")
        if tuple(grid.shape) != self.grid_shape or not grid.is_contiguous():
            raise ValueError(f"grid must be a contiguous {self.grid_shape} tensor")

    def grid(self, gcf: DeviceGCF, vis: torch.Tensor, grid: torch.Tensor | None = None,
             stream=None, adjoint: bool = False) -> torch.Tensor:
        """Gridding; adjoint=True conjugates every tap, which makes it the
        Hermitian adjoint of degrid()."""
        if gcf.data.dtype != torch.complex64:
            raise TypeError("FP32 path takes complex64 taps")
        self._check_vis(vis)
        if grid is None:
            grid = torch.empty(self.grid_shape, dtype=torch.complex64, device=vis.device)
        self._check_grid(grid)
        t = gcf.struct()
        fn = _lib.lib().rcg_grid_tiled_adjoint_f32 if adjoint else _lib.lib().rcg_grid_tiled_f32
        _lib.check(fn(self._h, C.byref(t), vis.data_ptr(), grid.data_ptr(), _stream(stream)),
                   "rcg_grid_tiled_f32")
        return grid

    def grid_f64(self, gcf: DeviceGCF, vis: torch.Tensor, grid: torch.Tensor | None = None,
                 stream=None) -> torch.Tensor:
        """FP64 gridding over the plan's pairs (rcg_grid_tiled_f64): complex128
        taps (the plan binned with a table of the same layout), visibilities
        and grid."""
        if gcf.data.dtype != torch.complex128:
            raise TypeError("FP64 gridding takes complex128 taps")
        if vis.dtype != torch.complex128:
            raise TypeError("FP64 gridding takes complex128 visibilities")
        if vis.numel() != self.n or not vis.is_contiguous():
            raise ValueError("vis must be a contiguous tensor of n points")
        if grid is None:
            grid = torch.empty(self.grid_shape, dtype=torch.complex128, device=vis.device)
        if grid.dtype != torch.complex128 or tuple(grid.shape) != self.grid_shape or \
                not grid.is_contiguous():
            raise ValueError(f"grid must be a contiguous complex128 {self.grid_shape} tensor")
        t = gcf.struct()
        _lib.check(_lib.lib().rcg_grid_tiled_f64(self._h, C.byref(t), vis.data_ptr(),
                                                 grid.data_ptr(), _stream(stream)),
                   "rcg_grid_tiled_f64")
        return grid

    def degrid_f64(self, gcf: DeviceGCF, grid: torch.Tensor, vis: torch.Tensor | None = None,
                   stream=None) -> torch.Tensor:
        """FP64 degridding over the plan's pairs (rcg_degrid_tiled_f64):
        complex128 taps, grid and visibilities."""
        if gcf.data.dtype != torch.complex128:
            raise TypeError("FP64 degridding takes complex128 taps")
        if grid.dtype != torch.complex128 or tuple(grid.shape) != self.grid_shape or \
                not grid.is_contiguous():
            raise ValueError(f"grid must be a contiguous complex128 {self.grid_shape} tensor")
        if vis is None:
            vis = torch.empty(self.n, dtype=torch.complex128, device=grid.device)
        if vis.dtype != torch.complex128 or vis.numel() != self.n or not vis.is_contiguous():
            raise ValueError("vis must be a contiguous complex128 tensor of n points")
        t = gcf.struct()
        _lib.check(_lib.lib().rcg_degrid_tiled_f64(self._h, C.byref(t), grid.data_ptr(),
                                                   vis.data_ptr(), _stream(stream)),
                   "rcg_degrid_tiled_f64")
        return vis

    def degrid(self, gcf: DeviceGCF, grid: torch.Tensor, vis: torch.Tensor | None = None,
               stream=None) -> torch.Tensor:
        if gcf.data.dtype != torch.complex64:
            raise TypeError("FP32 path takes complex64 taps")
        self._check_grid(grid)
        if vis is None:
            vis = torch.empty(self.n, dtype=torch.complex64, device=grid.device)
        self._check_vis(vis)
        t = gcf.struct()
        _lib.check(_lib.lib().rcg_degrid_tiled_f32(self._h, C.byref(t), grid.data_ptr(),
                                                   vis.data_ptr(), _stream(stream)),
                   "rcg_degrid_tiled_f32")
        return vis

    def stats(self, stream=None) -> dict:
        taps = C.c_ulonglong()
        pairs = C.c_longlong()
        items = C.c_longlong()
        err = C.c_int()
        _lib.check(_lib.lib().rcg_tiled_stats(self._h, C.byref(taps), C.byref(pairs),
                                              C.byref(items), C.byref(err),
                                              _stream(stream)), "rcg_tiled_stats")
        return {"taps": taps.value, "pairs": pairs.value, "items": items.value,
                "err": err.value}

    def compensated(self, enable: bool) -> None:
        """Compensated (TwoSum) tile flushes for gridding."""
        _lib.check(_lib.lib().rcg_tiled_compensated(self._h, int(enable)),
                   "rcg_tiled_compensated")

    def timing(self, enable: bool) -> None:
        _lib.check(_lib.lib().rcg_tiled_timing(self._h, int(enable)), "rcg_tiled_timing")

    def kernel_times(self, which: str) -> list[float]:
        """Durations (ms) of the main grid ('grid') or degrid ('degrid')
        kernel of the calls made since timing was switched on (at most 64)."""
        buf = (C.c_float * 64)()
        n = C.c_int()
        _lib.check(_lib.lib().rcg_tiled_kernel_times(self._h, 0 if which == "grid" else 1,
                                                     buf, 64, C.byref(n)),
                   "rcg_tiled_kernel_times")
        return [buf[k] for k in range(n.value)]

    def capture(self, gcf: DeviceGCF, stages: int, vis: torch.Tensor | None = None,
                grid: torch.Tensor | None = None,
                vis_out: torch.Tensor | None = None) -> "Graph":
        """A HIP graph of the chosen stages (STAGE_* bits) on these buffers."""
        if vis is not None:
            self._check_vis(vis)
        if vis_out is not None:
            self._check_vis(vis_out)
        if grid is not None:
            self._check_grid(grid)
        t = gcf.struct()
        h = C.c_void_p()
        ptr = lambda x: x.data_ptr() if x is not None else None  # noqa: E731
        _lib.check(_lib.lib().rcg_tiled_graph_create(self._h, C.byref(t), ptr(vis), ptr(grid),
                                                     ptr(vis_out), stages, C.byref(h)),
                   "rcg_tiled_graph_create")
        return Graph(h, keep=(self, gcf, vis, grid, vis_out))


class Graph(_lib.Handle):
    """An executable HIP graph (rcg_graph); keeps its plan and buffers alive."""

    DESTROY = "rcg_graph_destroy"

    def __init__(self, h: C.c_void_p, keep=()):
        self._keep = keep
        self._own(h)

    def launch(self, stream=None) -> None:
        _lib.check(_lib.lib().rcg_graph_launch(self._h, _stream(stream)), "rcg_graph_launch")

    def _release(self) -> None:
        self._keep = ()


class FFTPlan(_lib.Handle):
    """rcg_fft plan: the reference's fft_inplace_even transform (sign -1/+1
    C2C, 0 R2C, 2 C2R, fftshift before and after, unnormalised) on a device
    buffer of `size` rows of `pitch` complex elements."""

    def __init__(self, sign: int, size: int, pitch: int | None = None, f64: bool = False):
        h = C.c_void_p()
        self.sign, self.size, self.pitch, self.f64 = sign, size, pitch or size, f64
        _lib.check(_lib.lib().rcg_fft_plan_create(sign, int(f64), size, self.pitch, C.byref(h)),
                   "rcg_fft_plan_create")
        self._own(h)

    def __call__(self, data: torch.Tensor, stream=None) -> torch.Tensor:
        want = torch.complex128 if self.f64 else torch.complex64
        if data.dtype != want or not data.is_contiguous() or \
                data.numel() < self.size * self.pitch:
            raise ValueError(f"FFT buffer must be contiguous {want} of size*pitch elements")
        _lib.check(_lib.lib().rcg_fft_execute(self._h, data.data_ptr(), _stream(stream)),
                   "rcg_fft_execute")
        return data

    DESTROY = "rcg_fft_plan_destroy"


def hogbom(res: torch.Tensor, psf: torch.Tensor, niters: int, gain: float, threshold: float,
           model: torch.Tensor | None = None, stream=None) -> tuple[torch.Tensor, int]:
    """Hogbom CLEAN on device images (rcg_hogbom): res (float32/float64
    [siz, pitch]) is cleaned in place; returns (model, accepted peaks)."""
    if res.dtype != psf.dtype or res.shape != psf.shape or res.dim() != 2 or \
            res.dtype not in (torch.float32, torch.float64):
        raise ValueError("res and psf must be 2-D float tensors of one shape and dtype")
    if not (res.is_contiguous() and psf.is_contiguous()):
        raise ValueError("res and psf must be contiguous")
    siz, pitch = res.shape
    if model is None:
        model = torch.empty_like(res)
    h = _lib.lib()
    scratch = torch.empty(int(h.rcg_hogbom_scratch_bytes()), dtype=torch.uint8,
                          device=res.device)
    s = _stream(stream)
    _lib.check(h.rcg_hogbom(model.data_ptr(), res.data_ptr(), psf.data_ptr(),
                            int(res.dtype == torch.float64), siz, pitch, niters, gain,
                            threshold, scratch.data_ptr(), s), "rcg_hogbom")
    it = C.c_uint()
    _lib.check(h.rcg_hogbom_iterations(scratch.data_ptr(), C.byref(it), s),
               "rcg_hogbom_iterations")
    return model, int(it.value)


def unpack_vis5(rec: torch.Tensor, f64: bool = False, rotate_scale: float | None = None,
                stream=None) -> tuple[torch.Tensor, torch.Tensor]:
    """MS6 five-double records [n, 5] (u, v, w, re, im) on the device ->
    (uvw float64 [n, 3], visibilities complex64/complex128 [n]), optionally
    rotated by exp(2 pi i w scale)."""
    if rec.dtype != torch.float64 or rec.dim() != 2 or rec.shape[1] != 5 or \
            not rec.is_contiguous():
        raise ValueError("records must be a contiguous float64 [n, 5] tensor")
    n = rec.shape[0]
    uvw = torch.empty((n, 3), dtype=torch.float64, device=rec.device)
    vis = torch.empty(n, dtype=torch.complex128 if f64 else torch.complex64, device=rec.device)
    rot = rotate_scale is not None
    _lib.check(_lib.lib().rcg_unpack_vis5(rec.data_ptr(), n, int(rot),
                                          float(rotate_scale or 0.0), uvw.data_ptr(),
                                          vis.data_ptr(), int(f64), _stream(stream)),
               "rcg_unpack_vis5")
    return uvw, vis


def pack_vis5(uvw: torch.Tensor, vis: torch.Tensor, stream=None) -> torch.Tensor:
    """(uvw [n, 3], vis [n]) -> MS6 records [n, 5] float64."""
    n = vis.numel()
    f64 = vis.dtype == torch.complex128
    rec = torch.empty((n, 5), dtype=torch.float64, device=vis.device)
    _lib.check(_lib.lib().rcg_pack_vis5(uvw.contiguous().data_ptr(), vis.contiguous().data_ptr(),
                                        int(f64), n, rec.data_ptr(), _stream(stream)),
               "rcg_pack_vis5")
    return rec
