"""Replays of the reference callers' calls of gridKernelCPUFullGCF and
deGridKernelCPUFullGCF: the same pointer tables and arguments, built the way
each caller builds them (test infrastructure).

* cppcycle (MS4/misc/cppcycle.cpp:96-223, with mkBlWpMap of
  MS4/misc/stats_n_utils.cpp:7-35): w-planes -M..M with
  M = max(round(maxw / wstep), round(-minw / wstep)) (C round, half away
  from zero); layer supports lsize = min(max, min + growth |p|); the GCF data
  of every plane's 64 oversampled layers one after the other and a table of
  64 (2M + 1) layer pointers passed at entry 64 M ("Center at 0"); chunk
  support lsize(wp), wp = round(maxw / wstep) for an all-negative baseline,
  round(minw / wstep) for an all-positive one, else 0; (pitch, size) in
  that order.
* gridWrapper (MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:64-85, table of
  Kernel/CPU/GCF.hs:40-100, baselineMinWPlane of Data.hs:259-263): the same
  layer data and table, with M from Haskell's round (half to even); the
  table passed at entry tsiz / 2 = 64 M + 32, 32 entries (4 oversampled
  rows) past the centre; the supports list over -tsiz/2 .. tsiz/2, centred;
  chunk support size(baselineMinWPlane) = the support of the baseline's
  smallest |w| plane; and the grid's (width, pitch) passed as
  (grid_pitch, grid_size).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

OVER2 = 64


def c_round(x: float) -> int:
    """C's round(): halves away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def hs_round(x: float) -> int:
    """Haskell's round: halves to the even neighbour."""
    return int(round(x))


def support_rule(min_size: int, growth: int, max_size: int):
    """min(max, min + growth |p|) (cppcycle.cpp:101-103, GCF.hs:89-90,
    ScatterGrid.hs:79)."""
    return lambda p: min(max_size, min_size + growth * abs(p))


@dataclass
class Layers:
    """GCF layers of planes -M..M, 64 per plane, one after the other
    (cppcycle.cpp:150-172, GCF.hs:53-62), plus a zero pad past the end so
    that reads through shifted pointers stay inside the allocation."""
    M: int
    supp: object
    data: np.ndarray
    ptrs: np.ndarray = field(init=False)

    def __post_init__(self):
        base = self.data.ctypes.data
        offs, off = [], 0
        for p in range(-self.M, self.M + 1):
            S = self.supp(p)
            for _ in range(OVER2):
                offs.append(off)
                off += S * S
        # one more plane's worth of entries, all on the pad
        offs += [off] * OVER2
        self.ptrs = np.uint64(base) + np.asarray(offs, np.uint64) * np.uint64(16)

    @property
    def tsiz(self) -> int:
        return OVER2 * (2 * self.M + 1)

    def table_at(self, entry: int) -> int:
        return int(self.ptrs.ctypes.data + 8 * entry)


def make_layers(M: int, supp, rng=None, encoded: bool = False) -> Layers:
    """Random taps, or (encoded) tap(entry, r, c) = (entry + 100000) +
    i (r 1000 + c), entry = the layer's index in the table, so that a
    gridded value names the layer it was read from."""
    parts = []
    entry = 0
    for p in range(-M, M + 1):
        S = supp(p)
        for _ in range(OVER2):
            if encoded:
                r = np.arange(S)[:, None]
                c = np.arange(S)[None, :]
                parts.append(((entry + 100000) + 1j * (r * 1000 + c)).reshape(-1))
            else:
                parts.append(rng.standard_normal(S * S) + 1j * rng.standard_normal(S * S))
            entry += 1
    smax = max(supp(p) for p in range(-M, M + 1))
    parts.append(np.zeros(OVER2 * smax * smax + 1024))
    return Layers(M, supp, np.ascontiguousarray(np.concatenate(parts), np.complex128))


@dataclass
class Call:
    """One call's arguments; the arrays behind the pointers live as long as
    this object."""
    scale: float
    wstep: float
    chunk_supps: np.ndarray
    table: int
    gcf_supps: np.ndarray
    supps_entry: int
    uvw: np.ndarray
    vis: np.ndarray
    ts_ch: int
    pitch_arg: int
    size_arg: int
    rows: int            # grid rows the caller allocated
    row_len: int         # ... and their length
    layers: Layers

    def _tabs(self, vis):
        nb = self.uvw.shape[0]
        up = np.uint64(self.uvw.ctypes.data) + np.arange(nb, dtype=np.uint64) * np.uint64(
            self.uvw.strides[0])
        vpt = np.uint64(vis.ctypes.data) + np.arange(nb, dtype=np.uint64) * np.uint64(
            vis.strides[0])
        return up, vpt

    def args(self, grid: np.ndarray, vis: np.ndarray):
        up, vpt = self._tabs(vis)
        self._keep = (up, vpt)
        return (self.scale, self.wstep, self.uvw.shape[0], self.chunk_supps.ctypes.data,
                grid.ctypes.data, self.table, up.ctypes.data, vpt.ctypes.data, self.ts_ch,
                self.pitch_arg, self.size_arg,
                self.gcf_supps.ctypes.data + 4 * self.supps_entry)

    def grid(self, fn) -> tuple[np.ndarray, int]:
        """fn(*args) -> taps; the grid as the caller allocated it."""
        g = aligned_zeros((self.rows, self.row_len))
        taps = fn(*self.args(g, self.vis))
        return g, int(taps)

    def degrid(self, fn, g: np.ndarray) -> tuple[np.ndarray, int]:
        out = np.zeros_like(self.vis)
        g = np.ascontiguousarray(g, np.complex128)
        taps = fn(*self.args(g, out))
        return out, int(taps)


def aligned_zeros(shape, align: int = 64) -> np.ndarray:
    n = int(np.prod(shape))
    raw = np.zeros(n * 16 + align, np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + n * 16].view(np.complex128).reshape(shape)


def cppcycle_call(uvw, vis, wstep, supp, grid_size, grid_pad, layers: Layers,
                  scale: float | None = None) -> Call:
    """cppcycle.cpp:96-223 (scale: (gridSize/2 - 1) / maxx * 16 as there,
    unless given)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    vis = np.ascontiguousarray(vis, np.complex128)
    M = layers.M
    if scale is None:
        maxx = max(np.abs(uvw[..., 0]).max(), np.abs(uvw[..., 1]).max())
        scale = (grid_size // 2 - 1) / maxx * 16.0
    wp = []
    for bl in range(uvw.shape[0]):
        cmax = uvw[bl, :, 2].max() / wstep
        cmin = uvw[bl, :, 2].min() / wstep
        wp.append(c_round(cmax) if cmax < 0 else c_round(cmin) if cmin > 0 else 0)
    chunk = np.asarray([supp(p) for p in wp], np.int32)
    gsup = np.asarray([supp(p) for p in range(-M, M + 1)], np.int32)
    pitch = grid_size + grid_pad
    return Call(scale, wstep, chunk, layers.table_at(OVER2 * M), gsup, M, uvw, vis,
                uvw.shape[1], pitch, grid_size, grid_size, pitch, layers)


def wrapper_call(uvw, vis, wstep, supp, width, pitch, scale, layers: Layers) -> Call:
    """ScatterGrid.hs:64-85 on a square GridPar (width = height)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    vis = np.ascontiguousarray(vis, np.complex128)
    chunk = []
    for bl in range(uvw.shape[0]):
        mn, mx = uvw[bl, :, 2].min(), uvw[bl, :, 2].max()
        p = hs_round(mn / wstep) if mn > 0 else hs_round(-mx / wstep) if mx < 0 else 0
        chunk.append(supp(p))
    half = layers.tsiz // 2
    gsup = np.asarray([supp(i) for i in range(-half, half + 1)], np.int32)
    return Call(scale, wstep, np.asarray(chunk, np.int32), layers.table_at(half), gsup, half,
                uvw, vis, uvw.shape[1], width, pitch, width, pitch, layers)


def haskell_max_plane(uvw, wstep) -> int:
    """GCF.hs:91: max(round(maxW / wstep), round(-minW / wstep))."""
    return max(hs_round(uvw[..., 2].max() / wstep), hs_round(-uvw[..., 2].min() / wstep))


def cppcycle_max_plane(uvw, wstep) -> int:
    """cppcycle.cpp:99."""
    return max(c_round(uvw[..., 2].max() / wstep), c_round(-uvw[..., 2].min() / wstep))


def baselines(seed: int, n_bl: int, ts: int, wstep: float, w_planes: float,
              lim: float, far: float = 0.0):
    """n_bl baselines of ts points: u, v mostly in [-lim, lim] (a few out to
    `far`), w drifting along each baseline within +-w_planes wstep, some
    baselines crossing w = 0; halves of wstep on some baselines' ends."""
    rng = np.random.default_rng(seed)
    uvw = np.zeros((n_bl, ts, 3))
    t = np.linspace(0.0, 1.0, ts)
    for b in range(n_bl):
        u0, v0 = rng.uniform(-lim, lim, 2)
        du, dv = rng.uniform(-lim / 4, lim / 4, 2)
        uvw[b, :, 0] = np.clip(u0 + du * t, -lim, lim)
        uvw[b, :, 1] = np.clip(v0 + dv * t, -lim, lim)
        w0 = rng.uniform(-w_planes, w_planes) * wstep
        dw = rng.uniform(-0.8, 0.8) * wstep
        w = np.clip(w0 + dw * t, -w_planes * wstep, w_planes * wstep)
        if b % 5 == 3:       # ends on a half plane
            w[-1] = (np.trunc(w[-1] / wstep) + math.copysign(0.5, w[-1] or 1.0)) * wstep
            w[-1] = float(np.clip(w[-1], -w_planes * wstep, w_planes * wstep))
        uvw[b, :, 2] = w
    if far > 0:
        k = rng.random((n_bl, ts)) < 0.05
        uvw[..., 0] = np.where(k, rng.uniform(-far, far, (n_bl, ts)), uvw[..., 0])
    vis = rng.standard_normal((n_bl, ts)) + 1j * rng.standard_normal((n_bl, ts))
    return uvw, vis
