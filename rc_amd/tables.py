"""Host-side data formats of the gridder: GCF tables and visibility chunks.

GCF layout (the reference's, MS4/kernel/cpu/gcf/GCF.cpp:110-114,
MS4/misc/cppcycle.cpp:149-172, 216-222): w-plane p owns 64 layers, one per
oversampling offset (over_u, over_v) in [0,8)^2, each an S_p x S_p block of
complex taps stored row-major (row = u offset, column = v offset).  Layer
index L = (p*8 + over_u)*8 + over_v; the pointer table and the support table
are addressed from their w_plane = 0 centre, so negative indices reach the
negative planes.  A half table stores planes 0..M only; the gridder uses
layer |L| and conjugates the taps of negative-w points
(scatter_gridder_w_dependent_dyn_1p.cpp:130-137, 159-164).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

OVER = 8  # MS4/kernel/common/metrix.h:5


@dataclass
class GCFTable:
    """A GCF table in host memory.

    data       complex taps, all layers concatenated (complex128 or complex64)
    layer_off  element offset of every layer slot (int64)
    layer_len  taps in every layer slot (int32)
    center     slot of layer index 0 (full table); 0 for a half table
    supps      support of plane p at supps[p + max_w_plane] (int32)
    """
    data: np.ndarray
    layer_off: np.ndarray
    layer_len: np.ndarray
    center: int
    supps: np.ndarray
    max_w_plane: int
    half: bool
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_slots(self) -> int:
        return int(self.layer_off.shape[0])

    @staticmethod
    def from_planes(planes: dict[int, np.ndarray], half: bool,
                    dtype=np.complex128) -> "GCFTable":
        """Build a table from {plane: taps[8, 8, S, S]}.

        Full tables need every plane in [-M, M]; half tables every plane in
        [0, M].  The support table is symmetric (supps[-p] = supps[p]) for a
        half table, which is what the reference's HalfGCF gridder expects
        (it indexes gcf_supps by the signed plane).
        """
        ps = sorted(planes)
        M = max(abs(p) for p in ps)
        want = list(range(0, M + 1)) if half else list(range(-M, M + 1))
        if ps != want:
            raise ValueError(f"planes {ps} do not cover {want[0]}..{want[-1]}")
        supps = np.zeros(2 * M + 1, np.int32)
        chunks, offs, lens = [], [], []
        off = 0
        for p in want:
            t = np.asarray(planes[p])
            if t.ndim != 4 or t.shape[:2] != (OVER, OVER) or t.shape[2] != t.shape[3]:
                raise ValueError(f"plane {p}: taps must be [8, 8, S, S], got {t.shape}")
            S = t.shape[2]
            supps[p + M] = S
            if half:
                supps[-p + M] = S
            for ou in range(OVER):
                for ov in range(OVER):
                    chunks.append(np.ascontiguousarray(t[ou, ov]).reshape(-1))
                    offs.append(off)
                    lens.append(S * S)
                    off += S * S
        data = np.concatenate(chunks).astype(dtype)
        return GCFTable(data=data,
                        layer_off=np.asarray(offs, np.int64),
                        layer_len=np.asarray(lens, np.int32),
                        center=0 if half else M * OVER * OVER,
                        supps=supps, max_w_plane=M, half=half)

    def astype(self, dtype) -> "GCFTable":
        return GCFTable(self.data.astype(dtype), self.layer_off.copy(),
                        self.layer_len.copy(), self.center, self.supps.copy(),
                        self.max_w_plane, self.half)

    def layer(self, index: int) -> np.ndarray:
        """Layer of signed index L as an S x S array."""
        slot = abs(index) if self.half else index + self.center
        o, n = int(self.layer_off[slot]), int(self.layer_len[slot])
        S = int(round(n ** 0.5))
        return self.data[o:o + n].reshape(S, S)

    # -- reference-style pointer tables (host memory) ----------------------
    def pointer_tables(self):
        """(centred layer-pointer table, centred support pointer) as c_void_p.

        The arrays behind them stay alive as long as this object does.
        """
        item = self.data.dtype.itemsize
        base = self.data.ctypes.data
        ptrs = (C.c_void_p * self.n_slots)(
            *[base + int(o) * item for o in self.layer_off])
        supps = np.ascontiguousarray(self.supps, np.int32)
        self._keep[:] = [ptrs, supps]
        gcf_c = C.c_void_p(C.addressof(ptrs) + self.center * C.sizeof(C.c_void_p))
        sup_c = C.c_void_p(supps.ctypes.data + self.max_w_plane * 4)
        return gcf_c, sup_c


@dataclass
class VisChunks:
    """Visibilities in the reference's chunked form: n_chunks chunks of ts_ch
    points, chunk b with support supps[b] (one chunk = one baseline's
    time/channel samples in MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:76-82).
    """
    uvw: np.ndarray      # (n_chunks, ts_ch, 3) float64
    vis: np.ndarray      # (n_chunks, ts_ch) complex
    supps: np.ndarray    # (n_chunks,) int32
    scale: float
    wstep: float

    @property
    def n_chunks(self) -> int:
        return int(self.uvw.shape[0])

    @property
    def ts_ch(self) -> int:
        return int(self.uvw.shape[1])

    @property
    def n(self) -> int:
        return self.n_chunks * self.ts_ch

    def chunk_pointers(self, arr: np.ndarray) -> C.Array:
        """Per-chunk pointer table into a (n_chunks, ts_ch, ...) C-contiguous
        array."""
        if not arr.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        stride = arr.strides[0]
        base = arr.ctypes.data
        # built in numpy; the ctypes array shares (and keeps alive) its memory
        ptrs = np.uint64(base) + np.arange(self.n_chunks, dtype=np.uint64) * np.uint64(stride)
        return (C.c_void_p * self.n_chunks).from_buffer(ptrs)
