"""GCF tables generated on the device (librcgrid.so rcg_gcf_* / rcg_calc_accums).

make_gcf() follows the reference's GCF kernel (MS4/dna-programs/Kernel/CPU/
GCF.hs:36-89): per-plane supports min(max_size, min_size + growth*|p|) for
planes -M..M, the mean w of every plane's points (or the plane's nominal w
when it has none), and one mkGCFLayer per plane with w = mean * theta and
t2 = theta / 2.  The result is a full (signed-plane) table ready for the
gridders.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .device import DeviceGCF, _stream
from .tables import OVER


class GCFGenerator(_lib.Handle):
    """rcg_gcf_gen: FP64 arena and rocFFT plan for one max_support/src_pad."""

    def __init__(self, max_support: int, src_pad: int = 4):
        h = C.c_void_p()
        _lib.check(_lib.lib().rcg_gcf_gen_create(max_support, src_pad, C.byref(h)),
                   "rcg_gcf_gen_create")
        self.max_support, self.src_pad = max_support, src_pad
        self._own(h)

    def layer(self, dst: torch.Tensor, support: int, t2: float, w: float,
              pad_quirk: bool = False, stream=None) -> torch.Tensor:
        """One plane's 64 layers [8, 8, support, support] into dst
        (complex128 or complex64, contiguous, 64*support^2 elements)."""
        if dst.numel() != OVER * OVER * support * support or not dst.is_contiguous():
            raise ValueError("dst must hold 64*support^2 contiguous elements")
        f64 = dst.dtype == torch.complex128
        if not f64 and dst.dtype != torch.complex64:
            raise TypeError("dst must be complex128 or complex64")
        _lib.check(_lib.lib().rcg_gcf_gen_layer(self._h, dst.data_ptr(), int(f64), support,
                                                t2, w, int(pad_quirk), _stream(stream)),
                   "rcg_gcf_gen_layer")
        return dst

    DESTROY = "rcg_gcf_gen_destroy"


def calc_accums(uvw: torch.Tensor, wstep: float, n_planes: int,
                stream=None) -> tuple[torch.Tensor, torch.Tensor, int]:
    """Per-plane sum of w and point count (device tensors), and the error
    word (bit 0: points outside the planes, skipped)."""
    uvw = uvw.reshape(-1, 3)
    if uvw.dtype != torch.float64 or not uvw.is_contiguous():
        raise TypeError("uvw must be contiguous float64 [n, 3]")
    dev = uvw.device
    sums = torch.empty(n_planes, dtype=torch.float64, device=dev)
    npts = torch.empty(n_planes, dtype=torch.int32, device=dev)
    err = torch.empty(1, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().rcg_calc_accums(uvw.data_ptr(), uvw.shape[0], wstep, n_planes,
                                          sums.data_ptr(), npts.data_ptr(), err.data_ptr(),
                                          _stream(stream)), "rcg_calc_accums")
    return sums, npts, int(err.item())


def make_gcf(uvw: torch.Tensor, wstep: float, theta: float, min_size: int, growth: int,
             max_size: int, max_w_plane: int | None = None, src_pad: int = 4,
             pad_quirk: bool = False, dtype=torch.complex64, stream=None
             ) -> tuple[DeviceGCF, dict]:
    """Full GCF table for planes -M..M on the device; returns the table and
    {"supps", "means", "npts"} (host lists)."""
    uvw = uvw.reshape(-1, 3)
    if max_w_plane is None:
        w = uvw[:, 2]
        max_w_plane = max(int(torch.round(w.max() / wstep).item()),
                          int(torch.round(-w.min() / wstep).item()), 0)
    M = int(max_w_plane)
    n_planes = 2 * M + 1
    sums, npts, err = calc_accums(uvw, wstep, n_planes, stream)
    if err:
        raise _lib.RcgError("make_gcf: points outside the w-planes")
    sums, npts = sums.cpu().numpy(), npts.cpu().numpy()
    supps = [min(max_size, min_size + growth * abs(p)) for p in range(-M, M + 1)]
    means = []
    wmid = -wstep * M
    for k in range(n_planes):
        means.append(sums[k] / npts[k] if npts[k] > 0 else wmid)
        wmid = wmid + wstep
    offs, lens, off = [], [], 0
    plane_off = []
    for S in supps:
        plane_off.append(off)
        for _ in range(OVER * OVER):
            offs.append(off)
            lens.append(S * S)
            off += S * S
    data = torch.empty(off, dtype=dtype, device=uvw.device)
    gen = GCFGenerator(max_size, src_pad)
    try:
        for k, S in enumerate(supps):
            seg = data[plane_off[k]:plane_off[k] + OVER * OVER * S * S]
            gen.layer(seg, S, theta / 2.0, means[k] * theta, pad_quirk, stream)
        torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
    finally:
        gen.close()
    dev = uvw.device
    table = DeviceGCF(
        data=data,
        layer_off=torch.tensor(offs, dtype=torch.int64, device=dev),
        layer_len=torch.tensor(lens, dtype=torch.int32, device=dev),
        supps=torch.tensor(supps, dtype=torch.int32, device=dev),
        center=M * OVER * OVER, max_w_plane=M, half=False)
    return table, {"supps": supps, "means": means, "npts": npts.tolist()}
