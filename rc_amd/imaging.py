"""The reference's MS4 imaging sequence on the GPU: dirty image, PSF, Hogbom
CLEAN and major cycles, on device data through librcgrid.so.

Follows MS4/misc/cppcycle.cpp:190-292 (grid, normalizeCPU,
herm_padded_inplace, fft_inplace_even sign 2 into the real view of the same
buffer, the PSF from visibilities that are all 1, deconvolve) and the major
loop of MS4/dna-programs/imaging.hs:92-166 (dirty image -> clean -> degrid
the model and subtract it from the visibilities -> repeat; the residual of
the last minor cycle is the result).

Two deliberate differences, both because the reference's choice gives
wrong numbers:
  * the model image goes to the uv grid through the complex forward
    transform of the real image (sign -1), not the reference's R2C, whose
    in-place half spectrum leaves the other half of the grid unspecified
    for the degridder;
  * the subtraction out = in - degrid(model) is explicit (the MS4 GPU
    degridder subtracts, the CPU one overwrites; imaging.hs:55-64 states
    the subtraction as the intent).

f64=True runs the FP64 kernels (exact pregrid, FP64 atomics, FP64 FFT and
CLEAN) for parity with the reference; f64=False the tiled FP32 path.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from . import device as D


class Imager:
    def __init__(self, geom: D.DeviceGeom, gcf: D.DeviceGCF, f64: bool = False, tile: int = 32,
                 chunk: int = 512, max_support: int | None = None):
        self.geom, self.gcf, self.f64 = geom, gcf, f64
        self.N, self.P = geom.grid_size, geom.grid_pitch
        if self.N % 2:
            raise ValueError("the image transforms need an even grid size")
        want = torch.complex128 if f64 else torch.complex64
        if gcf.data.dtype != want:
            raise TypeError(f"{'FP64' if f64 else 'FP32'} imaging takes {want} taps")
        self.cdtype = want
        self.rdtype = torch.float64 if f64 else torch.float32
        dev = geom.uvw.device
        if f64:
            self.pg, taps, err = D.pregrid(geom, gcf)
            if int(err.item()):
                raise _lib.RcgError("pregrid: a point references a layer outside the table")
            self.taps = int(taps.item())
            self.plan = None
        else:
            self.plan = D.TiledPlan(geom, tile=tile, chunk=chunk, max_support=max_support)
            self.plan.bin(gcf)
            st = self.plan.stats()
            if st["err"]:
                raise _lib.RcgError(f"binning error word {st['err']}")
            self.taps = st["taps"]
        self._c2r = D.FFTPlan(2, self.N, self.P, f64=f64)
        self._fwd = D.FFTPlan(-1, self.N, self.P, f64=f64)
        self.device = dev

    # -- building blocks -------------------------------------------------
    def grid(self, vis: torch.Tensor) -> torch.Tensor:
        if self.f64:
            return D.grid_plain_f64(self.geom, self.gcf, self.pg, vis.to(self.cdtype).contiguous())
        return self.plan.grid(self.gcf, vis.to(self.cdtype).contiguous())

    def degrid(self, grid: torch.Tensor) -> torch.Tensor:
        if self.f64:
            return D.degrid_plain_f64(self.geom, self.gcf, self.pg, grid)
        return self.plan.degrid(self.gcf, grid)

    def _s(self):
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def real_view(self, grid: torch.Tensor) -> torch.Tensor:
        """The first N*P reals of a (N, P) complex buffer as a (N, P) image
        (the in-place real view fft_inplace_even writes)."""
        return torch.view_as_real(grid).reshape(-1)[: self.N * self.P].view(self.N, self.P)

    def image(self, vis: torch.Tensor) -> torch.Tensor:
        """grid -> normalise by 1/(N*P) -> hermitian fold -> C2R: the real
        image (N rows of P reals, a view of a fresh buffer)."""
        g = self.grid(vis)
        h = _lib.lib()
        f64 = int(self.f64)
        _lib.check(h.rcg_normalize(g.data_ptr(), f64, self.P, self.N, self._s()), "rcg_normalize")
        _lib.check(h.rcg_herm(g.data_ptr(), f64, self.N, self.P, self._s()), "rcg_herm")
        self._c2r(g)
        return self.real_view(g)

    def psf(self) -> torch.Tensor:
        ones = torch.empty(self.geom.n, dtype=self.cdtype, device=self.device)
        _lib.check(_lib.lib().rcg_const_vis(ones.data_ptr(), int(self.f64), ones.numel(), 1.0,
                                            0.0, self._s()), "rcg_const_vis")
        return self.image(ones)

    def predict(self, model: torch.Tensor) -> torch.Tensor:
        """Visibilities of a real model image (N x P reals, the first N
        columns used): complex forward transform, then degrid."""
        g = torch.zeros((self.N, self.P), dtype=self.cdtype, device=self.device)
        g.real[:, : self.N] = model[:, : self.N]
        self._fwd(g)
        return self.degrid(g)

    def subtract(self, vis: torch.Tensor, pred: torch.Tensor) -> None:
        """vis -= pred, in place."""
        _lib.check(_lib.lib().rcg_subtract_vis(vis.data_ptr(), pred.data_ptr(), int(self.f64),
                                               vis.numel(), self._s()), "rcg_subtract_vis")

    # -- the loop ----------------------------------------------------------
    def major_cycles(self, vis: torch.Tensor, n_major: int, niters: int, gain: float,
                     threshold: float) -> dict:
        """imaging.hs:140-160: n_major rounds of dirty image -> Hogbom ->
        (except after the last) subtract the degridded model.  Returns the
        last residual image, the summed model, the final visibilities and
        the CLEAN iterations per round."""
        psf = self.psf().clone()
        v = vis.to(self.cdtype).clone()
        total = torch.zeros((self.N, self.P), dtype=self.rdtype, device=self.device)
        iters = []
        residual = None
        for i in range(1, n_major + 1):
            residual = self.image(v).clone()
            model, it = D.hogbom(residual, psf, niters, gain, threshold)
            iters.append(it)
            total += model
            if i == n_major:
                break
            self.subtract(v, self.predict(model))
        return {"residual": residual, "model": total, "vis": v, "psf": psf, "iters": iters}

    def close(self) -> None:
        if self.plan is not None:
            self.plan.close()
        self._c2r.close()
        self._fwd.close()
