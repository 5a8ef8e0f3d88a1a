"""The differentiable measurement operator: sky image -> visibilities.

    A x = D(F(x))

F is the reference's image -> grid transform (fft_inplace_even with sign -1:
fftshift, unnormalised forward DFT, fftshift; MS4/kernel/cpu/fft/
fft_dyn_padded.cpp:77-115) on the image taken as complex, and D is the
w-projection degridder (the transpose of the gridder, as the reference's
deGridKernelCPU*GCF, MS4/kernel/cpu/scatter_grid/
scatter_gridder_w_dependent_dyn_1p.cpp:368-431).  Both run in librcgrid.so
(rocFFT plans and the tiled FP32 degridder).

The adjoint is A^H y = F^H(D^H(y)): D^H grids with conjugated taps
(rcg_grid_tiled_adjoint_f32) and F^H is the sign +1 transform with the same
shifts (fftshift is a symmetric permutation for even N, and the unnormalised
inverse DFT is the adjoint of the forward one).  For a real image, the
gradient of a real loss L with respect to x is Re(A^H g), where g is the
gradient PyTorch passes for the complex output (dL/d Re y + i dL/d Im y).

MeasurementOp.apply(x) is the torch.autograd.Function; its backward is the
hand-written adjoint above (the gridder), not an autograd trace.
"""
from __future__ import annotations

import torch

from . import device as D


class MeasurementOp:
    """A fixed geometry (pregrid + tile bins done once), GCF table and grid
    size; forward() and adjoint() on device tensors."""

    def __init__(self, geom: D.DeviceGeom, gcf: D.DeviceGCF, tile: int = 32, chunk: int = 512,
                 max_support: int | None = None):
        if gcf.data.dtype != torch.complex64:
            raise TypeError("the measurement operator runs the FP32 path: complex64 taps")
        if geom.grid_size != geom.grid_pitch or geom.grid_size % 2:
            raise ValueError("the image transform needs an even square grid (pitch = size)")
        self.geom, self.gcf, self.n = geom, gcf, geom.n
        self.size = geom.grid_size
        self.plan = D.TiledPlan(geom, tile=tile, chunk=chunk, max_support=max_support)
        self.plan.bin(gcf)
        st = self.plan.stats()
        if st["err"]:
            raise RuntimeError(f"binning error word {st['err']}")
        self.taps = st["taps"]
        self._fwd = D.FFTPlan(-1, self.size)
        self._bwd = D.FFTPlan(1, self.size)

    def forward(self, image: torch.Tensor) -> torch.Tensor:
        """image [N, N] (real or complex) -> visibilities [n] complex64."""
        if tuple(image.shape) != (self.size, self.size):
            raise ValueError(f"image must be {self.size} x {self.size}")
        g = image.to(torch.complex64).contiguous()
        if g.data_ptr() == image.data_ptr():
            g = g.clone()
        self._fwd(g)
        return self.plan.degrid(self.gcf, g)

    def adjoint(self, vis: torch.Tensor) -> torch.Tensor:
        """visibilities [n] -> complex image [N, N] (A^H vis)."""
        g = self.plan.grid(self.gcf, vis.to(torch.complex64).contiguous(), adjoint=True)
        return self._bwd(g)

    def grid_image(self, vis: torch.Tensor) -> torch.Tensor:
        """The imaging direction of the reference pipeline: grid with the
        gridder's own taps, then the sign +1 transform (a dirty image when
        vis are data, the PSF when vis are ones)."""
        g = self.plan.grid(self.gcf, vis.to(torch.complex64).contiguous())
        return self._bwd(g)

    def apply(self, image: torch.Tensor) -> torch.Tensor:
        return _Measure.apply(image, self)

    def close(self) -> None:
        self.plan.close()
        self._fwd.close()
        self._bwd.close()


class _Measure(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, op: MeasurementOp):
        ctx.op = op
        ctx.real = not image.is_complex()
        return op.forward(image.detach())

    @staticmethod
    def backward(ctx, grad_vis):
        g = ctx.op.adjoint(grad_vis.contiguous())
        return (g.real.contiguous() if ctx.real else g), None


def chi2(op: MeasurementOp, image: torch.Tensor, data: torch.Tensor) -> torch.Tensor:
    """sum |A x - d|^2, differentiable in x through the hand-written adjoint."""
    r = op.apply(image) - data
    return (r.real.square() + r.imag.square()).sum()
