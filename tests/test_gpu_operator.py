"""GPU tests of the differentiable measurement operator A = degrid o FFT
(rc_amd.operator): dot-product adjoint test, finite differences of chi^2,
and the chi^2 gradient against an FP64 reference built from the oracle
(FP64 degridder / gridder with conjugated taps) and numpy's FFT.

Bars: adjoint identity and the FP64 gradient comparison within 1e-5
relative (FP32 kernels on complex64-rounded inputs); finite differences
within 1e-3 (chi^2 is quadratic, so central differences have no truncation
error; the bar covers FP32 rounding of chi^2 itself)."""
import dataclasses

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import device as D
from rc_amd import operator as M
from rc_amd import synth
from rc_amd.tables import VisChunks

pytestmark = pytest.mark.gpu


def make_op(N=128, n_ant=24, n_times=24, n_planes=4, s_max=16, seed=3):
    vc, gcf, _ = synth.wproj(seed=seed, n_ant=n_ant, n_times=n_times, grid_size=N,
                             n_planes=n_planes, s_min=8, s_max=s_max, vis_dtype=np.complex128)
    gcf = gcf.astype(np.complex64).astype(np.complex128)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    return M.MeasurementOp(geom, dg, tile=32, chunk=128), vc, gcf


def fshift_fft(x, sign):
    """The reference's transform: fftshift, unnormalised DFT, fftshift."""
    s = np.fft.fftshift(x)
    s = np.fft.fft2(s) if sign < 0 else np.fft.ifft2(s) * s.size
    return np.fft.fftshift(s)


def a64(vc, gcf, x):
    g = fshift_fft(x.astype(np.complex128), -1)
    return O.degrid(vc, gcf, g, nthreads=8)[0].reshape(-1)


def ah64(vc, gcf, y, N):
    """A^H y in FP64: grid with conjugated taps, then the sign +1 transform."""
    conj = dataclasses.replace(gcf, data=np.conj(gcf.data), _keep=[])
    vcy = VisChunks(uvw=vc.uvw, vis=y.reshape(vc.vis.shape), supps=vc.supps, scale=vc.scale,
                    wstep=vc.wstep)
    g, _ = O.grid(vcy, conj, N, nthreads=8)
    return fshift_fft(g, 1)


def test_adjoint_dot_product(hip_device):
    op, vc, gcf = make_op()
    rng = np.random.default_rng(0)
    N = op.size
    for complex_image in (False, True):
        x = rng.standard_normal((N, N))
        if complex_image:
            x = x + 1j * rng.standard_normal((N, N))
        y = rng.standard_normal(op.n) + 1j * rng.standard_normal(op.n)
        xt = torch.from_numpy(x.astype(np.complex64 if complex_image else np.float32)).cuda()
        yt = torch.from_numpy(y.astype(np.complex64)).cuda()
        ax = op.forward(xt).cpu().numpy().astype(np.complex128)
        ahy = op.adjoint(yt).cpu().numpy().astype(np.complex128)
        x32 = xt.cpu().numpy().astype(np.complex128)
        y32 = yt.cpu().numpy().astype(np.complex128)
        lhs = np.vdot(y32, ax)          # y^H (A x)
        rhs = np.vdot(ahy, x32)         # (A^H y)^H x
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (complex_image, lhs, rhs)
    op.close()


def test_forward_and_adjoint_match_fp64(hip_device):
    op, vc, gcf = make_op(seed=4)
    rng = np.random.default_rng(1)
    N = op.size
    x = rng.standard_normal((N, N)).astype(np.float32)
    y = (rng.standard_normal(op.n) + 1j * rng.standard_normal(op.n)).astype(np.complex64)
    ax = op.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ahy = op.adjoint(torch.from_numpy(y).cuda()).cpu().numpy()
    assert rel_l2(ax, a64(vc, gcf, x)) <= 1e-5
    assert rel_l2(ahy, ah64(vc, gcf, y.astype(np.complex128), N)) <= 1e-5
    op.close()


def test_chi2_gradient_matches_fp64_reference(hip_device):
    """autograd gradient of chi^2 = sum |A x - d|^2 through the hand-written
    backward, against 2 Re(A^H (A x - d)) evaluated in FP64."""
    op, vc, gcf = make_op(seed=5)
    rng = np.random.default_rng(2)
    N = op.size
    x = rng.standard_normal((N, N)).astype(np.float32)
    d = (rng.standard_normal(op.n) + 1j * rng.standard_normal(op.n)).astype(np.complex64)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    loss = M.chi2(op, xt, torch.from_numpy(d).cuda())
    loss.backward()
    g32 = xt.grad.cpu().numpy().astype(np.float64)
    r = a64(vc, gcf, x) - d.astype(np.complex128)
    g64 = 2.0 * np.real(ah64(vc, gcf, r, N))
    assert rel_l2(g32, g64) <= 1e-5
    c64 = float(np.sum(np.abs(r) ** 2))
    assert abs(float(loss.detach()) - c64) <= 1e-5 * c64
    op.close()


def test_chi2_finite_differences(hip_device):
    op, vc, gcf = make_op(seed=6)
    rng = np.random.default_rng(3)
    N = op.size
    x = torch.from_numpy(rng.standard_normal((N, N)).astype(np.float32)).cuda()
    d = op.forward(x + 0.3 * torch.from_numpy(
        rng.standard_normal((N, N)).astype(np.float32)).cuda())
    xg = x.clone().requires_grad_(True)
    M.chi2(op, xg, d).backward()
    for k in range(3):
        v = torch.from_numpy(rng.standard_normal((N, N)).astype(np.float32)).cuda()
        h = 0.05

        def L(z):
            r = (op.forward(z) - d).to(torch.complex128)
            return float((r.real.square() + r.imag.square()).sum())

        fd = (L(x + h * v) - L(x - h * v)) / (2 * h)
        an = float((xg.grad.double() * v.double()).sum())
        assert abs(fd - an) <= 1e-3 * abs(an), (k, fd, an)
    op.close()


def test_gradient_steps_reduce_chi2(hip_device):
    """20 steps of gradient descent on chi^2 (step 1/L from a power
    iteration on A^H A) decrease chi^2 monotonically."""
    op, vc, gcf = make_op(seed=7)
    rng = np.random.default_rng(4)
    N = op.size
    truth = torch.zeros((N, N), device="cuda")
    idx = rng.integers(N // 4, 3 * N // 4, (20, 2))
    truth[idx[:, 0], idx[:, 1]] = torch.from_numpy(
        rng.lognormal(0.0, 1.0, 20).astype(np.float32)).cuda()
    d = op.forward(truth)
    v = torch.from_numpy(rng.standard_normal((N, N)).astype(np.float32)).cuda()
    for _ in range(15):
        v = op.adjoint(op.forward(v)).real
        lip = float(v.norm())
        v = v / lip
    step = 1.0 / (2.0 * lip * 1.05)
    x = torch.zeros((N, N), device="cuda", requires_grad=True)
    prev = None
    for _ in range(20):
        loss = M.chi2(op, x, d)
        x.grad = None
        loss.backward()
        with torch.no_grad():
            x -= step * x.grad
        cur = float(loss)
        if prev is not None:
            assert cur <= prev * (1 + 1e-6)
        prev = cur
    assert prev < 0.9 * float(M.chi2(op, torch.zeros_like(truth), d))
    op.close()
