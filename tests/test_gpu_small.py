"""GPU parity of the kernels next to the gridder: the reference's host entry
points (normalizeCPU, rotateCPU, grid0, reweight, herm_padded_inplace,
fft_inplace_even, deconvolve) against the golden vectors and the oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import GOLDEN, rel_l2
from oracle import fft_oracle as F
from oracle import oracle as O
from rc_amd import _lib, ms4

pytestmark = pytest.mark.gpu


def test_small_kernels_golden(hip_device):
    z = np.load(GOLDEN / "small_kernels.npz")
    N, P = int(z["N"]), int(z["P"])
    a = z["src"].copy()
    ms4.normalize(a)
    assert np.array_equal(a, z["norm"])
    uvw, vis, scale = z["uvw"], z["vis"], float(z["scale"])
    r = vis.copy()
    ms4.rotate(uvw, r, scale)
    assert rel_l2(r, z["rot"]) <= 1e-15
    g0 = np.zeros((N, P), np.complex128)
    ms4.grid0(uvw, vis, g0, scale)
    assert rel_l2(g0, z["grid0"]) <= 1e-15
    assert np.array_equal(g0 != 0, z["grid0"] != 0)
    rw = vis.copy()
    ms4.reweight(uvw, rw, scale, N)
    assert np.array_equal(rw, z["reweight"])


def test_herm_padded_vs_oracle(hip_device):
    rng = np.random.default_rng(1)
    N, P = 64, 70
    a = rng.standard_normal((N, P)) + 1j * rng.standard_normal((N, P))
    b, c = a.copy(), a.copy()
    ms4.herm_padded_inplace(b)
    O.oracle_lib().orc_herm_padded(c.ctypes.data, N, P)
    assert np.array_equal(b, c)


@pytest.mark.parametrize("sign", [-1, 1, 0, 2])
@pytest.mark.parametrize("n,p", [(64, 64), (48, 53)])
def test_fft_inplace_even_vs_oracle(hip_device, sign, n, p):
    rng = np.random.default_rng(10 + sign)
    buf = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    if sign == 0:
        # a real image in the real view, the rest arbitrary bytes
        F.real_view(buf, n, p)[:, :n] = rng.standard_normal((n, n))
    exp = F.fft_inplace_even(buf, sign, n)
    got = buf.copy()
    ms4.fft_inplace_even(sign, got, n, p)
    if sign == 2:
        a, b = F.real_view(got, n, p)[:, :n], F.real_view(exp, n, p)[:, :n]
    else:
        a, b = got[:, :n], exp[:, :n]
    assert rel_l2(a, b) <= 1e-12


def test_fft_device_fp32_c2c(hip_device):
    """rcg_fft (single precision) forward then backward = N^2 * identity."""
    n = 256
    rng = np.random.default_rng(2)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    d = torch.from_numpy(x).cuda()
    h = _lib.lib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    plans = []
    for sign in (-1, 1):
        pl = C.c_void_p()
        _lib.check(h.rcg_fft_plan_create(sign, 0, n, n, C.byref(pl)), "plan")
        plans.append(pl)
    _lib.check(h.rcg_fft_execute(plans[0], d.data_ptr(), s), "fwd")
    fwd = d.cpu().numpy()
    exp = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(x.astype(np.complex128))))
    assert rel_l2(fwd, exp) <= 1e-5
    _lib.check(h.rcg_fft_execute(plans[1], d.data_ptr(), s), "bwd")
    back = d.cpu().numpy() / (n * n)
    assert rel_l2(back, x) <= 1e-5
    for pl in plans:
        h.rcg_fft_plan_destroy(pl)


def _point_sources(n, p, srcs, sigma=1.5, noise=0.0, seed=0):
    x = np.arange(n)
    c = n // 2
    psf = np.zeros((n, p))
    psf[:, :n] = np.exp(-((x[:, None] - c) ** 2 + (x[None, :] - c) ** 2) / (2 * sigma ** 2))
    dirty = np.zeros((n, p))
    for (r, cc, f) in srcs:
        for i in range(n):
            si = i - r + c
            if not 0 <= si < n:
                continue
            j = np.arange(n)
            sj = j - cc + c
            ok = (sj >= 0) & (sj < n)
            dirty[i, j[ok]] += f * psf[si, sj[ok]]
    if noise:
        dirty[:, :n] += noise * np.random.default_rng(seed).standard_normal((n, n))
    return dirty, psf


def test_deconvolve_vs_oracle(hip_device):
    n, p = 96, 100
    dirty, psf = _point_sources(n, p, [(20, 30, 2.0), (50, 70, -1.0), (70, 20, 0.5)],
                                noise=1e-3)
    res_gpu = dirty.copy()
    mod_gpu = ms4.deconvolve(res_gpu, psf, 300, 0.1, 1e-3)
    res_ref = dirty.copy()
    mod_ref = np.zeros_like(dirty)
    O.oracle_lib().orc_hogbom(mod_ref.ctypes.data, res_ref.ctypes.data, psf.ctypes.data, n,
                              p, 300, 0.1, 1e-3)
    assert np.array_equal(mod_gpu != 0, mod_ref != 0)
    assert rel_l2(mod_gpu, mod_ref) <= 1e-12
    assert rel_l2(res_gpu[:, :n], res_ref[:, :n]) <= 1e-12


def test_hogbom_device_fp32_iterations_and_threshold(hip_device):
    n = 128
    dirty, psf = _point_sources(n, n, [(40, 40, 1.0), (90, 60, 0.6)])
    h = _lib.lib()
    res = torch.from_numpy(dirty.astype(np.float32)).cuda()
    ps = torch.from_numpy(psf.astype(np.float32)).cuda()
    mod = torch.empty_like(res)
    scratch = torch.empty(int(h.rcg_hogbom_scratch_bytes()), dtype=torch.uint8, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(h.rcg_hogbom(mod.data_ptr(), res.data_ptr(), ps.data_ptr(), 0, n, n, 1000, 0.5,
                            1e-2, scratch.data_ptr(), s), "rcg_hogbom")
    it = C.c_uint()
    _lib.check(h.rcg_hogbom_iterations(scratch.data_ptr(), C.byref(it), s), "iters")
    res_ref = dirty.copy()
    mod_ref = np.zeros_like(dirty)
    it_ref = O.oracle_lib().orc_hogbom(mod_ref.ctypes.data, res_ref.ctypes.data,
                                       psf.ctypes.data, n, n, 1000, 0.5, 1e-2)
    assert it.value == it_ref < 1000
    assert rel_l2(mod.cpu().numpy(), mod_ref) <= 1e-5
    assert np.abs(res.cpu().numpy()).max() < 1e-2


@pytest.mark.parametrize("f64", [True, False])
@pytest.mark.parametrize("n,p", [(2000, 2008), (2600, 2608)])
def test_hogbom_ragged_large_vs_oracle(hip_device, f64, n, p):
    """Ragged sizes in a padded pitch, with a PSF whose sidelobes cover the
    whole image, against the oracle's Hogbom (FP64 to 1e-12, FP32 to 1e-5):
    several rows per block and columns per thread, and at 2600^2 more than
    one batch of loads per thread."""
    from rc_amd import device as D
    yy, xx = np.mgrid[0:n, 0:n]
    c = n // 2
    psf = np.zeros((n, p))
    psf[:, :n] = (np.exp(-((yy - c) ** 2 + (xx - c) ** 2) / 4.5)
                  + 0.02 * np.cos(0.05 * (yy - c)) * np.cos(0.07 * (xx - c)))
    dirty = np.zeros((n, p))
    srcs = [(700, n - 500, 2.0), (n - 700, 300, -1.3)] if f64 else [(700, n - 500, 2.0)]
    for (r, cc, f) in srcs:
        i0, i1 = max(0, r - c), min(n, n + r - c)
        j0, j1 = max(0, cc - c), min(n, n + cc - c)
        dirty[i0:i1, j0:j1] += f * psf[i0 - r + c:i1 - r + c, j0 - cc + c:j1 - cc + c]
    dirty[:, :n] += 1e-3 * np.random.default_rng(5).standard_normal((n, n))
    niters = 25
    res_ref = dirty.copy()
    mod_ref = np.zeros_like(dirty)
    it_ref = O.oracle_lib().orc_hogbom(mod_ref.ctypes.data, res_ref.ctypes.data,
                                       psf.ctypes.data, n, p, niters, 0.1, 0.0)
    dt = torch.float64 if f64 else torch.float32
    res = torch.from_numpy(dirty).to(dt).cuda()
    ps = torch.from_numpy(psf).to(dt).cuda()
    mod, it = D.hogbom(res, ps, niters, 0.1, 0.0)
    assert it == it_ref == niters
    tol = 1e-12 if f64 else 1e-5
    m = mod.cpu().numpy().astype(np.float64)
    r = res.cpu().numpy().astype(np.float64)
    assert np.array_equal(m != 0, mod_ref != 0)
    assert rel_l2(m, mod_ref) <= tol
    assert rel_l2(r[:, :n], res_ref[:, :n]) <= tol


def test_vis5_unpack_rotate_and_pack(hip_device):
    """MS6 five-double records -> uvw + visibilities (rotated by
    exp(2 pi i w scale) in FP64 against the oracle's rotateCPU) and back."""
    from rc_amd import device as D
    rng = np.random.default_rng(12)
    n = 5000
    rec = rng.standard_normal((n, 5)) * np.array([100, 100, 30, 1, 1])
    t = torch.from_numpy(rec).cuda()
    uvw, vis = D.unpack_vis5(t, f64=True)
    assert np.array_equal(uvw.cpu().numpy(), rec[:, :3])
    assert np.array_equal(vis.cpu().numpy(), rec[:, 3] + 1j * rec[:, 4])
    uvw, vrot = D.unpack_vis5(t, f64=True, rotate_scale=0.013)
    exp = (rec[:, 3] + 1j * rec[:, 4]).copy()
    uvw_np = np.ascontiguousarray(rec[:, :3])  # kept alive across the call
    O.oracle_lib().orc_rotate(uvw_np.ctypes.data, exp.ctypes.data, n, 0.013)
    assert rel_l2(vrot.cpu().numpy(), exp) <= 1e-15
    back = D.pack_vis5(uvw, vis).cpu().numpy()
    assert np.array_equal(back, rec)
    _, v32 = D.unpack_vis5(t, f64=False)
    assert np.array_equal(v32.cpu().numpy(), (rec[:, 3] + 1j * rec[:, 4]).astype(np.complex64))
