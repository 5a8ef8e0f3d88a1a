"""GPU tests of the imaging sequence (rc_amd.imaging): the FP64 dirty image
and PSF against the oracle chain of the reference's cppcycle.cpp (oracle
grid, normalize, herm_padded, numpy C2R with the reference's shifts), the
major-cycle bookkeeping, and the subtract/const kernels.  Bars: FP64 images
within 1e-12 of the oracle chain (relative to the image maximum)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import fft_oracle as F
from oracle import oracle as O
from rc_amd import _lib
from rc_amd import device as D
from rc_amd import imaging as IM
from rc_amd import synth
from rc_amd.tables import VisChunks

pytestmark = pytest.mark.gpu


def small(f64, N=128, P=132, seed=3):
    vc, gcf, _ = synth.wproj(seed=seed, n_ant=24, n_times=24, grid_size=N, n_planes=4, s_min=8,
                             s_max=16, vis_dtype=np.complex128)
    geom = D.DeviceGeom.from_chunks(vc, N, P)
    dt = np.complex128 if f64 else np.complex64
    dg = D.DeviceGCF.from_table(gcf, dtype=dt)
    return vc, gcf, geom, dg


def cpu_image(vc, gcf, N, P):
    g, _ = O.grid(vc, gcf, N, pitch=P, nthreads=8)
    h = O.oracle_lib()
    h.orc_normalize(g.ctypes.data, P, N)
    h.orc_herm_padded(g.ctypes.data, N, P)
    out = F.fft_inplace_even(g, 2, N)
    return F.real_view(out, N, P)[:, :N].copy()


def test_fp64_dirty_image_and_psf_match_oracle_chain(hip_device):
    vc, gcf, geom, dg = small(True)
    N, P = geom.grid_size, geom.grid_pitch
    im = IM.Imager(geom, dg, f64=True)
    vis = torch.from_numpy(vc.vis.reshape(-1)).cuda()
    dirty = im.image(vis).cpu().numpy()[:, :N]
    exp = cpu_image(vc, gcf, N, P)
    assert np.abs(dirty - exp).max() <= 1e-12 * np.abs(exp).max()
    psf = im.psf().cpu().numpy()[:, :N]
    ones = VisChunks(uvw=vc.uvw, vis=np.ones_like(vc.vis), supps=vc.supps, scale=vc.scale,
                     wstep=vc.wstep)
    exp_psf = cpu_image(ones, gcf, N, P)
    assert np.abs(psf - exp_psf).max() <= 1e-12 * np.abs(exp_psf).max()
    im.close()


def test_major_cycle_subtracts_the_degridded_model(hip_device):
    vc, gcf, geom, dg = small(False, seed=4)
    N = geom.grid_size
    im = IM.Imager(geom, dg, f64=False, tile=32, chunk=128)
    sky = torch.zeros((N, geom.grid_pitch), device="cuda")
    rng = np.random.default_rng(0)
    for (r, c) in rng.integers(N // 4, 3 * N // 4, (5, 2)):
        sky[r, c] = float(rng.lognormal())
    vis = im.predict(sky)
    one = im.major_cycles(vis, 1, 150, 0.1, 0.0)
    two = im.major_cycles(vis, 2, 150, 0.1, 0.0)
    # the second round started from vis - predict(model of the first round)
    exp = vis.clone()
    im.subtract(exp, im.predict(one["model"]))
    assert rel_l2(two["vis"].cpu().numpy(), exp.cpu().numpy()) <= 1e-5
    assert len(two["iters"]) == 2 and two["iters"][0] == 150
    d0 = float(one["residual"][:, :N].abs().max())
    d1 = float(two["residual"][:, :N].abs().max())
    dirty = float(im.image(vis)[:, :N].abs().max())
    assert d0 < dirty and d1 < dirty
    im.close()


def test_subtract_and_const_vis(hip_device):
    h = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    for f64, dt in ((1, torch.complex128), (0, torch.complex64)):
        a = torch.randn(1001, dtype=dt, device="cuda")
        b = torch.randn(1001, dtype=dt, device="cuda")
        exp = (a - b).cpu()
        _lib.check(h.rcg_subtract_vis(a.data_ptr(), b.data_ptr(), f64, 1001, s), "sub")
        assert torch.equal(a.cpu(), exp)
        _lib.check(h.rcg_const_vis(a.data_ptr(), f64, 1001, 1.0, -2.0, s), "const")
        assert torch.equal(a.cpu(), torch.full((1001,), 1 - 2j, dtype=dt))
