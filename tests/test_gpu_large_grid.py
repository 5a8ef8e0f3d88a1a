"""Grids past the LDS histogram of the binning pass: a 16384^2 grid (262,144
tiles of 32 per instance, counted and ranked with global atomics) with a
sparse w-projection sample (64 antennas x 64 times = 129,024 points, 16
w-planes, supports 8..64, half table), against the oracle on the same
complex64-rounded inputs: the FP32 tiled path in both record layouts (grid
and degrid, relative L2 <= 1e-5), the FP64 tiled gridder (<= 1e-12), error
word 0 and tap counts equal, the FP64 tiled degridder (<= 1e-12); and the
MS4 host calls at that size within 1e-12: gridKernelCPUHalfGCF on the tiled
FP64 gridder (not the plain one), deGridKernelCPUHalfGCF on the plain FP64
degridder (the faster of the two).
The reference takes any grid its int16 coordinates allow
(MS4/kernel/common/common.h:128-134)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import _lib, ms4
from rc_amd import device as D
from rc_amd import synth
from rc_amd.tables import VisChunks

pytestmark = pytest.mark.gpu

N = 16384
RCG_PATH_TILED = 1
RCG_PATH_PLAIN = 2


@pytest.fixture(scope="module")
def case():
    vc, gcf, _ = synth.wproj(seed=9, n_ant=64, n_times=64, grid_size=N, n_planes=16, s_min=8,
                             s_max=64)
    vc = VisChunks(uvw=vc.uvw, vis=vc.vis.astype(np.complex64).astype(np.complex128),
                   supps=vc.supps, scale=vc.scale, wstep=vc.wstep)
    gcf = gcf.astype(np.complex64).astype(np.complex128)
    ref, taps = O.grid(vc, gcf, N, nthreads=16)
    return vc, gcf, ref, taps


@pytest.mark.parametrize("layout,tile", [("pairs", 32), ("halo", 32)])
def test_fp32_tiled_16384(hip_device, case, layout, tile):
    vc, gcf, ref, taps = case
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    plan = D.TiledPlan(geom, tile=tile, chunk=512, max_support=64, layout=layout)
    plan.bin(dg)
    st = plan.stats()
    assert st["err"] == 0 and st["taps"] == taps
    vis = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).cuda()
    g = plan.grid(dg, vis)
    torch.cuda.synchronize()
    assert rel_l2(g.cpu().numpy(), ref) <= 1e-5
    del g
    ref32 = ref.astype(np.complex64)
    v = plan.degrid(dg, torch.from_numpy(ref32).cuda()).cpu().numpy()
    plan.close()
    vref, _ = O.degrid(vc, gcf, ref32.astype(np.complex128), nthreads=16)
    assert rel_l2(v, vref.reshape(-1)) <= 1e-5


def test_fp64_tiled_16384(hip_device, case):
    vc, gcf, ref, taps = case
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex128)
    plan = D.TiledPlan(geom, tile=32, chunk=1024, max_support=64)
    plan.bin(dg)
    st = plan.stats()
    assert st["err"] == 0 and st["taps"] == taps
    vis = torch.from_numpy(vc.vis.reshape(-1)).cuda()
    g = plan.grid_f64(dg, vis)
    torch.cuda.synchronize()
    plan.close()
    assert rel_l2(g.cpu().numpy(), ref) <= 1e-12


def test_fp64_tiled_degrid_16384(hip_device, case):
    vc, gcf, ref, taps = case
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex128)
    plan = D.TiledPlan(geom, tile=32, chunk=1024, max_support=64)
    plan.bin(dg)
    v = plan.degrid_f64(dg, torch.from_numpy(ref).cuda()).cpu().numpy()
    plan.close()
    vref, _ = O.degrid(vc, gcf, ref, nthreads=16)
    assert rel_l2(v, vref.reshape(-1)) <= 1e-12


def test_ms4_host_calls_16384(hip_device, case):
    vc, gcf, ref, taps = case
    g, t = ms4.grid(vc, gcf, N)
    assert _lib.lib().rcg_legacy_path() == RCG_PATH_TILED
    assert t == taps
    assert rel_l2(g, ref) <= 1e-12
    v, dt = ms4.degrid(vc, gcf, ref)
    # the plain FP64 degridder, a gather without atomics, is the faster one
    assert _lib.lib().rcg_legacy_path() == RCG_PATH_PLAIN
    vref, _ = O.degrid(vc, gcf, ref, nthreads=16)
    assert dt == taps
    assert rel_l2(v, vref) <= 1e-12
