"""GPU parity of GCF generation (rcg_gcf_*, mkGCFLayer, calcAccums) against
the numpy oracle (oracle/gcf_oracle.py).  Bars: FP64 taps and w sums within
1e-12 (relative to the largest tap / sum), point counts exact; complex64
taps within 1e-6."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import gcf_oracle as G
from oracle import oracle as O
from rc_amd import _lib
from rc_amd import device as D
from rc_amd import gcf as GG
from rc_amd import synth
from rc_amd.tables import GCFTable, VisChunks

pytestmark = pytest.mark.gpu


def max_rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("support,max_support,pad,w,quirk", [
    (8, 8, 0, 0.0, False), (6, 8, 4, 12.5, False), (8, 8, 4, -7.25, True),
    (16, 32, 4, 40.0, False), (31, 64, 4, 150.0, True), (5, 9, 3, 3.3, False)])
def test_layer_vs_oracle(hip_device, support, max_support, pad, w, quirk):
    t2 = 0.02
    exp = G.mk_gcf_layer(support, max_support, pad, t2, w, pad_quirk=quirk)
    gen = GG.GCFGenerator(max_support, pad)
    d64 = torch.empty(64 * support * support, dtype=torch.complex128, device="cuda")
    d32 = torch.empty(64 * support * support, dtype=torch.complex64, device="cuda")
    gen.layer(d64, support, t2, w, quirk)
    gen.layer(d32, support, t2, w, quirk)
    got = d64.cpu().numpy().reshape(8, 8, support, support)
    assert max_rel(got, exp) <= 1e-12
    assert max_rel(d32.cpu().numpy().reshape(8, 8, support, support), exp) <= 1e-6
    gen.close()


def test_calc_accums_vs_oracle(hip_device):
    rng = np.random.default_rng(5)
    n, wstep, n_planes = 200_000, 0.75, 21
    uvw = rng.standard_normal((n, 3)) * 3.0
    uvw[:, 2] = np.clip(uvw[:, 2], -7.4, 7.4)  # planes -10..10
    uvw[:1000, 2] = (rng.integers(-14, 15, 1000)) * (wstep / 2)  # exact halves
    sums, npts, err = GG.calc_accums(torch.from_numpy(uvw).cuda(), wstep, n_planes)
    es, en = G.calc_accums(uvw, wstep, n_planes)
    assert err == 0
    assert np.array_equal(npts.cpu().numpy(), en)
    assert np.abs(sums.cpu().numpy() - es).max() <= 1e-12 * np.abs(es).max()


def test_calc_accums_reports_points_outside(hip_device):
    uvw = np.zeros((10, 3))
    uvw[:, 2] = [0, 1, 2, 3, 4, 5, 6, 7, 8, 100]
    sums, npts, err = GG.calc_accums(torch.from_numpy(uvw).cuda(), 1.0, 17)
    assert err & 1
    assert int(npts.sum()) == 9


def wproj_small():
    vc, _, N = synth.wproj(n_ant=16, n_times=12, grid_size=256, n_planes=4, s_min=8, s_max=16,
                           vis_dtype=np.complex128)
    return vc, N


def test_make_gcf_vs_oracle_kernel(hip_device):
    vc, N = wproj_small()
    uvw = vc.uvw.reshape(-1, 3)
    theta = vc.scale  # the field of view: pixel = u * theta
    dg, info = GG.make_gcf(torch.from_numpy(uvw).cuda(), vc.wstep, theta, min_size=4,
                           growth=2, max_size=16, dtype=torch.complex128)
    M = dg.max_w_plane
    supps, layers, means = G.gcf_kernel(uvw, vc.wstep, theta, 4, 2, 16, M, src_pad=4)
    assert info["supps"] == supps
    assert np.allclose(info["means"], means, rtol=1e-12, atol=0)
    data = dg.data.cpu().numpy()
    off = 0
    for k, S in enumerate(supps):
        got = data[off:off + 64 * S * S].reshape(8, 8, S, S)
        assert max_rel(got, layers[k]) <= 1e-12, k
        off += 64 * S * S


def test_generated_table_drives_the_gridder(hip_device):
    """The device-generated full table, used by the FP64 GPU gridder, gives
    the oracle's grid for the oracle-generated table."""
    vc, N = wproj_small()
    uvw = vc.uvw.reshape(-1, 3)
    theta = 0.02
    dg, info = GG.make_gcf(torch.from_numpy(uvw).cuda(), vc.wstep, theta, min_size=4,
                           growth=2, max_size=16, dtype=torch.complex128)
    M = dg.max_w_plane
    supps, layers, _ = G.gcf_kernel(uvw, vc.wstep, theta, 4, 2, 16, M, src_pad=4)
    table = GCFTable.from_planes({p: layers[p + M] for p in range(-M, M + 1)}, half=False)
    # chunk supports: the point's own plane support (one point per chunk)
    planes = np.rint(uvw[:, 2] / vc.wstep).astype(int)
    supp = np.asarray(supps)[planes + M].astype(np.int32)
    vc2 = VisChunks(uvw=vc.uvw, vis=vc.vis, supps=supp, scale=vc.scale, wstep=vc.wstep)
    ref, taps = O.grid(vc2, table, N, nthreads=8)
    geom = D.DeviceGeom.from_chunks(vc2, N)
    pg, tp, err = D.pregrid(geom, dg)
    g = D.grid_plain_f64(geom, dg, pg, torch.from_numpy(vc2.vis.reshape(-1)).cuda())
    assert int(tp.item()) == taps and int(err.item()) == 0
    assert rel_l2(g.cpu().numpy(), ref) <= 1e-12


def test_legacy_entry_points(hip_device):
    h = _lib.lib()
    S, ms, pad = 6, 8, 4
    dst = np.zeros(64 * S * S, np.complex128)
    table = (C.c_void_p * 64)()
    arena = np.zeros((ms * 8) * (ms * 8 + pad), np.complex128)
    p = h.mkGCFLayer(None, dst.ctypes.data, table, arena.ctypes.data, S, ms, pad, 0.02, 9.5)
    assert p and _lib.last_error()[0] == 0
    assert max_rel(dst.reshape(8, 8, S, S), G.mk_gcf_layer(S, ms, pad, 0.02, 9.5)) <= 1e-12
    assert [table[i] for i in (0, 1, 63)] == [dst.ctypes.data + i * S * S * 16 for i in (0, 1, 63)]
    # the handle is reused; the quirk switch selects the reference's sums
    h.rcg_gcf_pad_quirk(1)
    p2 = h.mkGCFLayer(p, dst.ctypes.data, table, arena.ctypes.data, S, ms, pad, 0.02, 9.5)
    h.rcg_gcf_pad_quirk(0)
    assert p2 == p
    assert max_rel(dst.reshape(8, 8, S, S),
                   G.mk_gcf_layer(S, ms, pad, 0.02, 9.5, pad_quirk=True)) <= 1e-12
    h.rcg_gcf_gen_destroy(p)
    rng = np.random.default_rng(8)
    uvw = rng.standard_normal((5000, 3)) * 2.0
    uvw[:, 2] = np.clip(uvw[:, 2], -6.4, 6.4)  # planes -6..6
    sums = np.zeros(13)
    npts = np.zeros(13, np.int32)
    h.calcAccums(uvw.ctypes.data, sums.ctypes.data, npts.ctypes.data, 1.0, 5000, 13)
    es, en = G.calc_accums(uvw, 1.0, 13)
    assert np.array_equal(npts, en)
    assert np.abs(sums - es).max() <= 1e-12 * np.abs(es).max()
