"""Full-size properties of the tiled FP32 path on the headline workload: the
4096^2 w-projection layout with 32 w-planes (supports 8..64, half table),
9,987,840 points from 256 antennas over 306 time samples.

The CPU oracle would need minutes of host time for grids of this size, so
parity here rests on properties that hold at any size; test_gpu_tiled.py
pins the same kernels against the oracle on smaller inputs.

- pregrid records bit-exact against the oracle's pregrid;
- the tap count equals the sum of S^2 over in-bound points;
- degridding an all-ones grid gives every point the sum T_i of its window's
  taps, within the worst-case FP32 summation bound (n + 64) * 2^-24 * sum|t|
  for n = S^2 terms (any order of the additions);
- the sum of all grid cells equals sum_i vis_i * T_i;
- the transpose identity sum_i v_i degrid(G)_i = sum_c G_c grid(v)_c
  (bilinear, no conjugation) for a random grid G;
- grid and degrid within 1e-5 relative L2 of the FP64 GPU path, which
  matches the oracle to 1e-12 on smaller inputs.

Inputs are rounded to complex64 first and the FP64 references use those
rounded values, so every difference measures FP32 accumulation only.
"""
import dataclasses

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import device as D
from rc_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5
U32 = 2.0 ** -24


@pytest.fixture(scope="module")
def work(hip_device):
    vc, gcf, N = synth.wproj(vis_dtype=np.complex128)
    vis = vc.vis.reshape(-1).astype(np.complex64)
    gcf32 = gcf.data.astype(np.complex64)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    plan = D.TiledPlan(geom, tile=32, chunk=512, max_support=64)
    plan.bin(dg)
    w = dict(vc=vc, gcf=gcf, N=N, vis=vis, gcf32=gcf32, geom=geom, dg=dg, plan=plan,
             dvis=torch.from_numpy(vis).cuda())
    yield w
    plan.close()


def oracle_pregrid(vc, N, half):
    """The oracle's per-point pregrid, one call per chunk support."""
    h = O.oracle_lib()
    uvw = np.ascontiguousarray(vc.uvw.reshape(-1, 3), np.float64)
    S = np.repeat(np.asarray(vc.supps, np.int64), vc.ts_ch)
    out = np.zeros((uvw.shape[0], 6), np.int32)
    for s in np.unique(S):
        idx = np.flatnonzero(S == s)
        part = np.ascontiguousarray(uvw[idx])
        res = np.zeros((idx.size, 6), np.int32)
        h.orc_pregrid_many(vc.scale, vc.wstep, part.ctypes.data, idx.size, int(s), N,
                           int(half), res.ctypes.data)
        out[idx] = res
    return out, S


def expected_records(vc, gcf, N, P):
    """Vectorised form of the record rule in test_gpu_pregrid.py."""
    pg, S = oracle_pregrid(vc, N, gcf.half)
    u, v, wp, idx = (pg[:, k].astype(np.int64) for k in (0, 1, 4, 5))
    inb = (idx != np.iinfo(np.int32).min) & (u >= 0) & (u < N - S) & (v >= 0) & (v < N - S)
    slot = np.where(gcf.half, np.abs(idx), idx + gcf.center)
    gs = np.asarray(gcf.supps, np.int64)[np.clip(wp + gcf.max_w_plane, 0, len(gcf.supps) - 1)]
    off = (gs - S) // 2 * (gs + 1)
    ok = inb & (slot >= 0) & (slot < gcf.n_slots) & (gs >= S)
    lo = np.asarray(gcf.layer_off, np.int64)[np.clip(slot, 0, gcf.n_slots - 1)]
    ll = np.asarray(gcf.layer_len, np.int64)[np.clip(slot, 0, gcf.n_slots - 1)]
    ok &= off + (S - 1) * (gs + 1) < ll
    assert np.array_equal(ok, inb), "the workload has no out-of-table points"
    conj = (gcf.half & (wp < 0)).astype(np.int64)
    return dict(inb=inb, cell=u * P + v, tap0=lo + off, stride=gs, supp=S, conj=conj)


@pytest.fixture(scope="module")
def records(work):
    return expected_records(work["vc"], work["gcf"], work["N"], work["geom"].grid_pitch)


def window_sums_of(r, taps32):
    """T_i (complex128, from the complex64 taps, conjugated for mirrored
    layers) and sum |t| per point; 0 for out-of-bound points."""
    taps = taps32.astype(np.complex128)
    inb = r["inb"]
    # a window is fixed by its first tap and support (the stride by its layer)
    key = r["tap0"][inb] * 128 + r["supp"][inb]
    uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
    stride = r["stride"][inb][first]
    tsum = np.empty(len(uk), np.complex128)
    tabs = np.empty(len(uk))
    for j, (k, gs) in enumerate(zip(uk, stride)):
        t0, s = divmod(int(k), 128)
        win = taps[t0 + (np.arange(s)[:, None] * gs + np.arange(s)[None, :])]
        tsum[j], tabs[j] = win.sum(), np.abs(win).sum()
    T = np.zeros(r["inb"].size, np.complex128)
    A = np.zeros(r["inb"].size)
    T[r["inb"]] = tsum[inv.reshape(-1)]
    A[r["inb"]] = tabs[inv.reshape(-1)]
    T = np.where(r["conj"] == 1, np.conj(T), T)
    return T, A


@pytest.fixture(scope="module")
def window_sums(work, records):
    return window_sums_of(records, work["gcf32"])


def test_pregrid_records_bit_exact(work, records):
    pg, taps, err = D.pregrid(work["geom"], work["dg"])
    rec = D.records_to_numpy(pg)
    r = records
    inb = r["inb"]
    assert np.array_equal(rec["cell"] >= 0, inb)
    assert np.array_equal(rec["cell"][inb].astype(np.int64), r["cell"][inb])
    assert np.array_equal(rec["tap0"][inb].astype(np.int64), r["tap0"][inb])
    assert np.array_equal(rec["stride"][inb].astype(np.int64), r["stride"][inb])
    assert np.array_equal(rec["supp"].astype(np.int64), r["supp"])
    assert np.array_equal(rec["flags"][inb].astype(np.int64), r["conj"][inb])
    assert int(taps.item()) == int((r["supp"][inb] ** 2).sum())
    assert int(err.item()) == 0


def test_tap_count(work, records):
    st = work["plan"].stats()
    r = records
    assert st["err"] == 0
    assert st["taps"] == int((r["supp"][r["inb"]] ** 2).sum()) == 3197826880


def test_degrid_of_ones_gives_window_sums(work, window_sums, records):
    T, A = window_sums
    N = work["N"]
    ones = torch.ones((N, work["geom"].grid_pitch), dtype=torch.complex64, device="cuda")
    got = work["plan"].degrid(work["dg"], ones).cpu().numpy().astype(np.complex128)
    n = records["supp"].astype(np.float64) ** 2
    bound = (n + 64) * U32 * A
    err = np.abs(got - T)
    assert np.all(err <= bound), (err / np.maximum(bound, 1e-300)).max()
    assert np.all(got[~records["inb"]] == 0)
    assert rel_l2(got, T) <= TOL


def test_grid_checksum(work, window_sums):
    T, A = window_sums
    g = work["plan"].grid(work["dg"], work["dvis"]).cpu().numpy()
    total = g.astype(np.complex128).sum()
    v = work["vis"].astype(np.complex128)
    exp = (v * T).sum()
    scale = (np.abs(v) * A).sum()
    assert abs(total - exp) <= TOL * scale


def test_transpose_identity(work):
    N, P = work["N"], work["geom"].grid_pitch
    gen = torch.Generator(device="cuda").manual_seed(7)
    G = torch.complex(torch.randn((N, P), generator=gen, device="cuda"),
                      torch.randn((N, P), generator=gen, device="cuda"))
    plan, dg = work["plan"], work["dg"]
    dv = plan.degrid(dg, G).cpu().numpy().astype(np.complex128)
    g = plan.grid(dg, work["dvis"]).cpu().numpy().astype(np.complex128)
    v = work["vis"].astype(np.complex128)
    Gh = G.cpu().numpy().astype(np.complex128)
    lhs = (v * dv).sum()
    rhs = (Gh[:, :N] * g[:, :N]).sum()
    scale = np.sqrt((np.abs(v) ** 2).sum() * (np.abs(dv) ** 2).sum())
    assert abs(lhs - rhs) <= TOL * scale


def test_fp32_vs_fp64_gpu_path(work):
    geom = work["geom"]
    # the complex64-rounded taps, held in FP64
    t64 = D.DeviceGCF.from_table(dataclasses.replace(work["gcf"], data=work["gcf32"]),
                                 dtype=np.complex128)
    pg, _, _ = D.pregrid(geom, t64)
    v64 = torch.from_numpy(work["vis"].astype(np.complex128)).cuda()
    ref = D.grid_plain_f64(geom, t64, pg, v64).cpu().numpy()
    got = work["plan"].grid(work["dg"], work["dvis"]).cpu().numpy()
    assert rel_l2(got, ref) <= TOL
    gref = torch.from_numpy(ref.astype(np.complex64).astype(np.complex128)).cuda()
    vref = D.degrid_plain_f64(geom, t64, pg, gref).cpu().numpy()
    vgot = work["plan"].degrid(work["dg"], gref.to(torch.complex64)).cpu().numpy()
    assert rel_l2(vgot, vref) <= TOL
