"""GPU parity of rcg_grid_tiled_f64 and rcg_degrid_tiled_f64 (FP64 gridding
and degridding over a tiled plan's pairs, the kernels of the MS4
host-memory entry points) against the FP64 oracle.  Every tap adds the
reference's complex product in FP64
(MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p.cpp:169 and
:368-431); only the summation order differs, so the bar is the MS4 one: tap
counts equal, relative L2 <= 1e-12."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import device as D
from rc_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-12


def run_f64(vc, gcf, N, tile, chunk, pitch=None, grid_in=None):
    """(grid, stats) of rcg_grid_tiled_f64, and with grid_in the
    visibilities rcg_degrid_tiled_f64 takes from it."""
    geom = D.DeviceGeom.from_chunks(vc, N, pitch)
    dg = D.DeviceGCF.from_table(gcf)
    plan = D.TiledPlan(geom, tile=tile, chunk=chunk)
    plan.bin(dg)
    vis = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex128)).cuda()
    g = plan.grid_f64(dg, vis)
    st = plan.stats()
    v = None
    if grid_in is not None:
        v = plan.degrid_f64(dg, torch.from_numpy(np.ascontiguousarray(grid_in)).cuda())
        v = v.cpu().numpy().reshape(vc.n_chunks, vc.ts_ch)
    torch.cuda.synchronize()
    out = g.cpu().numpy(), st
    plan.close()
    return out if grid_in is None else out + (v,)


@pytest.mark.parametrize("tile", [16, 32, 64])
@pytest.mark.parametrize("chunk", [3, 1024])
@pytest.mark.parametrize("seed,half,cs", [(0, False, "min"), (1, True, "min"),
                                          (2, True, "random")])
def test_tiled_f64_random_vs_oracle(hip_device, tile, chunk, seed, half, cs):
    vc, gcf, _ = synth.random_small(seed, grid_size=160, chunks=30, ts_ch=20,
                                    half=half, chunk_supp=cs, supports=(6, 17, 40))
    N = 160
    ref, taps = O.grid(vc, gcf, N, nthreads=4)
    vref, dtaps = O.degrid(vc, gcf, ref, nthreads=4)
    g, st, v = run_f64(vc, gcf, N, tile, chunk, grid_in=ref)
    assert st["err"] == 0 and st["taps"] == taps == dtaps
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL
    assert np.array_equal(v == 0, vref == 0)  # points left out are 0


def test_tiled_f64_padded_pitch_large_supports(hip_device):
    """Supports above one wavefront row (one lane per cell, several trips
    per lane) and a padded pitch."""
    vc, gcf, N = synth.random_small(21, grid_size=320, chunks=6, ts_ch=7, max_plane=2,
                                    half=True, supports=(64, 72, 80), chunk_supp="random")
    ref, taps = O.grid(vc, gcf, N, N + 5)
    vref, _ = O.degrid(vc, gcf, ref)
    g, st, v = run_f64(vc, gcf, N, 32, 64, N + 5, grid_in=ref)
    assert st["err"] == 0 and st["taps"] == taps
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_tiled_f64_ragged_grid_padded_pitch_vs_oracle(hip_device, tile):
    """A 600^2 grid in rows of 613 elements: the last row and column of
    tiles are partial at every tile size, windows reach the grid's edges and
    a tenth of the points lie outside it."""
    vc, gcf, _ = synth.random_small(23, grid_size=600, chunks=30, ts_ch=20, max_plane=2,
                                    half=True, supports=(6, 12, 20))
    N, P = 600, 613
    ref, taps = O.grid(vc, gcf, N, P, nthreads=4)
    vref, _ = O.degrid(vc, gcf, ref, nthreads=4)
    g, st, v = run_f64(vc, gcf, N, tile, 64, P, grid_in=ref)
    assert st["err"] == 0 and st["taps"] == taps
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL
    assert np.array_equal(v == 0, vref == 0)


def test_tiled_f64_all_points_outside(hip_device):
    """Every point outside the grid: no taps, a zero grid, zero
    visibilities."""
    vc, gcf, N = synth.random_small(24, grid_size=128, chunks=8, ts_ch=5, oob_frac=1.0)
    ref, taps = O.grid(vc, gcf, N)
    assert taps == 0 and not ref.any()
    g, st, v = run_f64(vc, gcf, N, 32, 64, grid_in=ref + 1.0)
    assert st["err"] == 0 and st["taps"] == 0 and st["pairs"] == 0
    assert not g.any() and not v.any()


def test_tiled_f64_wproj_slice_vs_oracle(hip_device):
    """128k points of the earth-rotation workload on the full 4096^2 grid
    with 32 w-planes (supports 8..64), complex128 throughout."""
    vc, gcf, N = synth.wproj(n_ant=64, n_times=64, vis_dtype=np.complex128)
    ref, taps = O.grid(vc, gcf, N, nthreads=8)
    vref, _ = O.degrid(vc, gcf, ref, nthreads=8)
    g, st, v = run_f64(vc, gcf, N, 32, 1024, grid_in=ref)
    assert st["err"] == 0 and st["taps"] == taps
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL


def test_tiled_f64_type_checks(hip_device):
    vc, gcf, N = synth.random_small(3)
    geom = D.DeviceGeom.from_chunks(vc, N)
    plan = D.TiledPlan(geom, tile=32, chunk=64)
    d64 = D.DeviceGCF.from_table(gcf)
    plan.bin(d64)
    with pytest.raises(TypeError):
        plan.grid_f64(D.DeviceGCF.from_table(gcf, dtype=np.complex64),
                      torch.zeros(vc.n, dtype=torch.complex128, device="cuda"))
    with pytest.raises(TypeError):
        plan.grid_f64(d64, torch.zeros(vc.n, dtype=torch.complex64, device="cuda"))
    plan.close()


def test_tiled_f64_batch_of_instances(hip_device):
    """A plan over two instances of one grid shape: each instance's grid
    matches the oracle on its own points."""
    N = 160
    cases = [synth.random_small(s, grid_size=N, chunks=20, ts_ch=15, half=True,
                                supports=(6, 17, 40)) for s in (31, 32)]
    gcf = cases[0][1]
    geoms = [D.DeviceGeom.from_chunks(vc, N) for vc, _, _ in cases]
    dg = D.DeviceGCF.from_table(gcf)
    plan = D.TiledPlan(geoms, tile=32, chunk=16)
    plan.bin(dg)
    vis = torch.from_numpy(np.concatenate([vc.vis.reshape(-1) for vc, _, _ in cases])
                           .astype(np.complex128)).cuda()
    g = plan.grid_f64(dg, vis).cpu().numpy()
    st = plan.stats()
    plan.close()
    assert g.shape == (2, N, N) and st["err"] == 0
    total = 0
    for k, (vc, _, _) in enumerate(cases):
        ref, taps = O.grid(vc, gcf, N)
        total += taps
        assert rel_l2(g[k], ref) <= TOL
    assert st["taps"] == total
