"""GPU parity of the tiled FP32 path against the FP64 oracle.

Inputs are rounded to complex64 first (visibilities, taps; the grid for
degridding) and the oracle runs in FP64 on those rounded values, so the
difference measures the FP32 accumulation only.  Bar: tap counts equal,
relative L2 <= 1e-5 for grids and visibilities."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import device as D
from rc_amd import synth
from rc_amd.tables import VisChunks

pytestmark = pytest.mark.gpu

TOL = 1e-5


def c64_round(vc, gcf):
    vc32 = VisChunks(uvw=vc.uvw, vis=vc.vis.astype(np.complex64).astype(np.complex128),
                     supps=vc.supps, scale=vc.scale, wstep=vc.wstep)
    return vc32, gcf.astype(np.complex64).astype(np.complex128)


def run_tiled(vc, gcf, N, tile, chunk, grid_in=None, layout="pairs", waves=0, degrid_waves=0,
              pitch=None):
    geom = D.DeviceGeom.from_chunks(vc, N, pitch)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    plan = D.TiledPlan(geom, tile=tile, chunk=chunk, layout=layout, waves=waves,
                       degrid_waves=degrid_waves)
    plan.bin(dg)
    vis = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).cuda()
    g = plan.grid(dg, vis)
    gin = g if grid_in is None else torch.from_numpy(grid_in.astype(np.complex64)).cuda()
    v = plan.degrid(dg, gin)
    st = plan.stats()
    torch.cuda.synchronize()
    out = g.cpu().numpy(), v.cpu().numpy().reshape(vc.n_chunks, vc.ts_ch), st
    plan.close()
    return out


LAYOUT_TILES = [("pairs", 16), ("pairs", 32), ("pairs", 64), ("halo", 16), ("halo", 32)]


@pytest.mark.parametrize("layout,tile", LAYOUT_TILES)
@pytest.mark.parametrize("chunk", [3, 512])
@pytest.mark.parametrize("seed,half,cs", [(0, False, "min"), (1, True, "min"),
                                          (2, True, "random")])
def test_tiled_random_vs_oracle(hip_device, layout, tile, chunk, seed, half, cs):
    """Supports 6, 17 and 40 with truncated windows (chunk support below the
    layer's) and, with 'random', odd chunk supports of every size up to the
    plane's: square whole windows, and in the halo layout near-square
    sub-windows of the 33..40-cell windows."""
    vc, gcf, _ = synth.random_small(seed, grid_size=160, chunks=30, ts_ch=20,
                                    half=half, chunk_supp=cs, supports=(6, 17, 40))
    N = 160
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=4)
    ref32 = ref.astype(np.complex64).astype(np.complex128)
    vref, _ = O.degrid(vc, gcf, ref32, nthreads=4)
    g, v, st = run_tiled(vc, gcf, N, tile, chunk, grid_in=ref32, layout=layout)
    assert st["err"] == 0
    assert st["taps"] == taps
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL


@pytest.mark.parametrize("tile", [16, 32, 64])
@pytest.mark.parametrize("waves", [1, 2, 3, 4])
def test_tiled_grid_waves_vs_oracle(hip_device, tile, waves):
    """Every choice of wavefronts (private tile copies) per gridding
    workgroup, on the truncated-window case with chunks of 7 pairs (items
    split over the wavefronts of a workgroup)."""
    vc, gcf, _ = synth.random_small(4, grid_size=160, chunks=30, ts_ch=20, half=True,
                                    chunk_supp="random", supports=(6, 17, 40))
    N = 160
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=4)
    g, v, st = run_tiled(vc, gcf, N, tile, 7, waves=waves)
    assert st["err"] == 0 and st["taps"] == taps
    assert rel_l2(g, ref) <= TOL


@pytest.mark.parametrize("tile", [16, 32, 64])
@pytest.mark.parametrize("degrid_waves", [2, 4, 8])
def test_tiled_degrid_waves_vs_oracle(hip_device, tile, degrid_waves):
    """Every choice of wavefronts per degridding workgroup (sharing one
    staged tile), items of 7 pairs split over them."""
    vc, gcf, _ = synth.random_small(5, grid_size=160, chunks=30, ts_ch=20, half=False,
                                    chunk_supp="random", supports=(6, 17, 40))
    N = 160
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=4)
    ref32 = ref.astype(np.complex64).astype(np.complex128)
    vref, _ = O.degrid(vc, gcf, ref32, nthreads=4)
    g, v, st = run_tiled(vc, gcf, N, tile, 7, grid_in=ref32, degrid_waves=degrid_waves)
    assert st["err"] == 0 and st["taps"] == taps
    assert rel_l2(v, vref) <= TOL


def test_tiled_grid_waves_rejected(hip_device):
    vc, gcf, N = synth.random_small(3)
    geom = D.DeviceGeom.from_chunks(vc, N)
    plan = D.TiledPlan(geom, tile=32, chunk=64)
    from rc_amd import _lib
    with pytest.raises(_lib.RcgError, match="waves"):
        plan.set_grid_waves(5)
    with pytest.raises(_lib.RcgError, match="waves"):
        plan.set_degrid_waves(3)
    plan.close()
    halo = D.TiledPlan(geom, tile=32, chunk=64, layout="halo")
    with pytest.raises(_lib.RcgError, match="halo"):
        halo.set_grid_waves(2)
    halo.close()


@pytest.mark.parametrize("layout,tile", [("pairs", 64), ("halo", 16), ("halo", 32)])
def test_tiled_wproj_slice_vs_oracle(hip_device, layout, tile):
    """128k points of the earth-rotation workload on the full 4096^2 grid
    with 32 w-planes (supports 8..64)."""
    vc, gcf, N = synth.wproj(n_ant=64, n_times=64, vis_dtype=np.complex128)
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=8)
    ref32 = ref.astype(np.complex64).astype(np.complex128)
    vref, _ = O.degrid(vc, gcf, ref32, nthreads=8)
    g, v, st = run_tiled(vc, gcf, N, tile, 512, grid_in=ref32, layout=layout)
    assert st["taps"] == taps and st["err"] == 0
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_tiled_ragged_grid_padded_pitch_vs_oracle(hip_device, tile):
    """A 600^2 grid (not a multiple of any tile: the last row and column of
    tiles are partial) in rows of 613 elements, windows up to the grid's
    edges and a tenth of the points pushed outside it."""
    vc, gcf, _ = synth.random_small(23, grid_size=600, chunks=30, ts_ch=20, max_plane=2,
                                    half=True, supports=(6, 12, 20))
    N, P = 600, 613
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, P, nthreads=4)
    ref32 = ref.astype(np.complex64).astype(np.complex128)
    vref, _ = O.degrid(vc, gcf, ref32, nthreads=4)
    g, v, st = run_tiled(vc, gcf, N, tile, 64, grid_in=ref32, pitch=P)
    assert g.shape == (N, P)
    assert st["err"] == 0 and st["taps"] == taps
    assert rel_l2(g[:, :N], ref[:, :N]) <= TOL
    assert rel_l2(v, vref) <= TOL
    assert np.array_equal(v == 0, vref == 0)  # points left out are 0


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_tiled_one_pair_per_item_vs_oracle(hip_device, tile):
    """chunk = 1: every pair is a work item of its own, so nearly every tile
    has several items (atomic flushes of the copies' sums) and the queues
    hand out one item per pair."""
    vc, gcf, _ = synth.random_small(25, grid_size=160, chunks=30, ts_ch=20, half=True,
                                    supports=(6, 17, 40))
    N = 160
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=4)
    ref32 = ref.astype(np.complex64).astype(np.complex128)
    vref, _ = O.degrid(vc, gcf, ref32, nthreads=4)
    g, v, st = run_tiled(vc, gcf, N, tile, 1, grid_in=ref32)
    assert st["err"] == 0 and st["taps"] == taps
    assert st["items"] == st["pairs"]
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL


def test_tiled_degrid_is_reproducible(hip_device):
    """The FP32 degridder sums each pair's products in a fixed lane order and
    a split point's parts in slot order (no atomics), so degridding is bit
    for bit the same after binning again, which reorders the pairs inside
    every tile (4096^2, 32 planes, 128k points, 64 pairs per item)."""
    vc, gcf, N = synth.wproj(n_ant=64, n_times=64)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    plan = D.TiledPlan(geom, tile=32, chunk=64)
    rng = np.random.default_rng(3)
    g = torch.from_numpy((rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
                         .astype(np.complex64)).cuda()
    outs = []
    for _ in range(3):
        plan.bin(dg)
        outs.append(plan.degrid(dg, g).cpu())
    assert plan.stats()["err"] == 0
    plan.close()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_tiled_all_points_outside(hip_device):
    """Every point outside the grid: no pairs and no taps, a zero grid, and
    zero visibilities degridded from a grid of ones."""
    vc, gcf, N = synth.random_small(24, grid_size=128, chunks=8, ts_ch=5, oob_frac=1.0)
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N)
    assert taps == 0 and not ref.any()
    g, v, st = run_tiled(vc, gcf, N, 32, 64, grid_in=np.ones((N, N)))
    assert st["err"] == 0 and st["taps"] == 0 and st["pairs"] == 0
    assert not g.any() and not v.any()


def test_tiled_support_above_plan_is_reported(hip_device):
    vc, gcf, N = synth.random_small(5, grid_size=96, supports=(6, 8, 10))
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    plan = D.TiledPlan(geom, tile=32, chunk=64, max_support=1)
    plan.bin(dg)
    assert plan.stats()["err"] & 2
    plan.close()


def test_compensated_flush_vs_oracle(hip_device):
    """Small work items (many flushes per tile) on the 4096^2 / 32-plane
    slice: the compensated flush stays within the bar and is at least as
    close to the FP64 oracle as the plain flush."""
    vc, gcf, N = synth.wproj(n_ant=64, n_times=64, vis_dtype=np.complex128)
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=8)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    plan = D.TiledPlan(geom, tile=32, chunk=8)
    plan.bin(dg)
    vis = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).cuda()
    plain = plan.grid(dg, vis).cpu().numpy()
    plan.compensated(True)
    comp = plan.grid(dg, vis).cpu().numpy()
    plan.compensated(False)
    again = plan.grid(dg, vis).cpu().numpy()
    e_plain, e_comp = rel_l2(plain, ref), rel_l2(comp, ref)
    assert plan.stats()["taps"] == taps
    assert e_comp <= TOL and e_plain <= TOL
    assert e_comp <= e_plain
    assert rel_l2(again, ref) <= TOL
    plan.close()


@pytest.mark.parametrize("n,chunk", [(192, 4096), (195, 4096), (387, 4096), (771, 4096),
                                     (771, 192)])
def test_tiled_record_batch_boundaries(hip_device, n, chunk):
    """Every window inside one 32x32 tile (cells 64..95 of a 160^2 grid), so
    the tile holds exactly n pairs; with 3 wavefronts per gridding and 4 per
    degridding workgroup a wavefront walks 64, 65, 129 or 257 pairs (gridder)
    and 48..193 (degridder): record batches of 64 that end exactly at, just
    before and past the item's last pair, and items of 192 pairs."""
    vc, gcf, _ = synth.random_small(11, grid_size=160, chunks=n, ts_ch=1, half=True,
                                    supports=(6, 8, 10), oob_frac=0.0)
    rng = np.random.default_rng(n)
    # u, v in [-10, 10] pixels (scale 1): a window's first cell is 80 - S/2 + floor(u),
    # 65..86 for S in {6, 8, 10}, so every S x S window lies in cells 64..95
    uv = rng.uniform(-10.0, 10.0, (n, 2))
    vc.uvw[:, 0, 0] = uv[:, 0] / vc.scale
    vc.uvw[:, 0, 1] = uv[:, 1] / vc.scale
    N = 160
    vc, gcf = c64_round(vc, gcf)
    ref, taps = O.grid(vc, gcf, N, nthreads=4)
    ref32 = ref.astype(np.complex64).astype(np.complex128)
    vref, _ = O.degrid(vc, gcf, ref32, nthreads=4)
    g, v, st = run_tiled(vc, gcf, N, 32, chunk, grid_in=ref32, waves=3, degrid_waves=4)
    assert st["err"] == 0
    assert st["taps"] == taps
    assert st["pairs"] == n
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(v, vref) <= TOL
