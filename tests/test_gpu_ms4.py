"""GPU parity of the MS4 entry points (host-memory ABI over HIP, FP64):
against the golden vectors of the reference and against the oracle.
Bar: tap counts equal, grids and visibilities within 1e-12 relative L2
(FP64 atomics make the summation order differ from the sequential one)."""
import numpy as np
import pytest

from _fixtures import RANDOM_FIXTURES, load_probe, load_random
from conftest import GOLDEN, rel_l2
from oracle import oracle as O
from rc_amd import _lib, ms4, synth
from rc_amd.tables import GCFTable, VisChunks

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("name", RANDOM_FIXTURES)
def test_grid_degrid_golden(hip_device, name):
    vc, gcf, N, z = load_random(name)
    g, taps = ms4.grid(vc, gcf, N)
    assert taps == int(z["taps"])
    assert rel_l2(g, z["grid"]) <= TOL
    vis, dtaps = ms4.degrid(vc, gcf, z["grid"])
    assert dtaps == int(z["dtaps"])
    assert rel_l2(vis, z["vis_out"]) <= TOL


def test_pregrid_probe_golden(hip_device):
    """Single points with an encoded GCF: the GPU must hit the same cell and
    the same tap as the reference (exact)."""
    from golden.make_golden import encoded_table
    for half, scale, wstep, N, pts, S, res in load_probe():
        sup = {p: 6 + 2 * abs(p) for p in range(-3 if not half else 0, 4)}
        gcf = encoded_table(sup, half)
        vc = VisChunks(uvw=pts.reshape(-1, 1, 3), vis=np.ones((len(pts), 1), np.complex128),
                       supps=S.astype(np.int32), scale=scale, wstep=wstep)
        for i in range(len(pts)):
            one = VisChunks(uvw=vc.uvw[i:i + 1], vis=vc.vis[i:i + 1], supps=vc.supps[i:i + 1],
                            scale=scale, wstep=wstep)
            g, taps = ms4.grid(one, gcf, N)
            nz = np.argwhere(g != 0)
            if len(nz) == 0:
                got = (-1, -1, 0.0, 0.0, taps)
            else:
                u0, v0 = nz.min(axis=0)
                got = (u0, v0, g[u0, v0].real, g[u0, v0].imag, taps)
            assert np.array_equal(np.asarray(got, np.float64), res[i]), (half, scale, pts[i])


def test_c1_full_size(hip_device):
    """512^2 grid, 100k points: GPU vs the oracle and the golden summary."""
    z = np.load(GOLDEN / "c1_summary.npz")
    vc, gcf, N = synth.plain_c1()
    g, taps = ms4.grid(vc, gcf, N)
    assert taps == int(z["taps"])
    ref, _ = O.grid(vc, gcf, N, nthreads=8)
    assert rel_l2(g, ref) <= TOL
    assert rel_l2(g.reshape(-1)[z["cells"]], z["cell_vals"]) <= TOL
    gn = ref / (N * N)
    vis, dtaps = ms4.degrid(vc, gcf, gn)
    assert dtaps == int(z["dtaps"])
    vref, _ = O.degrid(vc, gcf, gn, nthreads=8)
    assert rel_l2(vis, vref) <= TOL
    assert rel_l2(vis.reshape(-1)[z["pts"]], z["pt_vals"]) <= TOL


@pytest.mark.parametrize("supports,half,pitch_pad", [
    ((64, 72, 80), False, 0),     # supports above one wavefront row
    ((1, 3, 5), True, 3),         # tiny odd supports, padded pitch
    ((128, 130, 132), True, 0),   # the largest supports of the MS4 config
])
def test_support_shapes_vs_oracle(hip_device, supports, half, pitch_pad):
    vc, gcf, N = synth.random_small(21, grid_size=320, chunks=6, ts_ch=7, max_plane=2,
                                    half=half, supports=supports, chunk_supp="random")
    P = N + pitch_pad
    g, taps = ms4.grid(vc, gcf, N, P)
    ref, rtaps = O.grid(vc, gcf, N, P)
    assert taps == rtaps
    assert rel_l2(g, ref) <= TOL
    vis, dtaps = ms4.degrid(vc, gcf, ref, N)
    vref, _ = O.degrid(vc, gcf, ref, N)
    assert dtaps == rtaps
    assert rel_l2(vis, vref) <= TOL


def test_supports_above_225_bin_at_tile_64(hip_device):
    """Chunk supports above 225 span more than 64 tiles of 32, so the grid
    call bins its plan at T = 64 instead (supports 232 and 256, the largest
    the path accepts)."""
    vc, gcf, N = synth.random_small(22, grid_size=600, chunks=6, ts_ch=7, max_plane=1,
                                    half=True, supports=(232, 256), chunk_supp="min")
    g, taps = ms4.grid(vc, gcf, N)
    ref, rtaps = O.grid(vc, gcf, N)
    assert taps == rtaps > 0
    assert rel_l2(g, ref) <= TOL


def test_empty_and_all_out_of_bound(hip_device):
    vc, gcf, N = synth.random_small(3)
    empty = VisChunks(uvw=np.zeros((0, 5, 3)), vis=np.zeros((0, 5), np.complex128),
                      supps=np.zeros(0, np.int32), scale=1.0, wstep=1.0)
    g, taps = ms4.grid(empty, gcf, N)
    assert taps == 0 and not np.any(g)
    far = VisChunks(uvw=vc.uvw + np.array([10 * N, 0.0, 0.0]), vis=vc.vis, supps=vc.supps,
                    scale=1.0, wstep=vc.wstep)
    g, taps = ms4.grid(far, gcf, N)
    assert taps == 0 and not np.any(g)
    vis, dtaps = ms4.degrid(far, gcf, np.ones((N, N), np.complex128))
    assert dtaps == 0 and not np.any(vis)


def test_window_larger_than_layer_is_an_error(hip_device):
    """Chunk support above the plane support reads outside the layer in the
    reference (undefined); here the call fails with RCG_ERR_RANGE and a
    zeroed grid."""
    vc, gcf, N = synth.random_small(4, supports=(4, 6, 8))
    bad = VisChunks(uvw=vc.uvw, vis=vc.vis, supps=np.full_like(vc.supps, 9), scale=1.0,
                    wstep=vc.wstep)
    with pytest.raises(_lib.RcgError, match="GCF layer"):
        ms4.grid(bad, gcf, N)
    assert _lib.last_error()[0] == 4
