"""GPU parity of the MS6 Halide buffer_t entry points (kern_scatter,
kern_degrid and the _gpu1 variants) against oracle/halide_oracle.py.
Bar: FP64, within 1e-12 of the largest value."""
import numpy as np
import pytest
import torch

from oracle import halide_oracle as H
from rc_amd import halide as HB

pytestmark = pytest.mark.gpu


def case(S, n=3000, seed=0, N=256):
    rng = np.random.default_rng(seed + S)
    lim = 100.0
    vis = np.ascontiguousarray(np.column_stack([
        rng.uniform(-lim, lim, n), rng.uniform(-lim, lim, n), rng.standard_normal(n),
        rng.standard_normal(n), rng.standard_normal(n)]))
    gcf = np.ascontiguousarray(rng.standard_normal((64, S, S, 2)))
    return vis, gcf, N


def close(a, b):
    return np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("S", [8, 16, 32, 64])
@pytest.mark.parametrize("gpu1", [False, True])
def test_scatter_vs_oracle(hip_device, S, gpu1):
    vis, gcf, N = case(S)
    # a 200 x 180 window of the grid starting at (20, 30), row pitch 210
    origin, width, height, pitch = (20, 30), 200, 180, 210
    uvg = np.zeros((height, pitch, 2))
    uvg[..., 0] = 0.5  # the kernel adds to what is there
    exp = uvg.copy()
    H.scatter(vis, gcf, exp, 0.9, N, 6, width=width, origin=origin)
    rc = HB.scatter(0.9, N, 6, HB.vis_buffer(vis), HB.gcf_buffer(gcf),
                    HB.uvg_buffer(uvg, width, origin), gpu1=gpu1)
    assert rc == 0
    assert close(uvg, exp)
    assert np.count_nonzero(exp[..., 1]) > 0


@pytest.mark.parametrize("S", [8, 16, 32, 64])
@pytest.mark.parametrize("gpu1", [False, True])
def test_scatter_square_window_vs_oracle(hip_device, S, gpu1):
    """A square 224 x 224 window of the grid at (16, 24), row pitch 230: the
    tiled FP64 route (points binned by 32 x 32 tile, tile sums added to the
    grid).  The pad columns past the window stay as they were."""
    vis, gcf, N = case(S, n=20000, seed=3)
    origin, width, pitch = (16, 24), 224, 230
    uvg = np.zeros((width, pitch, 2))
    uvg[..., 1] = -0.25
    exp = uvg.copy()
    H.scatter(vis, gcf, exp, 0.9, N, 6, width=width, origin=origin)
    rc = HB.scatter(0.9, N, 6, HB.vis_buffer(vis), HB.gcf_buffer(gcf),
                    HB.uvg_buffer(uvg, width, origin), gpu1=gpu1)
    assert rc == 0
    assert close(uvg, exp)
    assert np.array_equal(uvg[:, width:], exp[:, width:])
    assert np.count_nonzero(exp[..., 0]) > S * S


@pytest.mark.parametrize("S", [8, 32])
@pytest.mark.parametrize("gpu1", [False, True])
def test_degrid_vs_oracle(hip_device, S, gpu1):
    vis, gcf, N = case(S, seed=5)
    rng = np.random.default_rng(9)
    origin, width, height, pitch = (10, 12), 230, 220, 236
    uvg = np.ascontiguousarray(rng.standard_normal((height, pitch, 2)))
    exp = H.degrid(gcf, uvg, vis, 0.9, N, 0, width=width, origin=origin)
    out = np.zeros_like(vis)
    rc = HB.degrid(0.9, N, 0, HB.gcf_buffer(gcf), HB.uvg_buffer(uvg, width, origin),
                   HB.vis_buffer(vis), HB.vis_buffer(out), gpu1=gpu1)
    assert rc == 0
    assert close(out, exp)
    assert np.array_equal(out[:, :3], vis[:, :3])
    assert np.count_nonzero(out[:, 3] - vis[:, 3]) > 0


def test_gpu1_reads_device_buffers_and_clears_dev(hip_device):
    vis, gcf, N = case(16, seed=2)
    dvis = torch.from_numpy(vis).cuda()
    dgcf = torch.from_numpy(gcf).cuda()
    uvg = np.zeros((N, N, 2))
    exp = uvg.copy()
    H.scatter(vis, gcf, exp, 0.9, N, 0)
    # host copies poisoned (points far outside the grid, NaN taps): the
    # device pointers must be used
    vb = HB.vis_buffer(np.full_like(vis, 1e300), dev=dvis.data_ptr())
    gb = HB.gcf_buffer(np.full_like(gcf, np.nan), dev=dgcf.data_ptr())
    ub = HB.uvg_buffer(uvg)
    assert HB.scatter(0.9, N, 0, vb, gb, ub, gpu1=True) == 0
    assert close(uvg, exp)
    assert vb.dev == 0 and gb.dev == 0 and ub.dev == 0
    # a dirty host copy wins over the device pointer (zeros there)
    zeros = torch.zeros_like(dvis)
    uvg2 = np.zeros((N, N, 2))
    vb2 = HB.vis_buffer(vis, dev=zeros.data_ptr(), host_dirty=True)
    assert HB.scatter(0.9, N, 0, vb2, HB.gcf_buffer(gcf), HB.uvg_buffer(uvg2), gpu1=True) == 0
    assert close(uvg2, exp)


def test_unsupported_gcf_size_codes(hip_device):
    vis, _, N = case(8, n=10)
    gcf = np.zeros((64, 12, 12, 2))
    uvg = np.zeros((N, N, 2))
    out = np.zeros_like(vis)
    args = (HB.vis_buffer(vis), HB.gcf_buffer(gcf), HB.uvg_buffer(uvg))
    assert HB.scatter(1.0, N, 0, *args) == -444
    assert HB.scatter(1.0, N, 0, *args, gpu1=True) == -222
    dargs = (HB.gcf_buffer(gcf), HB.uvg_buffer(uvg), HB.vis_buffer(vis), HB.vis_buffer(out))
    assert HB.degrid(1.0, N, 0, *dargs) == -555
    assert HB.degrid(1.0, N, 0, *dargs, gpu1=True) == -333


# --- nvGridder ---------------------------------------------------------------

def nv_case(n, N, scale, seed):
    rng = np.random.default_rng(seed)
    lim = N / scale
    vis = np.ascontiguousarray(np.column_stack([
        rng.uniform(0, lim, n), rng.uniform(0, lim, n), rng.standard_normal(n),
        rng.standard_normal(n), rng.standard_normal(n)]))
    gcf = np.ascontiguousarray(rng.standard_normal((64, 16, 16, 2)))
    return vis, gcf


@pytest.mark.parametrize("scale,N", [(1.0, 256), (0.5, 200)])
def test_nv_gridder_vs_oracle(hip_device, scale, N):
    vis, gcf = nv_case(3000, N, scale, 11)
    order = H.nv_sort_order(vis)
    exp = H.nv_grid(vis[order], gcf, scale, N)
    uvg = np.full((N, N, 2), 7.0)  # overwritten, not added to
    vb, gb, ub = HB.vis_buffer(vis), HB.nv_gcf_buffer(gcf), HB.uvg_buffer(uvg)
    vb.dev, ub.dev, ub.dev_dirty = 5, 5, True
    sorted_exp = vis[order].copy()
    assert HB.nv_gridder(scale, N, vb, gb, ub) == 0
    # the host visibility array is sorted in place
    assert np.array_equal(vis, sorted_exp)
    assert close(uvg, exp)
    assert vb.dev == 0 and ub.dev == 0 and not ub.dev_dirty and not vb.host_dirty


def test_nv_gridder_edges_and_errors(hip_device):
    N = 64
    # points near and beyond the edges, and far outside
    vis = np.ascontiguousarray(np.array([[0.2, 0.7, 0, 1.0, 0.5], [63.9, 10.0, 0, -1.0, 2.0],
                                         [70.5, 30.0, 0, 1.0, 1.0], [1e300, 5.0, 0, 1.0, 1.0],
                                         [30.0, 71.0, 0, 2.0, 0.0]]))
    gcf = np.ascontiguousarray(np.random.default_rng(2).standard_normal((64, 16, 16, 2)))
    exp = H.nv_grid(vis, gcf, 1.0, N)  # the far point contributes nothing
    uvg = np.zeros((N, N, 2))
    assert HB.nv_gridder(1.0, N, HB.vis_buffer(vis), HB.nv_gcf_buffer(gcf),
                         HB.uvg_buffer(uvg)) == 0
    assert close(uvg, exp)
    # a grid buffer smaller than grid_size^2 cells is refused
    small = np.zeros((N - 1, N, 2))
    assert HB.nv_gridder(1.0, N, HB.vis_buffer(vis), HB.nv_gcf_buffer(gcf),
                         HB.uvg_buffer(small)) == -1


def test_nv_gridder_subpixel_carry(hip_device):
    """x = -1e-300 (and y likewise): floor -1 with offset 8, carried to
    pixel 0, offset 0 as the reference's packed coordinates do; both the
    tiled route (a point inside) and the scatter route (points at the
    edge)."""
    N = 64
    gcf = np.ascontiguousarray(np.random.default_rng(4).standard_normal((64, 16, 16, 2)))
    vis = np.ascontiguousarray(np.array([[-1e-300, 20.0, 0, 1.0, 0.5],
                                         [30.0, -1e-300, 0, 2.0, -1.0],
                                         [40.0 - 2.0 ** -48, 33.0, 0, 1.0, 1.0]]))
    exp = H.nv_grid(vis, gcf, 1.0, N)
    assert H.nv_pixel(-1e-300, 20.0, 1.0) == (0, 20, 0)
    uvg = np.zeros((N, N, 2))
    assert HB.nv_gridder(1.0, N, HB.vis_buffer(vis), HB.nv_gcf_buffer(gcf),
                         HB.uvg_buffer(uvg)) == 0
    assert close(uvg, exp)


def test_bin_sort_and_raw_gridders(hip_device):
    """bin_sort3 / bin_sort5 and gridGPUc / gridGPUs (grid_gpu.cuh:27-43)."""
    vis, gcf = nv_case(2000, 128, 1.0, 4)
    order = H.nv_sort_order(vis)
    five = vis.copy()
    HB.bin_sort(five)
    assert np.array_equal(five, vis[order])
    three = np.ascontiguousarray(vis[:, :3])
    HB.bin_sort(three)
    assert np.array_equal(three, vis[order][:, :3])
    exp = H.nv_grid(vis, gcf, 1.0, 128)
    # gridGPUc does not sort: the caller's array is left as it is
    c = vis.copy()
    assert close(HB.grid_gpu(1.0, gcf, 128, combined=c), exp)
    assert np.array_equal(c, vis)
    pos = np.ascontiguousarray(vis[:, :3])
    vals = np.ascontiguousarray(vis[:, 3:])
    assert close(HB.grid_gpu(1.0, gcf, 128, pos=pos, vals=vals), exp)


# --- MS6 imaging glue --------------------------------------------------------

def test_init_and_psf_vis(hip_device):
    g = np.ones((6, 5, 2))
    assert HB.init(HB.buffer_of(g, (0, 3, 4))) == 0 and not g.any()
    im = np.ones((7, 9))
    assert HB.image_init(HB.buffer_of(im, (2, 2))) == 0 and not im.any()
    rng = np.random.default_rng(0)
    vis = rng.standard_normal((50, 5))
    out = np.zeros((20, 5))
    # output points 10..29 of an input holding points 5..54
    assert HB.psf_vis(HB.buffer_of(vis, (0, 5)), HB.buffer_of(out, (0, 10))) == 0
    assert np.array_equal(out, H.psf_vis(vis[5:25]))
    # an output that reaches past the input is refused
    assert HB.psf_vis(HB.buffer_of(vis, (0, 40)), HB.buffer_of(out, (0, 30))) == -1


@pytest.mark.parametrize("uvproj", [0, 1])
def test_rotate_vs_oracle(hip_device, uvproj):
    rng = np.random.default_rng(5 + uvproj)
    vis = np.column_stack([rng.uniform(-3e3, 3e3, (400, 3)), rng.standard_normal((400, 2))])
    vis = np.ascontiguousarray(vis)
    out = np.zeros_like(vis)
    args = (0.4, -0.8, 0.39, -0.805, 1e-3, 2e-3, uvproj)
    ob = HB.buffer_of(out)
    ob.min[2], ob.min[3] = 3, -2          # the facet (l, m)
    assert HB.rotate(*args, HB.buffer_of(vis), ob) == 0
    exp = H.rotate(vis, *args, 3, -2)
    assert close(out, exp)


def test_detile_defacet_image_sum_vs_oracle(hip_device):
    rng = np.random.default_rng(8)
    a = rng.standard_normal((30, 40, 2))
    out = rng.standard_normal((50, 35, 2))
    exp = out.copy()
    H.detile(a, (10, 5), exp, (25, 0))
    assert HB.detile(HB.buffer_of(a, (0, 10, 5)), HB.buffer_of(out, (0, 25, 0))) == 0
    assert close(out, exp)
    # facets at l, m in 0..2 shifted by x + l, y + m: neighbours overlap and
    # the right and bottom edges clamp onto the border, so many pixels have
    # several writers and the last one in Halide order must win
    facet = rng.standard_normal((3, 3, 16, 16))       # [m][l][y][x]
    img = rng.standard_normal((17, 17))
    expi = img.copy()
    H.defacet(facet, (0, 0, 0, 0), expi, (0, 0))
    assert HB.defacet(HB.buffer_of(facet), HB.buffer_of(img)) == 0
    assert np.array_equal(img, expi)
    src = rng.standard_normal((20, 20))
    s = np.zeros((10, 10))
    exps = s.copy()
    H.image_sum(src, (-3, -3), exps, (2, 2))
    assert HB.image_sum(HB.buffer_of(src, (-3, -3)), HB.buffer_of(s, (2, 2))) == 0
    assert close(s, exps)
