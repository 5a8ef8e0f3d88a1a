"""rcg_flow_bin_sizes / rcg_flow_binner (the MS6 visibility binning kernels,
MS6/programs/Kernel/Binning.hs) on the GPU against the oracle
(oracle/flow_oracle.py): counts, bin edges and the w range bit-exact, the
binned rows and the padding exactly; at 1 M points against the invariants
(counts add up, every box holds its rows in input order)."""
import math

import numpy as np
import pytest

from oracle import flow_oracle as F
from rc_amd import _lib, flow

pytestmark = pytest.mark.gpu


def grid_regions(k, lo=-10.0, hi=10.0):
    e = np.linspace(lo, hi, k + 1)
    return np.array([[e[i], e[i + 1], e[j], e[j + 1]] for i in range(k) for j in range(k)])


def random_vis(rng, n, width=5, span=12.0):
    p = np.zeros((n, width))
    p[:, 0] = rng.uniform(-span, span, n)
    p[:, 1] = rng.uniform(-span, span, n)
    p[:, 2] = rng.normal(0, 5, n)
    p[:, 3:] = rng.standard_normal((n, width - 3))
    return p


def same_rows(a, b):
    return a.shape == b.shape and np.array_equal(np.isnan(a), np.isnan(b)) and \
        np.array_equal(np.nan_to_num(a), np.nan_to_num(b))


@pytest.mark.parametrize("width", [5, 7])
def test_bin_sizes_matches_oracle(hip_device, width):
    rng = np.random.default_rng(81 + width)
    p = random_vis(rng, 20000, width)
    regions = grid_regions(8)
    low, high, se, counts = flow.bin_sizes(p, regions, 10)
    elow, ehigh, exp = F.bin_sizes(p, regions, 10)
    assert (low, high) == (elow, ehigh)
    assert np.array_equal(se, exp[:, :, :2])
    assert np.array_equal(counts, exp[:, :, 2].astype(np.int64))


def test_bin_sizes_kats(hip_device):
    """The oracle's hand-worked cases: gap between regions, all-zero w,
    points outside the bounding box, no points."""
    regions = np.array([[0, 4, 0, 10], [6, 10, 0, 10]], float)
    p = np.array([[5.0, 5.0, 1.0, 1, 1], [2.0, 5.0, 1.0, 2, 2], [11.0, 5.0, 9.0, 0, 0]])
    _, _, _, counts = flow.bin_sizes(p, regions, 1)
    assert counts.tolist() == [[2], [0]]
    z = p.copy()
    z[:, 2] = 0.0
    low, high, se, counts = flow.bin_sizes(z, regions, 3)
    assert low == high == 0.0 and counts[:, 0].sum() == 2 and counts[:, 1:].sum() == 0
    assert np.all(se == 0.0)
    low, high, se, counts = flow.bin_sizes(np.zeros((0, 5)), regions, 2)
    assert low == high == 0.0 and counts.sum() == 0


def test_binner_matches_oracle(hip_device):
    rng = np.random.default_rng(85)
    p = random_vis(rng, 20000)
    regions = grid_regions(8)
    _, _, se, counts = flow.bin_sizes(p, regions, 10)
    # the w domain split in two halves: every uv region appears twice, each
    # with half the bins (MS6/programs/binning.hs:46-47 splits the bin domain)
    boxes, box_bins = [], []
    for r in range(len(regions)):
        for half in (range(0, 5), range(5, 10)):
            boxes.append(regions[r])
            box_bins.append([(se[r, b, 0], se[r, b, 1], int(counts[r, b])) for b in half])
    boxes = np.array(boxes)
    outs, filled = flow.binner(p, boxes, box_bins)
    eouts, efilled = F.binner(p, boxes, box_bins, fill=np.nan)
    assert filled.tolist() == efilled
    for a, b in zip(outs, eouts):
        assert same_rows(a, b)
    assert filled.sum() == counts.sum()


def test_binner_kats(hip_device):
    quad = np.array([[-10, 0, -10, 0], [0, 10, -10, 0], [-10, 0, 0, 10], [0, 10, 0, 10]], float)
    b2 = [(-1.0, 1.5, 3), (1.5, 3.5, 1)]
    p = np.array([
        [5, 5, 3.0, 1, 0], [5, 5, 0.0, 2, 0], [5, 5, 3.5, 3, 0], [5, 5, -1.0, 4, 0],
        [-5, -5, 1.0, 5, 0], [5, 5, -2.0, 6, 0], [5, 5, math.nan, 7, 0],
    ], float)
    outs, filled = flow.binner(p, quad, [b2] * 4)
    eouts, efilled = F.binner(p, quad, [b2] * 4, fill=np.nan)
    assert filled.tolist() == efilled == [1, 0, 0, 3]
    for a, b in zip(outs, eouts):
        assert same_rows(a, b)
    # the first box with given lows wins
    boxes = np.array([[0, 10, 0, 10], [0, 5, 0, 5]], float)
    q = np.array([[1.0, 1.0, 0.5, 1, 1]])
    _, filled = flow.binner(q, boxes, [[(0.0, 1.0, 2)], [(0.0, 1.0, 2)]])
    assert filled.tolist() == [1, 0]


def test_flow_errors(hip_device):
    regions = np.array([[0, 4, 0, 10]], float)
    p = np.array([[1.0, 1.0, math.nan, 0, 0]])
    with pytest.raises(_lib.RcgError, match="NaN"):
        flow.bin_sizes(p, regions, 2)
    # NaN w outside the bounding box is not in the fold
    p[0, 0] = 50.0
    flow.bin_sizes(p, regions, 2)
    q = np.array([[1.0, 1.0, 0.5, 1, 1]] * 3)
    with pytest.raises(_lib.RcgError, match="holds"):
        flow.binner(q, regions, [[(0.0, 1.0, 2)]])
    with pytest.raises(_lib.RcgError, match="width"):
        flow.binner(np.zeros((3, 6)), regions, [[(0.0, 1.0, 2)]])
    assert _lib.last_error()[0] == 1


def test_flow_one_million_points(hip_device):
    """32 x 32 uv regions (a 2048 grid in 64-cell tiles), 10 w bins."""
    rng = np.random.default_rng(86)
    n = 1_000_000
    p = random_vis(rng, n, span=1100.0)
    p[:, 2] *= 40
    regions = grid_regions(32, -1024.0, 1024.0)
    low, high, se, counts = flow.bin_sizes(p, regions, 10)
    inside = (p[:, 0] >= -1024) & (p[:, 0] < 1024) & (p[:, 1] >= -1024) & (p[:, 1] < 1024)
    assert counts.sum() == inside.sum()
    assert low == min(0.0, p[inside, 2].min())
    assert high == p[inside, 2].max() + (p[inside, 2].max() - low) * 0.0001
    box_bins = [[(se[r, b, 0], se[r, b, 1], int(counts[r, b])) for b in range(10)]
                for r in range(len(regions))]
    outs, filled = flow.binner(p, regions, box_bins)
    assert np.array_equal(filled, counts.sum(axis=1))
    # per region: its rows in input order
    ui = np.floor((p[:, 0] + 1024) / 64).astype(int)
    vi = np.floor((p[:, 1] + 1024) / 64).astype(int)
    rid = np.where(inside, ui * 32 + vi, -1)
    order = np.argsort(rid, kind="stable")
    starts = np.searchsorted(rid[order], np.arange(len(regions)))
    for r in range(0, len(regions), 37):
        sel = order[starts[r]:starts[r] + filled[r]]
        assert np.array_equal(outs[r][:filled[r]], p[sel])
