"""The oracle of the MS6 visibility binning kernels (oracle/flow_oracle.py,
restating MS6/programs/Kernel/Binning.hs binSizer / binner) against
hand-worked KATs and invariants.  GHC is absent, so the Haskell cannot run;
these checks are the pinning."""
import math

import numpy as np
import pytest

from oracle import flow_oracle as F

QUAD = np.array([[-10, 0, -10, 0], [0, 10, -10, 0], [-10, 0, 0, 10], [0, 10, 0, 10]], float)


def kat_points():
    return np.array([
        [-5, -5, 2.0, 1, 0],
        [5, -5, -1.0, 2, 0],
        [5, 5, 3.0, 3, 0],
        [20, 5, 100.0, 4, 0],   # outside the bounding box: not in the fold
        [-5, 5, 0.5, 5, 0],
    ], float)


def test_bin_sizes_kat():
    low, high, out = F.bin_sizes(kat_points(), QUAD, 4)
    assert low == -1.0
    assert high == 3.0 + 4.0 * 0.0001
    # bins of (w + 1) / 4.0004 * 4: 2.9997 -> 2, 0 -> 0, 3.9996 -> 3, 1.49985 -> 1
    exp = np.zeros((4, 4))
    exp[0, 2] = exp[1, 0] = exp[3, 3] = exp[2, 1] = 1
    assert np.array_equal(out[:, :, 2], exp)
    assert out[0, 1, 0] == -1.0 + 1.0 * (high + 1.0) / 4.0
    assert out[0, 3, 1] == -1.0 + 4.0 * (high + 1.0) / 4.0
    assert np.array_equal(out[0, :, :2], out[3, :, :2])


def test_fold_starts_at_zero():
    """All w positive: low stays 0 (the fold starts from (0, 0))."""
    p = kat_points()
    p[:, 2] = [2.0, 1.0, 3.0, 100.0, 0.5]
    low, high, _ = F.bin_sizes(p, QUAD, 2)
    assert low == 0.0 and high == 3.0 + 3.0 * 0.0001


def test_all_zero_w_goes_to_bin_zero():
    """low = high = 0: (w - low) / (high - low) is NaN, which GHC's floor
    turns into minBound, clamped to bin 0."""
    p = kat_points()
    p[:, 2] = 0.0
    low, high, out = F.bin_sizes(p, QUAD, 3)
    assert low == high == 0.0
    assert out[:, :, 2].sum() == 4 and out[:, 0, 2].sum() == 4
    assert np.all(out[:, :, :2] == 0.0)


def test_gap_counted_by_sizer_dropped_by_binner():
    regions = np.array([[0, 4, 0, 10], [6, 10, 0, 10]], float)
    p = np.array([[5.0, 5.0, 1.0, 1, 1], [2.0, 5.0, 1.0, 2, 2]])
    _, _, out = F.bin_sizes(p, regions, 1)
    # u = 5 lies in the gap; lookupLE finds the region with u_low 0
    assert out[0, 0, 2] == 2 and out[1, 0, 2] == 0
    bins = [[(0.0, 2.0, 2)], [(0.0, 2.0, 0)]]
    outs, filled = F.binner(p, regions, bins, fill=np.nan)
    assert filled == [1, 0]
    assert np.array_equal(outs[0][0], p[1])
    # the second row of box 0 is zero padding (bin size 2, one row filled)
    assert np.array_equal(outs[0][1], np.zeros(5))


def test_binner_order_padding_and_ranges():
    boxes = QUAD
    b2 = [(-1.0, 1.5, 3), (1.5, 3.5, 1)]
    bins = [b2, b2, b2, b2]
    p = np.array([
        [5, 5, 3.0, 1, 0],    # box 3, upper bin
        [5, 5, 0.0, 2, 0],    # box 3, lower bin
        [5, 5, 3.5, 3, 0],    # w = w_high of the last bin: dropped
        [5, 5, -1.0, 4, 0],   # box 3, lower bin (w = w_low)
        [-5, -5, 1.0, 5, 0],  # box 0
        [5, 5, -2.0, 6, 0],   # below every w_low: dropped
    ], float)
    outs, filled = F.binner(p, boxes, bins, fill=np.nan)
    assert filled == [1, 0, 0, 3]
    # input order across the box's bins
    assert np.array_equal(outs[3][:3], p[[0, 1, 3]])
    assert np.isnan(outs[3][3]).all()          # beyond every bin's size: untouched
    assert np.array_equal(outs[0][0], p[4])
    assert np.array_equal(outs[0][1:3], np.zeros((2, 5)))   # padded to 3
    assert np.isnan(outs[0][3]).all()
    assert np.array_equal(outs[1][:3], np.zeros((3, 5)))


def test_binner_first_box_wins_and_overflow():
    boxes = np.array([[0, 10, 0, 10], [0, 5, 0, 5]], float)   # same lows
    bins = [[(0.0, 1.0, 2)], [(0.0, 1.0, 2)]]
    p = np.array([[1.0, 1.0, 0.5, 1, 1]])
    _, filled = F.binner(p, boxes, bins)
    assert filled == [1, 0]
    with pytest.raises(OverflowError):
        F.binner(np.repeat(p, 3, axis=0), boxes, bins)


def test_sizer_and_binner_agree_on_tiled_regions():
    rng = np.random.default_rng(71)
    n = 3000
    p = np.zeros((n, 5))
    p[:, 0] = rng.uniform(-12, 12, n)
    p[:, 1] = rng.uniform(-12, 12, n)
    p[:, 2] = rng.normal(0, 5, n)
    p[:, 3:] = rng.standard_normal((n, 2))
    edges = np.linspace(-10, 10, 5)
    regions = np.array([[edges[i], edges[i + 1], edges[j], edges[j + 1]]
                        for i in range(4) for j in range(4)])
    low, high, out = F.bin_sizes(p, regions, 6)
    inside = ((p[:, 0] >= -10) & (p[:, 0] < 10) & (p[:, 1] >= -10) & (p[:, 1] < 10))
    assert out[:, :, 2].sum() == inside.sum()
    assert low == min(0.0, p[inside, 2].min()) and high > p[inside, 2].max()
    bins = [[(out[r, b, 0], out[r, b, 1], int(out[r, b, 2])) for b in range(6)]
            for r in range(len(regions))]
    outs, filled = F.binner(p, regions, bins)
    assert filled == [int(out[r, :, 2].sum()) for r in range(len(regions))]
    # every box holds its points in input order
    for r, (ul, uh, vl, vh) in enumerate(regions):
        sel = (p[:, 0] >= ul) & (p[:, 0] < uh) & (p[:, 1] >= vl) & (p[:, 1] < vh)
        assert np.array_equal(outs[r][:filled[r]], p[sel])


def test_nan_w_inside_is_refused():
    p = kat_points()
    p[0, 2] = math.nan
    with pytest.raises(ValueError):
        F.bin_sizes(p, QUAD, 2)
