"""bin / binm (the MS4 uv binner, MS4/kernel/cpu/binner/binner.cpp:100-189)
on the GPU against the oracle's restatement (orc_bin): per-bin counts, the
8-byte records bit for bit, the rotated amplitudes to FP64 rounding (the
device sincos against the host's sin and cos), the order inside every bin,
and the reference's block layout (binner.cpp:18-85)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import _lib, ms4

pytestmark = pytest.mark.gpu

DIVS = 32


def supports(M, lo=8, step=4, hi=64):
    return np.array([min(hi, lo + step * abs(p)) for p in range(-M, M + 1)], np.int32)


def points(rng, n, N, M, wstep, neg=0.0):
    """n points with cells in [-neg, N), a third of them within 40 cells of
    a bin edge, planes -M..M."""
    ds = N // DIVS
    u = rng.uniform(-neg, N - 1, n)
    v = rng.uniform(-neg, N - 1, n)
    k = rng.integers(1, DIVS, n)
    sel = rng.random(n) < 0.33
    u[sel] = np.clip(k[sel] * ds + rng.uniform(-40, 40, sel.sum()), -neg, N - 1)
    w = rng.uniform(-(M + 0.49), M + 0.49, n) * wstep
    return np.stack([u, v, w], 1)


def check_against_oracle(uvw, vis, scale, wstep, supps, M, N, mirror):
    got = ms4.bin_points(uvw, vis, scale, wstep, supps, M, N, mirror)
    exp = O.bin_points(uvw, vis, scale, wstep, supps, M, N, mirror)
    assert exp is not None
    counts, pre, out, _ = exp
    total = int(counts.sum())
    assert np.array_equal(got.counts, counts)
    assert np.array_equal(got.pre, pre)
    assert rel_l2(got.vis, out) <= 1e-15
    assert np.max(np.abs(got.vis - out)) <= 4e-16 * max(1.0, np.max(np.abs(out)))
    # block layout of binner.cpp:38-85: amplitudes of all bins, then records
    assert got.data_off == ms4.BIN_TABLE_BYTES
    assert got.data_size == ms4.BIN_TABLE_BYTES + 24 * total
    start = np.concatenate([[0], np.cumsum(counts.ravel())])[:-1].reshape(DIVS, DIVS)
    full = counts > 0
    assert np.array_equal(got.vis_off[full], 16 * start[full])
    assert np.array_equal(got.pre_off[full], 16 * total + 8 * start[full])
    for qu, qv in zip(*np.nonzero(full)):
        bv, bp = got.bin(qu, qv)
        s0, s1 = start[qu, qv], start[qu, qv] + counts[qu, qv]
        assert np.array_equal(bp, pre[s0:s1])
        assert np.array_equal(bv, got.vis[s0:s1])
    return total


@pytest.mark.parametrize("mirror", [False, True])
def test_bin_4096_matches_oracle(hip_device, mirror):
    rng = np.random.default_rng(61)
    N, M, wstep, n = 4096, 15, 3.0, 200_000   # 49 segments of 4096 points
    uvw = points(rng, n, N, M, wstep, neg=100.0)
    uvw[:, :2] *= 0.5
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    total = check_against_oracle(uvw, vis, 2.0, wstep, supports(M), M, N, mirror)
    assert total > n


@pytest.mark.parametrize("mirror", [False, True])
def test_bin_512_wide_margins(hip_device, mirror):
    """16-cell bins, supports up to 64: margins wider than half a bin, where
    the if / else-if order picks the side, and negative cells in bin 0."""
    rng = np.random.default_rng(62)
    N, M, wstep, n = 512, 3, 1.0, 9000
    uvw = points(rng, n, N, M, wstep, neg=15.0)
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    check_against_oracle(uvw, vis, 1.0, wstep, supports(M, 8, 16, 64), M, N, mirror)


@pytest.mark.parametrize("n", [1, 63, 4096, 4097, 3 * 4096 + 17])
def test_bin_segment_edges(hip_device, n):
    rng = np.random.default_rng(63 + n)
    N, M = 1024, 2
    uvw = points(rng, n, N, M, 1.0)
    # many points in few bins, so that bins span segments
    uvw[: n // 2, :2] = rng.uniform(30, 34, (n // 2, 2))
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    check_against_oracle(uvw, vis, 1.0, 1.0, supports(M), M, N, False)


def test_bin_large_matches_oracle(hip_device):
    rng = np.random.default_rng(64)
    N, M, wstep, n = 4096, 31, 10.0, 4_000_000
    uvw = points(rng, n, N, M, wstep)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex128)
    total = check_against_oracle(uvw, vis, 1.0, wstep, supports(M, 8, 2, 64), M, N, True)
    assert total >= n


def test_bin_empty_and_errors(hip_device):
    supps = supports(1)
    got = ms4.bin_points(np.zeros((0, 3)), np.zeros(0, np.complex128), 1.0, 1.0, supps, 1, 512)
    assert got.counts.sum() == 0 and got.data_size == ms4.BIN_TABLE_BYTES
    vis = np.ones(1, np.complex128)
    # cell -16 of 16-cell bins: bin -1, the reference's assert fails
    with pytest.raises(_lib.RcgError, match="bins"):
        ms4.bin_points(np.array([[-16.0, 3.0, 0.0]]), vis, 1.0, 1.0, supps, 1, 512)
    assert _lib.last_error()[0] == 4
    # cell 512 of a 512 grid: bin 32
    with pytest.raises(_lib.RcgError):
        ms4.bin_points(np.array([[3.0, 512.0, 0.0]]), vis, 1.0, 1.0, supps, 1, 512)
    # a grid below 32 cells has zero-width bins
    with pytest.raises(_lib.RcgError, match="grid_size"):
        ms4.bin_points(np.array([[3.0, 3.0, 0.0]]), vis, 1.0, 1.0, supps, 1, 16)
    assert _lib.last_error()[0] == 1
    # coordinates outside int16
    with pytest.raises(_lib.RcgError, match="int16"):
        ms4.bin_points(np.array([[40000.0, 3.0, 0.0]]), vis, 1.0, 1.0, supps, 1, 512)
    # a good call after the failures clears the error
    ms4.bin_points(np.array([[3.0, 3.0, 0.0]]), vis, 1.0, 1.0, supps, 1, 512)
    assert _lib.last_error()[0] == 0
