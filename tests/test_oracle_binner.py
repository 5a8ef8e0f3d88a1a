"""The oracle's MS4 binner (orc_bin, restating MS4/kernel/cpu/binner/
binner.cpp:100-189) against checks that do not restate its if / else-if
form: the pinned pregrid (orc_pregrid) for the records, a geometric rule for
the margin copies, hand-computed KATs for the remainder edge cases, and the
input order inside every bin.  binner.cpp needs the CUDA runtime and cannot be
built here, so these checks are its pinning."""
import ctypes as C

import numpy as np

from oracle import oracle as O

DIVS = 32


def pregrid_untranslated(uvw, scale, wstep, mirror):
    """orc_pregrid with no translation (max_supp 0, grid_size 0): the
    binner's pregridPoint under __BINNER (MS4/kernel/common/common.h:
    120-166)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    out = np.zeros((len(uvw), 6), np.int32)
    O.oracle_lib().orc_pregrid_many(scale, wstep, uvw.ctypes.data, len(uvw), 0, 0,
                                    int(mirror), out.ctypes.data)
    return out


def edge_heavy_points(rng, n, N, M, wstep):
    """Points of a 2*M + 1 plane table, half of them within 40 cells of a bin
    edge, all with non-negative cells."""
    ds = N // DIVS
    u = rng.uniform(0, N - 1, n)
    v = rng.uniform(0, N - 1, n)
    k = rng.integers(1, DIVS, n)
    near = rng.random(n) < 0.5
    u[near] = np.clip(k[near] * ds + rng.uniform(-40, 40, near.sum()), 0, N - 1)
    v[~near] = np.clip(k[~near] * ds + rng.uniform(-40, 40, (~near).sum()), 0, N - 1)
    w = rng.uniform(-(M + 0.49), M + 0.49, n) * wstep
    return np.stack([u, v, w], 1)


def copy_sets(counts, src):
    bins = np.repeat(np.arange(DIVS * DIVS), counts.ravel())
    return np.sort(src.astype(np.int64) * 1024 + bins)


def geometric_sets(pg, supps, M, N):
    """Bins whose cell range meets the window [p.u - m, p.u + m) x
    [p.v - m, p.v + m), m = supp / 2, for non-negative cells and
    m <= N / 64 (then at most one neighbour per axis is reached)."""
    ds = N // DIVS
    keys = []
    for i, (u, v, _, _, wp, _) in enumerate(pg):
        m = supps[wp + M] // 2
        us = {b for b in ((u - m) // ds, u // ds, (u + m - 1) // ds) if 0 <= b < DIVS}
        vs = {b for b in ((v - m) // ds, v // ds, (v + m - 1) // ds) if 0 <= b < DIVS}
        keys += [i * 1024 + a * DIVS + b for a in us for b in vs]
    return np.sort(np.array(keys, np.int64))


def test_records_copies_and_order():
    rng = np.random.default_rng(51)
    N, M, wstep, n = 4096, 7, 2.5, 4000
    supps = np.array([8 + 8 * abs(p) for p in range(-M, M + 1)], np.int32)  # 8..64
    uvw = edge_heavy_points(rng, n, N, M, wstep)
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = 1.0
    for mirror in (False, True):
        counts, pre, out, src = O.bin_points(uvw, vis, scale, wstep, supps, M, N, mirror)
        pg = pregrid_untranslated(uvw, scale, wstep, mirror)
        # margin copies: the geometric rule
        assert np.array_equal(copy_sets(counts, src), geometric_sets(pg, supps, M, N))
        # each bin holds its copies in input order
        starts = np.concatenate([[0], np.cumsum(counts.ravel())])
        for b in range(DIVS * DIVS):
            seg = src[starts[b]:starts[b + 1]]
            assert np.all(np.diff(seg) > 0)
        # records: the pinned pregrid, translated by N/2 - supp/2
        s = supps[pg[src, 4] + M]
        assert np.array_equal(pre[:, 0], pg[src, 0] + N // 2 - s // 2)
        assert np.array_equal(pre[:, 1], pg[src, 1] + N // 2 - s // 2)
        assert np.array_equal(pre[:, 2], pg[src, 5])
        assert np.array_equal(pre[:, 3], s)
        # amplitudes: rotateCPU (pinned by the compiled reference's vectors)
        rot = vis.copy()
        uvw_c = np.ascontiguousarray(uvw)
        O.oracle_lib().orc_rotate(uvw_c.ctypes.data, rot.ctypes.data, n, scale)
        assert np.array_equal(out, rot[src])
    # mirrored index is the negated |w| index for w < 0
    pg0 = pregrid_untranslated(uvw, scale, wstep, False)
    pg1 = pregrid_untranslated(uvw, scale, wstep, True)
    neg = pg0[:, 4] < 0
    assert np.array_equal(pg1[neg, 5], -((-pg0[neg, 4] * 8 + pg0[neg, 2]) * 8 + pg0[neg, 3]))


def kat_table():
    # planes -1, 0, 1: supports 8, 64, 8
    return np.array([8, 64, 8], np.int32)


def bins_of(res, i):
    counts, _, _, src = res
    bins = np.repeat(np.arange(DIVS * DIVS), counts.ravel())
    return sorted((int(b) // DIVS, int(b) % DIVS) for b in bins[src == i])


def test_remainder_edge_kats():
    """grid 512: 16 cells per bin.  Hand-computed from binner.cpp:135-196."""
    uvw = np.array([
        [5.25, 100.5, 0.0],    # supp 64, m 32: ru 5 < m at qu 0 -> else-if right;
                               # rv 4 < m at qv 6 -> left
        [-3.5, 40.0, -1.0],    # cell -4: bin 0 by truncation, rem -4; supp 8
        [12.0, 40.0, 1.0],     # ru 12 = 16 - 4: strict test, no right copy
        [13.0, 40.0, 1.0],     # ru 13 > 12: right copy
        [20.0, 40.0, 0.0],     # supp 64: ru 4 < 32 at qu 1 -> left
    ])
    vis = np.ones(len(uvw), np.complex128)
    for mirror in (False, True):
        res = O.bin_points(uvw, vis, 1.0, 1.0, kat_table(), 1, 512, mirror)
        assert res is not None
        assert bins_of(res, 0) == [(0, 5), (0, 6), (1, 5), (1, 6)]
        assert bins_of(res, 1) == [(0, 2)]
        assert bins_of(res, 2) == [(0, 2)]
        assert bins_of(res, 3) == [(0, 2), (1, 2)]
        # cell (20, 40), supp 64: ru 4 < 32 at qu 1 -> left only (the else-if
        # skips the right test); rv 8 < 32 at qv 2 -> left only
        assert bins_of(res, 4) == [(0, 1), (0, 2), (1, 1), (1, 2)]
        counts, pre, out, src = res
        r2 = pre[np.flatnonzero(src == 2)[0]]
        assert list(r2) == [12 + 256 - 4, 40 + 256 - 4, (1 * 8 + 0) * 8 + 0, 8]
        r1 = pre[np.flatnonzero(src == 1)[0]]
        # cell -4, over_u = floor(8 * 0.5) = 4, plane -1
        assert list(r1) == [-4 + 252, 40 + 252, -96 if mirror else -32, 8]
        # exp(2 pi i w scale) for w = -1, 0, 1 is 1 to rounding
        assert np.allclose(out, 1.0, atol=1e-15, rtol=0)


def test_assert_cases_fail():
    vis = np.ones(1, np.complex128)
    # cell -16 of a 16-cell bin: quotient -1, the reference's assert fails
    assert O.bin_points(np.array([[-16.0, 3.0, 0.0]]), vis, 1.0, 1.0, kat_table(), 1,
                        512) is None
    # cell 512: bin 32
    assert O.bin_points(np.array([[3.0, 512.0, 0.0]]), vis, 1.0, 1.0, kat_table(), 1,
                        512) is None
    # grid smaller than 32 cells: zero-width bins
    assert O.bin_points(np.array([[3.0, 3.0, 0.0]]), vis, 1.0, 1.0, kat_table(), 1,
                        16) is None


def test_empty_input():
    counts, pre, out, src = O.bin_points(np.zeros((0, 3)), np.zeros(0, np.complex128), 1.0,
                                         1.0, kat_table(), 1, 512)
    assert counts.sum() == 0 and len(pre) == 0
