"""The reference callers' tables, replayed on the oracle (CPU): the calls of
MS4/misc/cppcycle.cpp:190-223 and of the Haskell gridWrapper
(MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:64-85), built as those callers
build them (tests/callers.py), against goldens written by the reference's
own compiled gridder with the same arguments; and what the wrapper's table
pointer, 32 entries past the centre, does: every layer is read 32 entries
(4 oversampled rows) further on, the last 4 rows of a plane's table from the
next plane's first layers."""
import numpy as np
import pytest

import callers as K
from conftest import GOLDEN
from oracle import oracle as O
from golden.make_golden import caller_cases


def orc(h, name):
    fn = getattr(h, name)
    return lambda *a: fn(0, *a, 1)


@pytest.mark.parametrize("name", ["cpp", "hs"])
def test_caller_tables_oracle_vs_reference_golden(name):
    z = np.load(GOLDEN / "callers.npz")
    call = dict(caller_cases())[name]
    # the generated inputs are the fixture's
    assert np.array_equal(call.uvw, z[f"{name}_uvw"])
    assert np.array_equal(call.vis, z[f"{name}_vis"])
    assert np.array_equal(call.chunk_supps, z[f"{name}_chunk_supps"])
    import hashlib
    assert hashlib.sha256(call.layers.data.tobytes()).hexdigest() == str(z[f"{name}_layers_sha256"])
    h = O.oracle_lib()
    g, taps = call.grid(orc(h, "orc_grid"))
    assert taps == int(z[f"{name}_taps"])
    assert np.array_equal(g, z[f"{name}_grid"])
    v, dtaps = call.degrid(orc(h, "orc_degrid"), z[f"{name}_grid"])
    assert dtaps == int(z[f"{name}_dtaps"])
    assert np.array_equal(v, z[f"{name}_vis_out"])


def test_caller_rules():
    """The two callers' chunk-support and plane-count rules on their edge
    cases: C round (halves away from zero) in cppcycle, Haskell round
    (halves to even) in the wrapper; a baseline crossing w = 0 gets plane 0;
    cppcycle keeps the sign of an all-negative baseline's plane."""
    assert [K.c_round(x) for x in (0.5, 1.5, 2.5, -0.5, -2.5)] == [1, 2, 3, -1, -3]
    assert [K.hs_round(x) for x in (0.5, 1.5, 2.5, -0.5, -2.5)] == [0, 2, 2, 0, -2]
    supp = K.support_rule(3, 16, 128)
    assert [supp(p) for p in (-9, -1, 0, 1, 8)] == [128, 19, 3, 19, 128]
    wstep = 1.0
    uvw = np.zeros((4, 2, 3))
    uvw[0, :, 2] = (-2.5, -1.6)    # all negative: plane round(-1.6) = -2
    uvw[1, :, 2] = (0.5, 2.5)      # all positive: plane round(0.5) = 1 / 0
    uvw[2, :, 2] = (-1.0, 1.0)     # crosses 0
    uvw[3, :, 2] = (2.5, 2.5)      # 3 (C) / 2 (Haskell)
    vis = np.ones((4, 2), np.complex128)
    lay = K.make_layers(3, supp, np.random.default_rng(0))
    cpp = K.cppcycle_call(uvw, vis, wstep, supp, 64, 2, lay, scale=1.0)
    assert cpp.chunk_supps.tolist() == [supp(2), supp(1), supp(0), supp(3)]
    hs = K.wrapper_call(uvw, vis, wstep, supp, 64, 64, 1.0, lay)
    assert hs.chunk_supps.tolist() == [supp(2), supp(0), supp(0), supp(2)]
    assert K.cppcycle_max_plane(uvw, wstep) == 3 and K.haskell_max_plane(uvw, wstep) == 2
    # table entry passed: 64 M (cppcycle) and tsiz / 2 = 64 M + 32 (wrapper)
    assert cpp.table == lay.table_at(64 * 3) and hs.table == lay.table_at(64 * 3 + 32)
    # argument order: (pitch, size) vs the wrapper's (width, pitch)
    assert (cpp.pitch_arg, cpp.size_arg) == (66, 64)
    assert (hs.pitch_arg, hs.size_arg) == (64, 64)


def encoded_points(seed, n, M, wstep, lim):
    rng = np.random.default_rng(seed)
    uvw = np.zeros((n, 1, 3))
    uvw[:, 0, 0] = rng.uniform(-lim, lim, n)
    uvw[:, 0, 1] = rng.uniform(-lim, lim, n)
    # planes -M..M, away from half planes
    uvw[:, 0, 2] = (rng.integers(-M, M + 1, n) + rng.uniform(-0.3, 0.3, n)) * wstep
    return uvw


def first_entry(call, fn, b):
    """Table entry of the layer point b was gridded with: the real part of
    its window's first cell, less 100000 (vis = 1, encoded taps); None when
    nothing was gridded."""
    sub = K.Call(call.scale, call.wstep, call.chunk_supps[b:b + 1], call.table, call.gcf_supps,
                 call.supps_entry, call.uvw[b:b + 1], np.ones((1, 1), np.complex128), 1,
                 call.pitch_arg, call.size_arg, call.rows, call.row_len, call.layers)
    g, taps = sub.grid(fn)
    nz = np.flatnonzero(g)
    if len(nz) == 0:
        return None
    return int(round(g.reshape(-1)[nz[0]].real)) - 100000


def offby32_cases():
    M, wstep = 2, 1.0
    supp = K.support_rule(6, 4, 40)
    lay = K.make_layers(M, supp, encoded=True)
    uvw = encoded_points(5, 40, M, wstep, 20.0)
    vis = np.ones((40, 1), np.complex128)
    cpp = K.cppcycle_call(uvw, vis, wstep, supp, 64, 0, lay, scale=1.0)
    hs = K.wrapper_call(uvw, vis, wstep, supp, 64, 64, 1.0, lay)
    return lay, cpp, hs


def check_offby32(fn):
    lay, cpp, hs = offby32_cases()
    n_shift = 0
    for b in range(cpp.uvw.shape[0]):
        assert cpp.chunk_supps[b] == hs.chunk_supps[b]
        e = first_entry(cpp, fn, b)
        assert e is not None and 0 <= e < lay.tsiz
        w = first_entry(hs, fn, b)
        if e + 32 < lay.tsiz:
            assert w == e + 32, (b, e, w)
            n_shift += 1
        else:  # past the table: the pad's zero taps
            assert w is None
    return n_shift


def test_wrapper_offby32_on_oracle():
    assert check_offby32(orc(O.oracle_lib(), "orc_grid")) >= 30


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_wrapper_offby32_on_reference():
    ref = O.ref_lib()
    assert check_offby32(ref.gridKernelCPUFullGCF) >= 30
