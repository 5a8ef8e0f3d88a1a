"""The reference callers' calls replayed on librcgrid (GPU): the tables and
arguments of MS4/misc/cppcycle.cpp:190-223 and of the Haskell gridWrapper
(MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:64-85), built as those callers
build them (tests/callers.py), against goldens written by the reference's
own gridder with the same arguments (tap counts equal, relative L2 <= 1e-12,
the FP64 bar); the wrapper's table pointer 32 entries past the centre reads
every layer 32 entries further on, as the reference does; and a padded grid
whose (width, pitch) the wrapper passes as (grid_pitch, grid_size) is
refused (the reference would index rows of `width` cells with `pitch` as the
bound, past the caller's grid for windows at the last row)."""
import numpy as np
import pytest

import callers as K
from conftest import GOLDEN, rel_l2
from golden.make_golden import caller_cases
from rc_amd import _lib
from test_callers import check_offby32

pytestmark = pytest.mark.gpu

RCG_ERR_ARG = 1


@pytest.mark.parametrize("name", ["cpp", "hs"])
def test_caller_tables_vs_reference_golden(hip_device, name):
    z = np.load(GOLDEN / "callers.npz")
    call = dict(caller_cases())[name]
    assert np.array_equal(call.uvw, z[f"{name}_uvw"])
    h = _lib.lib()
    g, taps = call.grid(h.gridKernelCPUFullGCF)
    assert _lib.last_error()[0] == 0, _lib.last_error()
    assert taps == int(z[f"{name}_taps"])
    assert rel_l2(g, z[f"{name}_grid"]) <= 1e-12
    assert np.array_equal(g != 0, z[f"{name}_grid"] != 0)
    v, dtaps = call.degrid(h.deGridKernelCPUFullGCF, z[f"{name}_grid"])
    assert _lib.last_error()[0] == 0, _lib.last_error()
    assert dtaps == int(z[f"{name}_dtaps"])
    assert rel_l2(v, z[f"{name}_vis_out"]) <= 1e-12


def test_wrapper_offby32_on_gpu(hip_device):
    assert check_offby32(_lib.lib().gridKernelCPUFullGCF) >= 30


def test_wrapper_padded_grid_refused(hip_device):
    uvw, vis = K.baselines(23, 4, 8, 2.0, 1.4, lim=20.0)
    supp = K.support_rule(8, 8, 40)
    lay = K.make_layers(K.haskell_max_plane(uvw, 2.0), supp, np.random.default_rng(3))
    call = K.wrapper_call(uvw, vis, 2.0, supp, 96, 100, 1.3, lay)
    g, taps = call.grid(_lib.lib().gridKernelCPUFullGCF)
    code, msg = _lib.last_error()
    assert taps == 0 and code == RCG_ERR_ARG and "pitch" in msg
