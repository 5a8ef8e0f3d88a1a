"""rcg_reweight_static, the entry point that allocates nothing (its count
buffer is the library's own per-device array), against rcg_reweight with a
caller buffer and against the oracle: bit-exact, on two calls in a row (the
second on a smaller grid and fewer points, so that stale counts of the first
call would show), plus its refusal of grids above its buffer."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rc_amd import _lib

pytestmark = pytest.mark.gpu

RCG_ERR_ARG = 1
MAX_SIZE = 4096


def points(seed, n, N, scale):
    """n points whose rounded cells fall on a grid of N cells (a few past
    its edge), many cells shared by several points."""
    rng = np.random.default_rng(seed)
    uvw = np.empty((n, 3))
    half = N // 2 + 3
    uvw[:, 0] = rng.integers(-half, half, n) / scale + rng.uniform(-0.4, 0.4, n) / scale
    uvw[:, 1] = rng.integers(-half, half, n) / scale + rng.uniform(-0.4, 0.4, n) / scale
    uvw[:, 2] = rng.uniform(-100, 100, n)
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return uvw, vis


def run(fn_name, uvw, vis, scale, N, with_buffer):
    lib = _lib.lib()
    du = torch.from_numpy(uvw).cuda()
    dv = torch.from_numpy(vis.astype(np.complex128)).cuda()
    args = [C.c_void_p(du.data_ptr()), C.c_void_p(dv.data_ptr()), scale, len(vis), N]
    if with_buffer:
        cnt = torch.full((N * N,), 12345, dtype=torch.int32, device="cuda")
        args.append(C.c_void_p(cnt.data_ptr()))
    args.append(None)
    st = getattr(lib, fn_name)(*args)
    torch.cuda.synchronize()
    return st, dv.cpu().numpy()


@pytest.mark.parametrize("calls", [[(1, 300_000, 1024, 0.37), (2, 40_000, 256, 1.9)]])
def test_reweight_static_matches(hip_device, calls):
    for seed, n, N, scale in calls:
        uvw, vis = points(seed, n, N, scale)
        st0, ref = run("rcg_reweight", uvw, vis, scale, N, True)
        st1, got = run("rcg_reweight_static", uvw, vis, scale, N, False)
        assert st0 == 0 and st1 == 0, _lib.last_error() if hasattr(_lib, "last_error") else ""
        assert np.array_equal(got, ref), f"call seed={seed} N={N}"
        # the oracle's reweight on the same points
        h = O.oracle_lib()
        orc = vis.astype(np.complex128).copy()
        h.orc_reweight(np.ascontiguousarray(uvw).ctypes.data, orc.ctypes.data, scale, n, N)
        assert np.array_equal(got, orc), f"call seed={seed} N={N} vs oracle"
        assert not np.array_equal(got, vis)  # some cells hold several points


def test_reweight_static_refuses_large_grid(hip_device):
    uvw, vis = points(3, 1000, 64, 1.0)
    st, got = run("rcg_reweight_static", uvw, vis, 1.0, MAX_SIZE + 1, False)
    assert st == RCG_ERR_ARG
    assert np.array_equal(got, vis)
