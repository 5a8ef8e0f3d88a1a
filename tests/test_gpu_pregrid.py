"""The device pregrid records (rcg_pregrid) against the oracle's pregrid:
integer fields bit-exact."""
import numpy as np
import pytest

from _fixtures import load_probe
from oracle import oracle as O
from rc_amd import device as D
from rc_amd import synth
from rc_amd.tables import VisChunks

pytestmark = pytest.mark.gpu


def expected_records(vc, gcf, N, P):
    pg = O.pregrid(vc, N, gcf.half)
    S = np.repeat(vc.supps, vc.ts_ch)
    out = []
    for (u, v, ou, ov, wp, idx), s in zip(pg, S):
        inb = idx != np.iinfo(np.int32).min and not (u < 0 or u >= N - s or v < 0 or v >= N - s)
        if not inb:
            out.append((-1, 0, 0, s, 0))
            continue
        slot = abs(idx) if gcf.half else idx + gcf.center
        gs = int(gcf.supps[wp + gcf.max_w_plane])
        off = (gs - s) // 2 * (gs + 1) if gs >= s else -1
        ok = 0 <= slot < gcf.n_slots and off >= 0 and off + (s - 1) * (gs + 1) < gcf.layer_len[slot]
        if not ok:
            out.append((-1, 0, 0, s, 2))
            continue
        out.append((u * P + v, int(gcf.layer_off[slot]) + off, gs, s,
                    1 if (gcf.half and wp < 0) else 0))
    return np.asarray(out, np.int64)


def check(vc, gcf, N, P):
    geom = D.DeviceGeom.from_chunks(vc, N, P)
    dg = D.DeviceGCF.from_table(gcf)
    pg, taps, err = D.pregrid(geom, dg)
    rec = D.records_to_numpy(pg)
    got = np.stack([rec["cell"], rec["tap0"], rec["stride"], rec["supp"], rec["flags"]], 1)
    exp = expected_records(vc, gcf, N, P)
    # out-of-bound records carry no tap offset or stride
    oob = exp[:, 0] < 0
    got[oob, 1:3] = 0
    assert np.array_equal(got.astype(np.int64), exp)
    inb = exp[:, 0] >= 0
    assert int(taps.item()) == int((exp[inb, 3] ** 2).sum())
    assert bool(err.item()) == bool((exp[:, 4] == 2).any())


@pytest.mark.parametrize("seed,half,cs", [(0, False, "min"), (1, True, "min"),
                                          (2, False, "random"), (3, True, "random")])
def test_pregrid_random(hip_device, seed, half, cs):
    vc, gcf, N = synth.random_small(seed, chunks=40, ts_ch=25, half=half, chunk_supp=cs)
    check(vc, gcf, N, N + 5)


def test_pregrid_probe_points(hip_device):
    from golden.make_golden import encoded_table
    for half, scale, wstep, N, pts, S, res in load_probe():
        sup = {p: 6 + 2 * abs(p) for p in range(-3 if not half else 0, 4)}
        gcf = encoded_table(sup, half)
        vc = VisChunks(uvw=pts.reshape(-1, 1, 3), vis=np.ones((len(pts), 1)),
                       supps=S.astype(np.int32), scale=scale, wstep=wstep)
        check(vc, gcf, N, N)


def test_pregrid_wproj_sample(hip_device):
    """A slice of the w-projection workload (earth-rotation tracks)."""
    vc, gcf, N = synth.wproj(n_ant=40, n_times=64)
    check(vc, gcf, N, N)
