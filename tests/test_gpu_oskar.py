"""OSKAR visibility files read and reshuffled by librcgrid (the reshuffle
and the extremes on the GPU) against the arrays the reference's reader
(MS4/dep/oskar/OskarBinReader.cpp:254-413, compiled from its sources)
returned for the same files (tests/golden/oskar_expected.npz): amplitudes,
uvw, Metrix and per-baseline w extremes bit for bit, over several blocks
with a partial last one and several channels; and the CRC failure of a
changed payload byte."""
import numpy as np
import pytest

from conftest import GOLDEN
from rc_amd import oskar

pytestmark = pytest.mark.gpu

EXP = GOLDEN / "oskar_expected.npz"


@pytest.mark.parametrize("name", ["oskar_blocks", "oskar_oneblock"])
def test_reshuffle_matches_reference(hip_device, name):
    z = np.load(EXP)
    assert int(z[f"{name}__read_status"]) == 0
    v = oskar.read(GOLDEN / f"{name}.vis")
    assert np.array_equal(v.amps.view(np.float64).reshape(-1), z[f"{name}__amps"])
    assert np.array_equal(v.uvw.reshape(-1), z[f"{name}__uvws"])
    met = np.asarray([v.metrix[k] for k in ("maxu", "maxv", "maxw", "minu", "minv", "minw")])
    assert np.array_equal(met, z[f"{name}__metrix"])
    assert np.array_equal(v.bl_w.reshape(-1), z[f"{name}__bl_ws"])


def test_payload_crc_failure(hip_device):
    z = np.load(EXP)
    assert int(z["oskar_badcrc__read_status"]) == -118
    with pytest.raises(oskar.OskarError) as e:
        oskar.read(GOLDEN / "oskar_badcrc.vis")
    assert e.value.status == -118
