"""OSKAR visibility headers through librcgrid's reader (host part, no GPU):
mkFromFile against what the reference's reader (MS4/dep/oskar/
OskarBinReader.cpp, compiled from its sources with its vendored oskar_binary
library) returned for the same files (tests/golden/oskar_expected.npz,
written by tests/golden/make_oskar.py): header fields equal, the CRC-32C
failure of a changed header byte, the refusal of single-precision
amplitudes, a missing file, truncated and corrupted files (a status, never a
crash); and the CRC-32C the reader applies, checked on the standard test
string."""
import ctypes as C

import numpy as np
import pytest

from conftest import GOLDEN
from rc_amd import oskar

EXP = GOLDEN / "oskar_expected.npz"
OPEN_FAIL, CRC_FAIL = -100, -118


@pytest.mark.parametrize("name", ["oskar_blocks", "oskar_oneblock"])
def test_header_matches_reference(name):
    z = np.load(EXP)
    assert int(z[f"{name}__mk_status"]) == 0
    vd, hdr = oskar.open_header(GOLDEN / f"{name}.vis")
    try:
        for k in ("num_baselines", "num_channels", "num_stations", "num_times", "num_points"):
            assert hdr[k] == int(z[f"{name}__{k}"]), k
        doubles = (hdr["phase_centre"] + hdr["telescope_centre"] + hdr["freq_start_inc"]
                   + hdr["time_start_inc"] + [hdr["channel_bandwidth_hz"],
                                              hdr["time_average_sec"]])
        assert np.array_equal(np.asarray(doubles), z[f"{name}__doubles"])
        assert hdr["num_times_baselines"] == hdr["num_times"] * hdr["num_baselines"]
    finally:
        oskar._fns().freeBinHandler(C.byref(vd))


@pytest.mark.parametrize("name", ["oskar_badhdr", "oskar_single"])
def test_header_errors_match_reference(name):
    z = np.load(EXP)
    want = int(z[f"{name}__mk_status"])
    assert want in (CRC_FAIL, int(np.int32(np.uint32(0xdeadbeef))))
    with pytest.raises(oskar.OskarError) as e:
        oskar.open_header(GOLDEN / f"{name}.vis")
    assert e.value.status == want


def test_missing_file():
    with pytest.raises(oskar.OskarError) as e:
        oskar.open_header(GOLDEN / "no_such_file.vis")
    assert e.value.status == OPEN_FAIL


def crc32c(data: bytes) -> int:
    """Bitwise CRC-32C (reflected 0x82F63B78, init and final xor all ones)."""
    c = 0xFFFFFFFF
    for b in data:
        c ^= b
        for _ in range(8):
            c = (c >> 1) ^ (0x82F63B78 if c & 1 else 0)
    return c ^ 0xFFFFFFFF


def test_crc32c_of_the_fixture_chunks():
    """The standard check value, then every chunk of a fixture: the stored
    CRC equals CRC-32C(tag, names, payload) computed here bit by bit."""
    assert crc32c(b"123456789") == 0xE3069283
    raw = (GOLDEN / "oskar_blocks.vis").read_bytes()
    assert raw[:9] == b"OSKARBIN\0"
    pos, n, ext = 64, 0, 0
    while pos + 20 <= len(raw):
        flags = raw[pos + 4]
        size = int.from_bytes(raw[pos + 12:pos + 20], "little")
        assert flags & 0x40  # CRC present
        body = raw[pos:pos + 20 + size - 4]
        stored = int.from_bytes(raw[pos + 20 + size - 4:pos + 20 + size], "little")
        assert crc32c(body) == stored
        ext += bool(flags & 0x80)
        pos += 20 + size
        n += 1
    assert pos == len(raw) and n > 20 and ext == 1


def test_truncated_and_corrupted_files(tmp_path):
    """Every prefix of a fixture shorter than its first chunk, prefixes cut
    through the later chunks, and copies with random bytes changed open with
    an OSKAR status or succeed; none crashes or reports success for a file
    cut inside the 64-byte header."""
    rng = np.random.default_rng(7)
    data = (GOLDEN / "oskar_blocks.vis").read_bytes()
    p = tmp_path / "f.vis"
    cases = [data[:k] for k in range(0, 400)] + [data[:k] for k in range(400, len(data), 211)]
    for _ in range(300):
        b = bytearray(data)
        for i in rng.integers(0, len(b), rng.integers(1, 7)):
            b[i] = int(rng.integers(0, 256))
        cases.append(bytes(b))
    for k, c in enumerate(cases):
        p.write_bytes(c)
        try:
            vd, _ = oskar.open_header(p)
        except oskar.OskarError as e:
            assert e.status != 0
            continue
        oskar._fns().freeBinHandler(C.byref(vd))
        assert len(c) >= 64, k
