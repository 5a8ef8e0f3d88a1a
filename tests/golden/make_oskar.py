"""Writes the OSKAR visibility fixtures under tests/golden/: small .vis files
written with the reference's vendored oskar_binary library and the arrays
its reader (MS4/dep/oskar/OskarBinReader.cpp: mkFromFile + readAndReshuffle)
returns for them, both from oracle/_ref/liboskarref.so (compiled from the
reference sources by oracle/Makefile).

    python tests/golden/make_oskar.py

Files:
  oskar_blocks.vis    5 stations (10 baselines), 3 channels, 7 times in
                      blocks of 3 (the last block partial), double complex
                      matrix amplitudes, an extended tag and other header
                      tags the reader skips.
  oskar_oneblock.vis  4 stations, 1 channel, 4 times in one block sized for 8.
  oskar_badcrc.vis    oskar_blocks.vis with one byte of block 1's BASELINE_UU
                      payload changed (the reader reports a CRC failure).
  oskar_badhdr.vis    oskar_oneblock.vis with one bit of the NUM_STATIONS header
                      value flipped (mkFromFile reports a CRC failure).
  oskar_single.vis    single-precision amplitudes (the reader refuses them).
  oskar_expected.npz  per file: mkFromFile status and header fields; for the
                      readable ones readAndReshuffle's status, amps, uvws,
                      Metrix and per-baseline w range.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
LIB = ROOT / "oracle" / "_ref" / "liboskarref.so"

# OSKAR binary codes (MS4/dep/oskar/oskar_binary/oskar_binary.h,
# oskar_binary_data_types.h) and the visibility tags of OskarBinReader.cpp
CHAR, INT, SINGLE, DOUBLE, COMPLEX, MATRIX = 0x01, 0x02, 0x04, 0x08, 0x20, 0x40
GROUP_HEADER, GROUP_BLOCK = 11, 12
H = dict(TELESCOPE_PATH=1, NUM_TAGS_PER_BLOCK=2, WRITE_AUTOCORRELATIONS=3, AMP_TYPE=4,
         MAX_TIMES_PER_BLOCK=5, NUM_TIMES_TOTAL=6, NUM_CHANNELS=7, NUM_STATIONS=8,
         FREQ_START_HZ=9, FREQ_INC_HZ=10, CHANNEL_BANDWIDTH_HZ=11, TIME_START_MJD_UTC=12,
         TIME_INC_SEC=13, TIME_AVERAGE_SEC=14, PHASE_CENTRE=15, TELESCOPE_CENTRE=16,
         STATION_X_OFFSET_ECEF=17)
B = dict(DIM_START_AND_SIZE=1, FREQ_REF_INC_HZ=2, TIME_REF_INC_MJD_UTC=3,
         CROSS_CORRELATIONS=5, BASELINE_UU=6, BASELINE_VV=7, BASELINE_WW=8)


class VisData(C.Structure):
    """OskarBinReader.h:36-53."""
    _fields_ = [("h", C.c_void_p), ("num_baselines", C.c_int), ("num_channels", C.c_int),
                ("num_stations", C.c_int), ("num_times", C.c_int),
                ("num_times_baselines", C.c_int), ("num_points", C.c_int),
                ("phase_centre", C.c_double * 2), ("telescope_centre", C.c_double * 3),
                ("freq_start_inc", C.c_double * 2), ("time_start_inc", C.c_double * 2),
                ("channel_bandwidth_hz", C.c_double), ("time_average_sec", C.c_double)]


def lib():
    h = C.CDLL(str(LIB))
    vp = C.c_void_p
    h.oskar_binary_create.restype = vp
    h.oskar_binary_create.argtypes = [C.c_char_p, C.c_char, C.POINTER(C.c_int)]
    h.oskar_binary_write.argtypes = [vp, C.c_ubyte, C.c_ubyte, C.c_ubyte, C.c_int, C.c_size_t,
                                     vp, C.POINTER(C.c_int)]
    h.oskar_binary_write_ext.argtypes = [vp, C.c_ubyte, C.c_char_p, C.c_char_p, C.c_int,
                                         C.c_size_t, vp, C.POINTER(C.c_int)]
    h.oskar_binary_free.argtypes = [vp]
    h.mkFromFile.argtypes = [C.POINTER(VisData), C.c_char_p]
    h.freeBinHandler.argtypes = [C.POINTER(VisData)]
    h.readAndReshuffle.argtypes = [C.POINTER(VisData), vp, vp, vp, vp]
    return h


def write_vis(path: Path, n_st, n_ch, n_t, per_block, seed, amp_type=DOUBLE | COMPLEX | MATRIX,
              extended=False):
    """A visibility file in the OSKAR 2.6 layout the reader expects: header
    tags (group 11), then per block its 7 tags (group 12, index = block)."""
    h = lib()
    st = C.c_int(0)
    f = h.oskar_binary_create(str(path).encode(), b"w", C.byref(st))
    assert st.value == 0 and f
    rng = np.random.default_rng(seed)
    nb = n_st * (n_st - 1) // 2

    def w(group, dt, tag, idx, arr):
        a = np.ascontiguousarray(arr)
        h.oskar_binary_write(f, dt, group, tag, idx, a.nbytes, a.ctypes.data, C.byref(st))
        assert st.value == 0

    def wi(tag, *v):
        w(GROUP_HEADER, INT, H[tag], 0, np.asarray(v, np.int32))

    def wd(tag, *v):
        w(GROUP_HEADER, DOUBLE, H[tag], 0, np.asarray(v, np.float64))

    w(GROUP_HEADER, CHAR, H["TELESCOPE_PATH"], 0, np.frombuffer(b"telescope/model\0", np.uint8))
    wi("NUM_TAGS_PER_BLOCK", 7)
    wi("WRITE_AUTOCORRELATIONS", 0)
    wi("AMP_TYPE", amp_type)
    wi("MAX_TIMES_PER_BLOCK", per_block)
    wi("NUM_TIMES_TOTAL", n_t)
    wi("NUM_CHANNELS", n_ch)
    wi("NUM_STATIONS", n_st)
    f0, df = 1.4e8 + 0.125, 2.5e6 / 3
    wd("FREQ_START_HZ", f0)
    wd("FREQ_INC_HZ", df)
    wd("CHANNEL_BANDWIDTH_HZ", 1e5)
    wd("TIME_START_MJD_UTC", 57000.25)
    wd("TIME_INC_SEC", 10.0)
    wd("TIME_AVERAGE_SEC", 9.5)
    wd("PHASE_CENTRE", 20.0, -30.0)
    wd("TELESCOPE_CENTRE", 116.7, -26.8, 300.0)
    w(GROUP_HEADER, DOUBLE, H["STATION_X_OFFSET_ECEF"], 0, rng.uniform(-1e3, 1e3, n_st))
    if extended:
        a = np.arange(5, dtype=np.int32)
        h.oskar_binary_write_ext(f, INT, b"extra_group", b"extra_tag", 0, a.nbytes,
                                 a.ctypes.data, C.byref(st))
        assert st.value == 0
    n_blocks = (n_t + per_block - 1) // per_block
    single = (amp_type & 0x0F) == SINGLE
    for blk in range(n_blocks):
        t0 = blk * per_block
        nt = min(per_block, n_t - t0)
        w(GROUP_BLOCK, INT, B["DIM_START_AND_SIZE"], blk,
          np.asarray([t0, 0, nt, n_ch, nb, n_st], np.int32))
        w(GROUP_BLOCK, DOUBLE, B["FREQ_REF_INC_HZ"], blk, np.asarray([f0, df]))
        w(GROUP_BLOCK, DOUBLE, B["TIME_REF_INC_MJD_UTC"], blk,
          np.asarray([57000.25 + t0 * 10.0 / 86400, 10.0 / 86400]))
        amps = rng.standard_normal((nt, n_ch, nb, 8))
        w(GROUP_BLOCK, amp_type, B["CROSS_CORRELATIONS"], blk,
          amps.astype(np.float32) if single else amps)
        for tag in ("BASELINE_UU", "BASELINE_VV", "BASELINE_WW"):
            w(GROUP_BLOCK, DOUBLE, B[tag], blk, rng.uniform(-3000.0, 3000.0, (nt, nb)))
    h.oskar_binary_free(f)


def read_ref(path: Path) -> dict:
    h = lib()
    vd = VisData()
    s = h.mkFromFile(C.byref(vd), str(path).encode())
    out = {"mk_status": s}
    if s != 0:
        return out
    out.update(num_baselines=vd.num_baselines, num_channels=vd.num_channels,
               num_stations=vd.num_stations, num_times=vd.num_times,
               num_points=vd.num_points,
               doubles=np.asarray(list(vd.phase_centre) + list(vd.telescope_centre)
                                  + list(vd.freq_start_inc) + list(vd.time_start_inc)
                                  + [vd.channel_bandwidth_hz, vd.time_average_sec]))
    amps = np.zeros(vd.num_points * 8)
    uvws = np.zeros(vd.num_points * 3)
    met = np.zeros(6)
    blw = np.zeros(vd.num_baselines * 2)
    out["read_status"] = h.readAndReshuffle(C.byref(vd), amps.ctypes.data, uvws.ctypes.data,
                                            met.ctypes.data, blw.ctypes.data)
    out.update(amps=amps, uvws=uvws, metrix=met, bl_ws=blw)
    h.freeBinHandler(C.byref(vd))
    return out


def main() -> int:
    if not LIB.exists():
        print(f"{LIB} missing: make -C oracle ref")
        return 2
    write_vis(HERE / "oskar_blocks.vis", 5, 3, 7, 3, seed=1, extended=True)
    write_vis(HERE / "oskar_oneblock.vis", 4, 1, 4, 8, seed=2)
    write_vis(HERE / "oskar_single.vis", 4, 1, 2, 2, seed=3, amp_type=SINGLE | COMPLEX | MATRIX)
    raw = bytearray((HERE / "oskar_blocks.vis").read_bytes())
    # one byte inside block 1's BASELINE_UU payload: find its tag (group 12,
    # tag 6, index 1) by scanning the 20-byte tag headers
    pos, hit = 64, None
    while pos + 20 <= len(raw):
        flags, grp, tag = raw[pos + 4], raw[pos + 6], raw[pos + 7]
        idx = int.from_bytes(raw[pos + 8:pos + 12], "little")
        size = int.from_bytes(raw[pos + 12:pos + 20], "little")
        if not flags & 0x80 and grp == GROUP_BLOCK and tag == B["BASELINE_UU"] and idx == 1:
            hit = pos + 20 + 16
            break
        pos += 20 + size
    assert hit is not None
    raw[hit] ^= 0x40
    (HERE / "oskar_badcrc.vis").write_bytes(bytes(raw))
    # one byte of the NUM_STATIONS header payload (group 11, tag 8)
    raw = bytearray((HERE / "oskar_oneblock.vis").read_bytes())
    pos, hit = 64, None
    while pos + 20 <= len(raw):
        flags, grp, tag = raw[pos + 4], raw[pos + 6], raw[pos + 7]
        size = int.from_bytes(raw[pos + 12:pos + 20], "little")
        if not flags & 0x80 and grp == GROUP_HEADER and tag == H["NUM_STATIONS"]:
            hit = pos + 20
            break
        pos += 20 + size
    assert hit is not None
    raw[hit] ^= 0x01
    (HERE / "oskar_badhdr.vis").write_bytes(bytes(raw))
    res = {}
    for name in ("oskar_blocks", "oskar_oneblock", "oskar_badcrc", "oskar_badhdr",
                 "oskar_single"):
        for k, v in read_ref(HERE / f"{name}.vis").items():
            res[f"{name}__{k}"] = np.asarray(v)
    np.savez_compressed(HERE / "oskar_expected.npz", **res)
    for k in sorted(res):
        if res[k].size < 4:
            print(k, res[k])
    return 0


if __name__ == "__main__":
    sys.exit(main())
