"""OSKAR 2.6 visibility files through librcgrid's reader (mkFromFile,
readAndReshuffle, freeBinHandler of include/rcgrid_oskar.h, the drop-in for
MS4/dep/oskar/OskarBinReader.h): header fields, and the amplitudes and uvw
reshuffled on the GPU into the [baseline][time][channel] layout the
gridder's callers use (MS4/misc/cppcycle.cpp:60-77)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


class VisDataC(C.Structure):
    """VisData of include/rcgrid_oskar.h (OskarBinReader.h:36-53)."""
    _fields_ = [("h", C.c_void_p), ("num_baselines", C.c_int), ("num_channels", C.c_int),
                ("num_stations", C.c_int), ("num_times", C.c_int),
                ("num_times_baselines", C.c_int), ("num_points", C.c_int),
                ("phase_centre", C.c_double * 2), ("telescope_centre", C.c_double * 3),
                ("freq_start_inc", C.c_double * 2), ("time_start_inc", C.c_double * 2),
                ("channel_bandwidth_hz", C.c_double), ("time_average_sec", C.c_double)]


class OskarError(RuntimeError):
    def __init__(self, what: str, status: int):
        super().__init__(f"{what}: OSKAR status {status}")
        self.status = status


@dataclass
class OskarVis:
    header: dict
    amps: np.ndarray      # complex128 [baselines, times, channels, 4] (2x2 matrix a, b, c, d)
    uvw: np.ndarray       # float64 [baselines, times, channels, 3], wavelengths
    metrix: dict          # extreme u, v, w
    bl_w: np.ndarray      # float64 [baselines, 2]: max w, min w


def _fns():
    h = _lib.lib()
    h.mkFromFile.restype = C.c_int
    h.mkFromFile.argtypes = [C.POINTER(VisDataC), C.c_char_p]
    h.readAndReshuffle.restype = C.c_int
    h.readAndReshuffle.argtypes = [C.POINTER(VisDataC), C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
    h.freeBinHandler.restype = None
    h.freeBinHandler.argtypes = [C.POINTER(VisDataC)]
    return h


def open_header(path) -> tuple[VisDataC, dict]:
    """mkFromFile: the open handle and the header fields; raises OskarError."""
    h = _fns()
    vd = VisDataC()
    s = h.mkFromFile(C.byref(vd), str(path).encode())
    if s != 0:
        raise OskarError("mkFromFile", s)
    hdr = {k: getattr(vd, k) for k in ("num_baselines", "num_channels", "num_stations",
                                       "num_times", "num_times_baselines", "num_points")}
    for k in ("phase_centre", "telescope_centre", "freq_start_inc", "time_start_inc"):
        hdr[k] = list(getattr(vd, k))
    hdr["channel_bandwidth_hz"] = vd.channel_bandwidth_hz
    hdr["time_average_sec"] = vd.time_average_sec
    return vd, hdr


def read(path) -> OskarVis:
    """The whole file: header, then readAndReshuffle (GPU); raises OskarError."""
    h = _fns()
    vd, hdr = open_header(path)
    try:
        n, nb = vd.num_points, vd.num_baselines
        amps = np.empty(n * 8)
        uvws = np.empty(n * 3)
        met = np.empty(6)
        blw = np.empty(max(nb, 1) * 2)
        s = h.readAndReshuffle(C.byref(vd), amps.ctypes.data, uvws.ctypes.data, met.ctypes.data,
                               blw.ctypes.data)
        if s != 0:
            raise OskarError("readAndReshuffle", s)
    finally:
        h.freeBinHandler(C.byref(vd))
    shape = (nb, vd.num_times, vd.num_channels)
    return OskarVis(header=hdr, amps=amps.view(np.complex128).reshape(shape + (4,)),
                    uvw=uvws.reshape(shape + (3,)),
                    metrix=dict(zip(("maxu", "maxv", "maxw", "minu", "minv", "minw"), met)),
                    bl_w=blw[:2 * nb].reshape(nb, 2))
