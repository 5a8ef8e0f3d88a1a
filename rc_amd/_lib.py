"""ctypes binding of rc_amd/librcgrid.so (C ABI declared in include/rcgrid.h).

The library is the product: nothing here falls back to a CPU or PyTorch
implementation.  If the shared library is missing or fails to load, every
call raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import sys
import threading
import weakref
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librcgrid.so"

ull = C.c_ulonglong
c_int_p = C.POINTER(C.c_int)
vp = C.c_void_p


class RcgError(RuntimeError):
    """A librcgrid call reported an error (rcg_last_error)."""


class Geom(C.Structure):
    """struct rcg_geom"""
    _fields_ = [
        ("uvw", vp),
        ("chunk_supps", vp),
        ("n_chunks", C.c_int32),
        ("ts_ch", C.c_int32),
        ("scale", C.c_double),
        ("wstep", C.c_double),
        ("grid_size", C.c_int32),
        ("grid_pitch", C.c_int32),
    ]


class Gcf(C.Structure):
    """struct rcg_gcf"""
    _fields_ = [
        ("data", vp),
        ("layer_off", vp),
        ("layer_len", vp),
        ("n_slots", C.c_int32),
        ("center", C.c_int32),
        ("supps", vp),
        ("max_w_plane", C.c_int32),
        ("half", C.c_int32),
        ("n_taps", C.c_int64),
    ]


_lock = threading.Lock()
_lib: C.CDLL | None = None

# name -> (restype, argtypes)
_LEGACY_GRID_ARGS = [C.c_double, C.c_double, C.c_int, vp, vp, vp, vp, vp,
                     C.c_int, C.c_int, C.c_int, vp]
SIGNATURES: dict[str, tuple] = {
    "rcg_legacy_path": (C.c_int, []),
    "gridKernelCPUFullGCF": (ull, _LEGACY_GRID_ARGS),
    "gridKernelCPUHalfGCF": (ull, _LEGACY_GRID_ARGS),
    "deGridKernelCPUFullGCF": (ull, _LEGACY_GRID_ARGS),
    "deGridKernelCPUHalfGCF": (ull, _LEGACY_GRID_ARGS),
    "rcg_last_error": (C.c_int, []),
    "rcg_last_error_string": (C.c_char_p, []),
    "rcg_pregrid": (C.c_int, [C.POINTER(Geom), C.POINTER(Gcf), vp, vp, vp, vp]),
    "rcg_grid_plain_f64": (C.c_int, [C.POINTER(Geom), C.POINTER(Gcf), vp, vp, vp, vp]),
    "rcg_degrid_plain_f64": (C.c_int, [C.POINTER(Geom), C.POINTER(Gcf), vp, vp, vp, vp]),
    "rcg_tiled_plan_create": (C.c_int, [C.POINTER(Geom), C.c_int, C.c_int, C.c_int,
                                        C.POINTER(vp)]),
    "rcg_tiled_plan_create_batch": (C.c_int, [C.POINTER(Geom), C.c_int, C.c_int, C.c_int,
                                              C.c_int, C.POINTER(vp)]),
    "rcg_tiled_plan_create_ex": (C.c_int, [C.POINTER(Geom), C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.POINTER(vp)]),
    "rcg_tiled_plan_destroy": (C.c_int, [vp]),
    "rcg_tiled_bin": (C.c_int, [vp, C.POINTER(Gcf), vp]),
    "rcg_tiled_timing": (C.c_int, [vp, C.c_int]),
    "rcg_tiled_compensated": (C.c_int, [vp, C.c_int]),
    "rcg_tiled_kernel_times": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int,
                                         C.POINTER(C.c_int)]),
    "rcg_tiled_graph_create": (C.c_int, [vp, C.POINTER(Gcf), vp, vp, vp, C.c_int,
                                         C.POINTER(vp)]),
    "rcg_graph_launch": (C.c_int, [vp, vp]),
    "rcg_graph_destroy": (C.c_int, [vp]),
    "rcg_tiled_set_degrid_waves": (C.c_int, [vp, C.c_int]),
    "rcg_tiled_set_grid_waves": (C.c_int, [vp, C.c_int]),
    "rcg_grid_tiled_f32": (C.c_int, [vp, C.POINTER(Gcf), vp, vp, vp]),
    "rcg_grid_tiled_adjoint_f32": (C.c_int, [vp, C.POINTER(Gcf), vp, vp, vp]),
    "rcg_degrid_tiled_f32": (C.c_int, [vp, C.POINTER(Gcf), vp, vp, vp]),
    "rcg_grid_tiled_f64": (C.c_int, [vp, C.POINTER(Gcf), vp, vp, vp]),
    "rcg_degrid_tiled_f64": (C.c_int, [vp, C.POINTER(Gcf), vp, vp, vp]),
    "rcg_tiled_stats": (C.c_int, [vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_longlong),
                                  C.POINTER(C.c_longlong), C.POINTER(C.c_int), vp]),
    # small kernels (device)
    "rcg_normalize": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "rcg_rotate": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_double, vp]),
    "rcg_grid0": (C.c_int, [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int, vp]),
    "rcg_reweight": (C.c_int, [vp, vp, C.c_double, C.c_int, C.c_int, vp, vp]),
    "rcg_reweight_static": (C.c_int, [vp, vp, C.c_double, C.c_int, C.c_int, vp]),
    "rcg_herm_padded": (C.c_int, [vp, C.c_int, C.c_int, vp]),
    "rcg_herm": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "rcg_subtract_vis": (C.c_int, [vp, vp, C.c_int, C.c_longlong, vp]),
    "rcg_const_vis": (C.c_int, [vp, C.c_int, C.c_longlong, C.c_double, C.c_double, vp]),
    "rcg_unpack_vis5": (C.c_int, [vp, C.c_longlong, C.c_int, C.c_double, vp, vp, C.c_int, vp]),
    "rcg_pack_vis5": (C.c_int, [vp, vp, C.c_int, C.c_longlong, vp, vp]),
    "rcg_fftshift_even": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "rcg_fft_plan_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "rcg_fft_execute": (C.c_int, [vp, vp, vp]),
    "rcg_fft_plan_destroy": (C.c_int, [vp]),
    "rcg_shutdown": (C.c_int, []),
    "rcg_hogbom_scratch_bytes": (C.c_size_t, []),
    "rcg_hogbom": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_double,
                             C.c_double, vp, vp]),
    "rcg_hogbom_iterations": (C.c_int, [vp, C.POINTER(C.c_uint), vp]),
    # reference host entry points next to the gridder
    "normalizeCPU": (None, [vp, C.c_int, C.c_int]),
    "rotateCPU": (None, [vp, vp, C.c_int, C.c_double]),
    "grid0": (None, [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int]),
    "reweight": (None, [vp, vp, C.c_double, C.c_int, C.c_int]),
    "herm_padded_inplace": (None, [vp, C.c_int, C.c_int]),
    "fft_inplace_even": (vp, [vp, C.c_int, vp, C.c_int, C.c_int]),
    "fftInitThreading": (None, []),
    "rcg_fft_destroy_plan": (None, [vp]),
    "deconvolve": (None, [vp, vp, vp, C.c_int, C.c_int, C.c_uint, C.c_double, C.c_double]),
    "mkGCFLayer": (vp, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
    "rcg_gcf_pad_quirk": (None, [C.c_int]),
    "calcAccums": (None, [vp, vp, vp, C.c_double, C.c_int, C.c_int]),
    "bin": (vp, [C.c_int, C.c_double, C.c_double, vp, vp, vp, C.c_int]),
    "binm": (vp, [C.c_int, C.c_double, C.c_double, vp, vp, vp, C.c_int]),
    "rcg_bin_free": (None, [vp]),
    # MS6 Halide buffer_t entry points (buffer_t passed by pointer)
    "kern_scatter": (C.c_int, [C.c_double, C.c_int32, C.c_int32, vp, vp, vp]),
    "kern_degrid": (C.c_int, [C.c_double, C.c_int32, C.c_int32, vp, vp, vp, vp]),
    "kern_scatter_gpu1": (C.c_int, [C.c_double, C.c_int32, C.c_int32, vp, vp, vp]),
    "kern_degrid_gpu1": (C.c_int, [C.c_double, C.c_int32, C.c_int32, vp, vp, vp, vp]),
    "nvGridder": (C.c_int, [C.c_double, C.c_int32, vp, vp, vp]),
    "bin_sort3": (None, [vp, C.c_int]),
    "bin_sort5": (None, [vp, C.c_int]),
    "gridGPUc": (None, [C.c_double, vp, vp, C.c_size_t, C.c_size_t, vp]),
    "gridGPUs": (None, [C.c_double, vp, vp, vp, C.c_size_t, C.c_size_t, vp]),
    "kern_init": (C.c_int, [vp]),
    "kern_image_init": (C.c_int, [vp]),
    "kern_psf_vis": (C.c_int, [vp, vp]),
    "kern_rotate": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_int32, vp, vp]),
    "kern_detile": (C.c_int, [vp, vp]),
    "kern_defacet": (C.c_int, [vp, vp]),
    "kern_image_sum": (C.c_int, [vp, vp]),
    "rcg_flow_bin_sizes": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, vp]),
    "rcg_flow_binner": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_int, vp, vp, vp, vp, vp]),
    # GCF generation (device)
    "rcg_gcf_gen_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
    "rcg_gcf_gen_layer": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                    vp]),
    "rcg_gcf_gen_destroy": (C.c_int, [vp]),
    "rcg_calc_accums": (C.c_int, [vp, C.c_longlong, C.c_double, C.c_int, vp, vp, vp, vp]),
}


def lib() -> C.CDLL:
    """Load librcgrid.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RcgError(
                    f"{LIB_PATH} is not built; run `python -m rc_amd.build` "
                    "(the HIP library has no CPU fallback)")
            h = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(h, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            _lib = h
            atexit.register(shutdown)
    return _lib


class Handle:
    """Base of the Python objects that own a librcgrid handle (plans, graphs,
    FFT plans, GCF generators).  Subclasses name their destroy function in
    DESTROY and keep the handle in self._h.  Every live owner is registered
    so that shutdown() (an atexit hook) releases its device memory while the
    HIP runtime is still up; after that, and while the interpreter is
    finalising, close() and __del__ make no library call."""

    DESTROY = ""
    _h = None

    def _own(self, h: C.c_void_p) -> None:
        self._h = h
        _owners.add(self)

    def _release(self) -> None:
        """Extra clean-up of a subclass after its handle is destroyed."""

    def close(self) -> None:
        h, self._h = self._h, C.c_void_p()
        if h and not _shut:
            getattr(lib(), self.DESTROY)(h)
        self._release()

    def __del__(self):
        if _shut or sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


_owners: "weakref.WeakSet[Handle]" = weakref.WeakSet()
_shut = False


def shutdown() -> None:
    """Destroy every live handle, then release the library's process-wide
    state (rcg_shutdown: rocFFT's rocfft_cleanup).  Registered with atexit
    when the library is loaded; later calls do nothing."""
    global _shut
    if _shut:
        return
    for o in list(_owners):
        try:
            o.close()
        except Exception:
            pass
    _shut = True
    if _lib is not None:
        _lib.rcg_shutdown()


def last_error() -> tuple[int, str]:
    h = lib()
    msg = h.rcg_last_error_string()
    return int(h.rcg_last_error()), (msg.decode() if msg else "")


def check(status: int, what: str) -> None:
    if status != 0:
        code, msg = last_error()
        raise RcgError(f"{what} failed with status {status}: {msg}")
