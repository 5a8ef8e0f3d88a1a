"""The C ABI: librcgrid.so loads and exports every function the headers in
include/ declare (no compute call; runs without a GPU)."""
import ctypes
import re
import subprocess
import threading

import pytest

from conftest import ROOT

HEADERS = sorted((ROOT / "include").glob("*.h"))
LIB = ROOT / "rc_amd" / "librcgrid.so"

MS4_NAMES = {"gridKernelCPUFullGCF", "gridKernelCPUHalfGCF", "deGridKernelCPUFullGCF",
             "deGridKernelCPUHalfGCF", "grid0", "normalizeCPU", "reweight", "rotateCPU",
             "herm_padded_inplace", "fft_inplace_even", "fftInitThreading", "deconvolve"}


def declared_functions():
    text = "\n".join(h.read_text() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    skip = {"if", "defined", "typedef"}
    return sorted({n for n in names if n not in skip})


def test_header_declares_the_reference_entry_points():
    assert MS4_NAMES <= set(declared_functions())
    assert {"kern_scatter", "kern_degrid", "kern_scatter_gpu1", "kern_degrid_gpu1",
            "mkGCFLayer", "calcAccums"} <= set(declared_functions())


def test_library_exports_every_declared_function():
    assert LIB.exists(), "build the library first: python -m rc_amd.build"
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_library_loads_without_a_gpu():
    ctypes.CDLL(str(LIB))
    from rc_amd import _lib
    h = _lib.lib()
    # the error state is per thread: a thread that made no call has none,
    # whatever earlier tests left on this one
    seen = []
    t = threading.Thread(target=lambda: seen.append(h.rcg_last_error()))
    t.start()
    t.join()
    assert seen == [0]
