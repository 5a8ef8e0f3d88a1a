"""Builds rc_amd/librcgrid.so (HIP, gfx950) and the checker libraries.

    python -m rc_amd.build            # library only
    python -m rc_amd.build --oracle   # also oracle/ (and oracle/_ref when the
                                      # reference tree is present)

Every source under rc_amd/csrc/*.hip is compiled with hipcc for gfx950 into
one shared library with a plain C ABI (include/rcgrid.h).  Packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "librcgrid.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
LDFLAGS = ["-shared", "-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")
    # hipcc's host pass does not know the packed-fp32 feature; drop that noise
    noise = ("packed-fp32-ops", "warning generated", "warnings generated")
    for line in (res.stdout + res.stderr).splitlines():
        if line.strip() and not any(k in line for k in noise):
            sys.stderr.write(line + "\n")


def build_lib(jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIPFLAGS, "-c", str(s), "-o", str(o)])
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIPFLAGS, *[str(o) for o in objs], *LDFLAGS, "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "all"])


def main(argv: list[str]) -> int:
    build_lib()
    if "--oracle" in argv:
        build_oracle()
    print(f"built {LIB}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
