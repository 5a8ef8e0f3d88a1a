"""numpy restatement of the reference's GCF (w-projection kernel) generator
(TEST INFRASTRUCTURE ONLY: loaded by tests/, never by the product).

MS4/kernel/cpu/gcf/GCF.cpp:16-219 and its caller
MS4/dna-programs/Kernel/CPU/GCF.hs:36-89.  GCF.cpp needs FFTW (absent
here), so it cannot be compiled; this restatement is pinned by the closed
form of the w = 0 layer (a Dirichlet kernel), a direct DFT on a tiny arena,
and the normalisation invariant (see tests/test_oracle_gcf.py).

mk_gcf_layer   GCF.cpp:58-120 (__mkGCFLayer) + :16-54
               (__transpose_and_normalize_and_extract), OVER = 8
calc_accums    GCF.cpp:148-219
gcf_kernel     GCF.hs:36-89 (per-plane supports, mean w per plane, layers of
               planes -maxWPlane..maxWPlane in order)
"""
from __future__ import annotations

import math

import numpy as np

from .fft_oracle import fft_inplace_even

OVER = 8


def w_screen(max_support: int, t2: float, w: float) -> np.ndarray:
    """The max_support^2 block of exp(-2 pi i w (1 - sqrt(1 - x^2 - y^2)))
    at x = i * t2 / radius, i = -radius .. radius + max_support % 2 - 1
    (GCF.cpp:80-97)."""
    radius = max_support // 2
    normer = t2 / float(radius)
    i = np.arange(-radius, radius + max_support % 2, dtype=np.float64)
    x = i * normer
    ph = w * (1.0 - np.sqrt(1.0 - x[:, None] * x[:, None] - x[None, :] * x[None, :]))
    a = 2.0 * math.pi * ph
    return np.cos(a) - 1j * np.sin(a)


def mk_gcf_layer(support: int, max_support: int, src_pad: int, t2: float, w: float,
                 pad_quirk: bool = False) -> np.ndarray:
    """The 64 oversampled layers [over_u][over_v][support][support] of one
    w-plane.  pad_quirk=True reproduces the reference's normalisation sums,
    which walk the padded arena as if it were contiguous (GCF.cpp:28-32)."""
    L = max_support * OVER
    pitch = L + src_pad
    arena = np.zeros((L, pitch), np.complex128)
    radius = max_support // 2
    r0 = L // 2 - radius
    arena[r0:r0 + max_support, r0:r0 + max_support] = w_screen(max_support, t2, w)
    arena = fft_inplace_even(arena, 1, L)  # FFTW_BACKWARD, with fftshifts
    if pad_quirk:
        flat = arena.reshape(-1)[: L * L]
    else:
        flat = arena[:, :L].reshape(-1)
    sums = flat.real.reshape(max_support, OVER, max_support, OVER).sum(axis=(0, 2))
    start = (max_support // 2 - support // 2) * OVER
    sub = arena[start:start + support * OVER, start:start + support * OVER]
    lay = sub.reshape(support, OVER, support, OVER).transpose(1, 3, 0, 2)
    return lay / sums[:, :, None, None]


def c_round(q: np.ndarray) -> np.ndarray:
    """C round(): nearest integer, halves away from zero, exactly (no
    q + 0.5 rounding step)."""
    t = np.trunc(q)
    f = q - t  # exact
    return t + np.where(np.abs(f) >= 0.5, np.sign(q), 0.0)


def calc_accums(uvw: np.ndarray, wstep: float, n_planes: int) -> tuple[np.ndarray, np.ndarray]:
    """Per w-plane sum of w and point count, planes centred at n_planes // 2
    (plane = round(w / wstep), half away from zero)."""
    w = np.asarray(uvw, np.float64).reshape(-1, 3)[:, 2]
    plane = c_round(w / wstep).astype(np.int64) + n_planes // 2
    if plane.size and (plane.min() < 0 or plane.max() >= n_planes):
        raise ValueError("a point's w-plane is outside the table")
    sums = np.zeros(n_planes)
    np.add.at(sums, plane, w)
    npts = np.bincount(plane, minlength=n_planes).astype(np.int64)
    return sums, npts


def plane_supports(max_w_plane: int, min_size: int, growth: int, max_size: int) -> list[int]:
    """supp(i) = min(max_size, min_size + growth * |i|), i = -M..M (GCF.hs:80-82)."""
    return [min(max_size, min_size + growth * abs(i)) for i in range(-max_w_plane, max_w_plane + 1)]


def gcf_kernel(uvw: np.ndarray, wstep: float, theta: float, min_size: int, growth: int,
               max_size: int, max_w_plane: int, src_pad: int = 4,
               pad_quirk: bool = False) -> tuple[list[int], list[np.ndarray], np.ndarray]:
    """All planes -M..M: (supports, layers per plane [8, 8, S, S], mean w per
    plane).  The layers use w = mean(w of the plane's points) * theta, or
    the plane's nominal w (p * wstep) * theta when it has no points."""
    n_planes = 2 * max_w_plane + 1
    sums, npts = calc_accums(uvw, wstep, n_planes)
    supps = plane_supports(max_w_plane, min_size, growth, max_size)
    layers, means = [], np.zeros(n_planes)
    wmid = -wstep * max_w_plane  # advanced by wstep per plane, as GCF.hs:59
    for k in range(n_planes):
        avg = sums[k] / npts[k] if npts[k] > 0 else wmid
        means[k] = avg
        layers.append(mk_gcf_layer(supps[k], max_size, src_pad, theta / 2.0, avg * theta,
                                   pad_quirk))
        wmid = wmid + wstep
    return supps, layers, means
