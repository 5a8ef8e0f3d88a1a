"""numpy restatement of the MS6 Halide gridding kernels (TEST
INFRASTRUCTURE ONLY: loaded by tests/, never by the product).

MS6/kernel/cpu/gridding/scatter.cpp:17-145 (kern_scatter_<S>) and
degrid.cpp:17-124 (kern_degrid_<S>).  Halide is absent here, so the
generated kernels cannot be built; this restatement is pinned by delta-tap
KATs, the bounds KAT, a scalar-loop brute force and the bilinear transpose
identity between the two (tests/test_oracle_halide.py).  One assumption is
not pinnable here: Halide's round() takes halves to even (numpy.rint), as
Halide documents it; the pixel and oversampling tests use values off the
halves except one KAT that states the assumption.

Arrays: vis [n, 5] (u, v, w, re, im); gcf [64, S, S, 2] ([over][y][x][re,im],
over = ou + 8 ov); uvg [Y, P, 2] ([y][x][re,im]) whose first cell is grid
position origin = (x, y), X = width columns in use.
"""
from __future__ import annotations

import numpy as np

OVER = 8


def pixel(u: float, v: float, scale: float, grid_size: int, S: int):
    """(x0, y0, over): scatter.cpp:70-72."""
    us, vs = u * scale, v * scale
    ou = int(np.clip(np.rint(OVER * (us - np.floor(us))), 0, OVER - 1))
    ov = int(np.clip(np.rint(OVER * (vs - np.floor(vs))), 0, OVER - 1))
    x0 = int(np.rint(us) + (grid_size // 2 - S // 2))
    y0 = int(np.rint(vs) + (grid_size // 2 - S // 2))
    return x0, y0, ou + OVER * ov


def in_bounds(x0, y0, S, width, height, origin, margin):
    """scatter.cpp:60-64, 75-77."""
    m = max(0, (margin - S) // 2)
    return (origin[0] + m <= x0 <= origin[0] + width - S - 1 - m and
            origin[1] + m <= y0 <= origin[1] + height - S - 1 - m)


def scatter(vis: np.ndarray, gcf: np.ndarray, uvg: np.ndarray, scale: float, grid_size: int,
            margin: int, width: int | None = None, origin=(0, 0)) -> None:
    """uvg += every in-bound point's vis * gcf[over] placed at (x0, y0)."""
    S = gcf.shape[1]
    Y, P = uvg.shape[:2]
    X = width or P
    g = gcf[..., 0] + 1j * gcf[..., 1]
    for u, v, _, re, im in vis:
        x0, y0, over = pixel(u, v, scale, grid_size, S)
        if not in_bounds(x0, y0, S, X, Y, origin, margin):
            continue
        c = (re + 1j * im) * g[over]
        ys, xs = y0 - origin[1], x0 - origin[0]
        uvg[ys:ys + S, xs:xs + S, 0] += c.real
        uvg[ys:ys + S, xs:xs + S, 1] += c.imag


def degrid(gcf: np.ndarray, uvg: np.ndarray, vis: np.ndarray, scale: float, grid_size: int,
           margin: int, width: int | None = None, origin=(0, 0)) -> np.ndarray:
    """vis_out = vis; vis_out re/im -= sum of the window of uvg times gcf."""
    S = gcf.shape[1]
    Y, P = uvg.shape[:2]
    X = width or P
    g = gcf[..., 0] + 1j * gcf[..., 1]
    grid = uvg[..., 0] + 1j * uvg[..., 1]
    out = vis.copy()
    for k, (u, v, _, re, im) in enumerate(vis):
        x0, y0, over = pixel(u, v, scale, grid_size, S)
        if not in_bounds(x0, y0, S, X, Y, origin, margin):
            continue
        ys, xs = y0 - origin[1], x0 - origin[0]
        s = np.sum(grid[ys:ys + S, xs:xs + S] * g[over])
        out[k, 3] -= s.real
        out[k, 4] -= s.imag
    return out
