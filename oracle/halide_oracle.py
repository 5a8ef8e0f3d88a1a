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

nvGridder (MS6/kernel/nvidia/gridder: halide_api.cu:4-29, grid_gpu.cu:37-51
and 83-155, binsort.cpp:6-32) is restated by nv_sort_order and nv_grid: a
16 x 16 GCF [over][b][a] with over = 8 sub_y + sub_x, floor-based pixels
without a shift to the grid centre, tap [7][7] on the point's pixel, the
grid overwritten, and the host visibility array sorted in place.  Pinned by
delta-tap KATs, hand-computed sort keys and a scalar-loop brute force
(tests/test_oracle_halide.py).  The reference's gather finds a point
through bookmarks of its sort order; the restatement is the sum that gather
computes where the reference is defined (non-negative pixel coordinates,
scale 1, grids up to 2048); outside that the reference indexes the GCF with
negative offsets and no parity is possible.
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


NV_GCF = 16


def _c_div(a: np.ndarray, b: int) -> np.ndarray:
    """C integer division (truncates toward zero)."""
    q = np.abs(a) // b
    return np.where(a < 0, -q, q)


def nv_sort_order(vis: np.ndarray) -> np.ndarray:
    """Stable order of binsort.cpp:8-31: by block key (gy * 2055) / 8 + gx
    (gx = floor(x) / 8, C integer arithmetic, IMG_SIZE 2048, unscaled
    coordinates), then by 8 (x - floor x) + (y - floor y)."""
    bs = NV_GCF // 2
    mx = np.floor(vis[:, 0]).astype(np.int64)
    my = np.floor(vis[:, 1]).astype(np.int64)
    blk = _c_div(_c_div(my, bs) * (2048 + bs - 1), bs) + _c_div(mx, bs)
    sub = OVER * (vis[:, 0] - mx) + (vis[:, 1] - my)
    return np.lexsort((sub, blk))


def nv_pixel(x: float, y: float, scale: float):
    """grid_gpu.cu:37-51: (main_x, main_y, over) by floor, over = 8 sub_y + sub_x."""
    xs, ys = x * scale, y * scale
    mx, my = int(np.floor(xs)), int(np.floor(ys))
    sx, sy = int(np.floor(OVER * (xs - mx))), int(np.floor(OVER * (ys - my)))
    return mx, my, OVER * sy + sx


def nv_grid(vis: np.ndarray, gcf: np.ndarray, scale: float, grid_size: int) -> np.ndarray:
    """grid_gpu.cu:83-155 as a scatter: out[y][x] (x fast, complex as [..., 2])
    = sum over points of vis * gcf[over][b][a] at x = main_x + a - 7,
    y = main_y + b - 7, taps outside the grid dropped."""
    out = np.zeros((grid_size, grid_size), np.complex128)
    g = gcf[..., 0] + 1j * gcf[..., 1]
    for x, y, _, re, im in vis:
        mx, my, over = nv_pixel(x, y, scale)
        for b in range(NV_GCF):
            py = my + b - 7
            if not 0 <= py < grid_size:
                continue
            for a in range(NV_GCF):
                px = mx + a - 7
                if 0 <= px < grid_size:
                    out[py, px] += (re + 1j * im) * g[over, b, a]
    return np.stack([out.real, out.imag], -1)
