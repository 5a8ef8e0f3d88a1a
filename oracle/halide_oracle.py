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


def nv_split(xs: float):
    """grid_gpu.cu:44-49: m = floor(xs), s = floor(8 (xs - m)), packed as
    8 m + s; s = 8 (xs a hair below an integer, where xs - m rounds to 1.0)
    carries into pixel m + 1, offset 0, as the packed value does."""
    m = int(np.floor(xs))
    s = int(np.floor(OVER * (xs - m)))
    return (m + 1, s - OVER) if s >= OVER else (m, s)


def nv_pixel(x: float, y: float, scale: float):
    """grid_gpu.cu:37-51: (main_x, main_y, over), over = 8 sub_y + sub_x."""
    mx, sx = nv_split(x * scale)
    my, sy = nv_split(y * scale)
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


# --- MS6 imaging glue (MS6/kernel/cpu/gridding) --------------------------------
# Arrays are numpy views of Halide buffers with the last index fastest:
# a buffer of dimensions (x0, x1, ...) is an array [..., x1, x0]; `mn` gives
# the Halide min of each dimension (x0 first).

def psf_vis(vis: np.ndarray) -> np.ndarray:
    """psf_vis.cpp:17-20: re = 1, im = 0, u v w kept (vis [n, 5])."""
    out = vis.copy()
    out[:, 3] = 1.0
    out[:, 4] = 0.0
    return out


def _rx(p):
    s, c = np.sin(p), np.cos(p)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _rz(p):
    s, c = np.sin(p), np.cos(p)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def rotation(in_lon, in_lat, out_lon, out_lat, uvproj):
    """rotate.cpp:33-52 with utils.h:137-158: (R, mtx, pos) where
    R = xyz2uvw(out) uvw2xyz(in), mtx the uvw transform (identity unless
    uvproj) and pos = w - R w."""
    pi = 4 * np.arctan(1.0)
    xyz2uvw = _rx(-(pi / 2) + out_lat) @ _rz(-(pi / 2) - out_lon)
    uvw2xyz = _rz((pi / 2) + in_lon) @ _rx((pi / 2) - in_lat)
    R = xyz2uvw @ uvw2xyz
    if uvproj:
        det = R[0, 0] * R[1, 1] - R[0, 1] * R[1, 0]
        inv = np.array([[R[1, 1] / det, -R[0, 1] / det, 0], [-R[1, 0] / det, R[0, 0] / det, 0],
                        [0, 0, 1]])
        mtx = R.T @ (inv.T @ R)
    else:
        mtx = np.eye(3)
    pos = np.array([0.0, 0.0, 1.0]) - R @ np.array([0.0, 0.0, 1.0])
    return R, mtx, pos


def rotate(vis: np.ndarray, in_lon, in_lat, out_lon0, out_lat0, lon_incr, lat_incr, uvproj,
           l_min: int, m_min: int) -> np.ndarray:
    """rotate.cpp:28-68: the facet at (l_min, m_min) (the output buffer's
    min[2], min[3]) has centre (out_lon0 + l_min lon_incr, out_lat0 + m_min
    lat_incr); uvw' = mtx uvw, vis' = vis exp(2 pi i pos.uvw)."""
    out_lon = out_lon0 + l_min * lon_incr
    out_lat = out_lat0 + m_min * lat_incr
    _, mtx, pos = rotation(in_lon, in_lat, out_lon, out_lat, uvproj)
    uvw = vis[:, :3]
    out = np.empty_like(vis)
    out[:, :3] = uvw @ mtx.T
    phi = 2 * (4 * np.arctan(1.0)) * (uvw @ pos)
    v = (vis[:, 3] + 1j * vis[:, 4]) * (np.cos(phi) + 1j * np.sin(phi))
    out[:, 3], out[:, 4] = v.real, v.imag
    return out


def detile(in_grid: np.ndarray, in_mn, out_grid: np.ndarray, out_mn) -> None:
    """detile.cpp:19-30: out += in over the overlap of the (x, y) ranges;
    grids [Y, X, 2], mn = (x, y) of the first cell."""
    x0, y0 = max(in_mn[0], out_mn[0]), max(in_mn[1], out_mn[1])
    x1 = min(in_mn[0] + in_grid.shape[1], out_mn[0] + out_grid.shape[1])
    y1 = min(in_mn[1] + in_grid.shape[0], out_mn[1] + out_grid.shape[0])
    if x1 <= x0 or y1 <= y0:
        return
    out_grid[y0 - out_mn[1]:y1 - out_mn[1], x0 - out_mn[0]:x1 - out_mn[0]] += \
        in_grid[y0 - in_mn[1]:y1 - in_mn[1], x0 - in_mn[0]:x1 - in_mn[0]]


def defacet(facet: np.ndarray, f_mn, image: np.ndarray, i_mn) -> None:
    """defacet.cpp:12-20: image(clamp(x + l, left, right), clamp(y + m, left,
    right)) = facet(x, y, l, m) in Halide order (x fastest), so the last
    write wins; left / right are the image's x bounds, used for y as well.
    facet [M, L, Y, X], f_mn = (x, y, l, m); image [Y, X], i_mn = (x, y)."""
    left, right = i_mn[0], i_mn[0] + image.shape[1] - 1
    M, L, Y, X = facet.shape
    for m in range(f_mn[3], f_mn[3] + M):
        for l in range(f_mn[2], f_mn[2] + L):
            for y in range(f_mn[1], f_mn[1] + Y):
                for x in range(f_mn[0], f_mn[0] + X):
                    ix = min(max(x + l, left), right)
                    iy = min(max(y + m, left), right)
                    image[iy - i_mn[1], ix - i_mn[0]] = \
                        facet[m - f_mn[3], l - f_mn[2], y - f_mn[1], x - f_mn[0]]


def image_sum(image: np.ndarray, i_mn, out: np.ndarray, o_mn) -> None:
    """image_sum.cpp:12-20: out(clamp(x, left, right), clamp(y, left, right))
    += image(x, y); left / right are out's x bounds, used for y as well."""
    left, right = o_mn[0], o_mn[0] + out.shape[1] - 1
    Y, X = image.shape
    for y in range(i_mn[1], i_mn[1] + Y):
        for x in range(i_mn[0], i_mn[0] + X):
            ox, oy = min(max(x, left), right), min(max(y, left), right)
            out[oy - o_mn[1], ox - o_mn[0]] += image[y - i_mn[1], x - i_mn[0]]
