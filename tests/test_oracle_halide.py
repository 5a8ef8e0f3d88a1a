"""Pinning of the MS6 Halide-kernel oracle (oracle/halide_oracle.py; the
Halide generators cannot be built here): delta-tap KATs through the pixel
and oversampling mapping, the bounds KAT, a scalar-loop brute force, and
the bilinear transpose identity between scatter and degrid."""
import numpy as np

from oracle import halide_oracle as H


def delta_gcf(S, over, y, x, val=1.0 + 0.0j):
    g = np.zeros((64, S, S, 2))
    g[over, y, x] = (val.real, val.imag)
    return g


def test_delta_tap_lands_on_its_cell():
    S, N = 8, 64
    for ou, ov, a, b in ((0, 0, 0, 0), (3, 5, 2, 7), (7, 1, 7, 0), (4, 6, 5, 3)):
        # fractions (o + 0.25)/8 away from rounding halves; round() of the
        # pixel goes up from o >= 4
        u, v = 10 + (ou + 0.25) / 8, -7 + (ov + 0.25) / 8
        vis = np.array([[u, v, 0.0, 2.0, -1.0]])
        uvg = np.zeros((N, N, 2))
        H.scatter(vis, delta_gcf(S, ou + 8 * ov, a, b), uvg, 1.0, N, 0)
        x0 = 10 + (1 if ou >= 4 else 0) + N // 2 - S // 2
        y0 = -7 + (1 if ov >= 4 else 0) + N // 2 - S // 2
        nz = np.argwhere(uvg[..., 0] != 0)
        assert nz.tolist() == [[y0 + a, x0 + b]]
        assert tuple(uvg[y0 + a, x0 + b]) == (2.0, -1.0)


def test_pixel_rounds_halves_to_even():
    # Halide's round: 2.5 -> 2, 3.5 -> 4 (numpy.rint); oversampling offset
    # 8 * 0.5 = 4 exactly stays 4
    assert H.pixel(2.5, 3.5, 1.0, 16, 8)[:2] == (2 + 4, 4 + 4)
    assert H.pixel(2.5, 3.5, 1.0, 16, 8)[2] == 4 + 8 * 4


def test_bounds_are_inclusive_of_extent_minus_s_minus_one():
    S, N = 8, 32
    gcf = np.ones((64, S, S, 2))
    for x0, inside in ((N - S - 1, True), (N - S, False), (0, True), (-1, False)):
        u = x0 - (N // 2 - S // 2) + 0.1
        vis = np.array([[u, 0.1, 0, 1.0, 0.0]])
        uvg = np.zeros((N, N, 2))
        H.scatter(vis, gcf, uvg, 1.0, N, 0)
        assert bool(uvg.any()) == inside, x0
    # a margin of 12 with S = 8 shrinks the range by 2 on each side
    assert H.in_bounds(2, 2, 8, 32, 32, (0, 0), 12)
    assert not H.in_bounds(1, 2, 8, 32, 32, (0, 0), 12)
    assert not H.in_bounds(N - S - 2, 2, 8, 32, 32, (0, 0), 12)


def brute_scatter(vis, gcf, uvg, scale, N, margin, origin):
    S = gcf.shape[1]
    Y, X = uvg.shape[:2]
    for u, v, _, re, im in vis:
        us, vs = u * scale, v * scale
        ou = min(max(int(np.rint(8 * (us - np.floor(us)))), 0), 7)
        ov = min(max(int(np.rint(8 * (vs - np.floor(vs)))), 0), 7)
        x0 = int(np.rint(us)) + N // 2 - S // 2
        y0 = int(np.rint(vs)) + N // 2 - S // 2
        m = max(0, (margin - S) // 2)
        if not (origin[0] + m <= x0 <= origin[0] + X - S - 1 - m and
                origin[1] + m <= y0 <= origin[1] + Y - S - 1 - m):
            continue
        for sy in range(S):
            for sx in range(S):
                gr, gi = gcf[ou + 8 * ov, sy, sx]
                uvg[y0 + sy - origin[1], x0 + sx - origin[0], 0] += re * gr - im * gi
                uvg[y0 + sy - origin[1], x0 + sx - origin[0], 1] += re * gi + im * gr


def test_scatter_matches_scalar_loops():
    rng = np.random.default_rng(0)
    S, N = 8, 48
    vis = np.column_stack([rng.uniform(-30, 30, 60), rng.uniform(-30, 30, 60),
                           rng.standard_normal(60), rng.standard_normal(60),
                           rng.standard_normal(60)])
    gcf = rng.standard_normal((64, S, S, 2))
    a = np.zeros((40, 44, 2))
    b = np.zeros((40, 44, 2))
    H.scatter(vis, gcf, a, 0.7, N, 10, origin=(3, 5))
    brute_scatter(vis, gcf, b, 0.7, N, 10, origin=(3, 5))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)
    assert np.count_nonzero(a) > 0


def test_degrid_is_the_transpose_of_scatter():
    rng = np.random.default_rng(1)
    S, N = 16, 64
    vis = np.column_stack([rng.uniform(-20, 20, 40), rng.uniform(-20, 20, 40),
                           np.zeros(40), rng.standard_normal(40), rng.standard_normal(40)])
    gcf = rng.standard_normal((64, S, S, 2))
    g = rng.standard_normal((N, N, 2))
    sc = np.zeros((N, N, 2))
    H.scatter(vis, gcf, sc, 1.0, N, 0)
    lhs = np.sum((sc[..., 0] + 1j * sc[..., 1]) * (g[..., 0] + 1j * g[..., 1]))
    out = H.degrid(gcf, g, vis, 1.0, N, 0)
    dg = (vis[:, 3] - out[:, 3]) + 1j * (vis[:, 4] - out[:, 4])
    rhs = np.sum((vis[:, 3] + 1j * vis[:, 4]) * dg)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    assert np.array_equal(out[:, :3], vis[:, :3])


# --- nvGridder (MS6/kernel/nvidia/gridder) ---------------------------------

def nv_delta(over, row, col, val=1.0 + 0.0j):
    g = np.zeros((64, 16, 16, 2))
    g[over, row, col] = (val.real, val.imag)
    return g


def test_nv_delta_tap_lands_on_its_pixel():
    # grid_gpu.cu:131-132 with d_gcf advanced by 16*15/2 - 1 = 119
    # (grid_gpu.cu:201): pixel (main_x + a, main_y + b) reads flat tap
    # 256*over + 16*b + a + 119 = [over][b + 7][a + 7]
    N = 64
    for scale, (x, y), (sx, sy), (row, col) in ((1.0, (20.3, 30.9), (2, 7), (3, 12)),
                                                (0.5, (41.0, 61.5), (4, 6), (15, 0))):
        vis = np.array([[x, y, 0.0, 2.0, -1.0]])
        g = nv_delta(8 * sy + sx, row, col)
        out = H.nv_grid(vis, g, scale, N)
        mx, my = int(np.floor(x * scale)), int(np.floor(y * scale))
        nz = np.argwhere(out[..., 0] != 0)
        assert nz.tolist() == [[my + row - 7, mx + col - 7]]
        assert tuple(out[my + row - 7, mx + col - 7]) == (2.0, -1.0)
    # the sub-pixel offsets of those points, by hand
    assert H.nv_pixel(20.3, 30.9, 1.0) == (20, 30, 8 * 7 + 2)
    assert H.nv_pixel(41.0, 61.5, 0.5) == (20, 30, 8 * 6 + 4)


def test_nv_subpixel_carry():
    """x a hair below an integer: 8 (x - floor x) rounds to 8, which the
    reference's packing 8 main + sub (grid_gpu.cu:44-49) carries into the
    next pixel with offset 0: x = -1e-300 is pixel 0, offset 0 (not pixel
    -1 with offset 7 or 8)."""
    x = -1e-300
    assert np.floor(x) == -1.0 and np.floor(8 * (x - np.floor(x))) == 8.0
    assert H.nv_split(x) == (0, 0)
    assert H.nv_pixel(x, 5.25, 1.0) == (0, 5, 8 * 2 + 0)
    assert H.nv_pixel(7.0 - 2.0 ** -50, 3.0, 1.0) == (6, 3, 7)  # an ordinary offset 7
    # the point lands where its packed value says: taps of layer 8 sy + 0
    N = 24
    vis = np.array([[x, 10.0, 0.0, 1.0, 0.0]])
    g = nv_delta(8 * 0 + 0, 7, 7)
    out = H.nv_grid(vis, g, 1.0, N)
    assert np.argwhere(out[..., 0] != 0).tolist() == [[10, 0]]


def test_nv_taps_outside_the_grid_are_dropped():
    N = 32
    vis = np.array([[2.1, 2.1, 0.0, 1.0, 0.0]])
    g = np.ones((64, 16, 16, 2))
    out = H.nv_grid(vis, g, 1.0, N)
    # pixels 2 + a - 7 >= 0 need a >= 5: 11 columns and 11 rows
    assert np.count_nonzero(out[..., 0]) == 11 * 11
    assert out[0, 0, 0] == 1.0 and out[2 + 8, 2 + 8, 0] == 1.0


def test_nv_sort_order_keys():
    # block key (gy * 2055) / 8 + gx on unscaled floor coordinates, C
    # division: x = -3.5 has floor -4 and -4 / 8 = 0 in C (block 0, not -1),
    # so it follows x = 1.0 of the same block, whose sub-key is smaller
    vis = np.array([[-3.5, 0.25, 0, 0, 0],    # block 0, sub 8*0.5 + 0.25 = 4.25
                    [1.0, 0.5, 0, 0, 0],      # block 0, sub 0.5
                    [0.5, 8.5, 0, 0, 0],      # gy 1: block 2055 / 8 = 256
                    [9.0, 0.0, 0, 0, 0],      # block 1, sub 0
                    [2060.5, 0.5, 0, 0, 0]])  # gx 257: block 257, after gy = 1
    assert H.nv_sort_order(vis).tolist() == [1, 0, 3, 2, 4]
    # equal keys keep their input order
    same = np.array([[1.25, 2.5, 0, 0, k] for k in range(4)], float)
    assert H.nv_sort_order(same).tolist() == [0, 1, 2, 3]


def test_nv_grid_matches_the_reference_gather_loops():
    """The pixel-centric gather of grid_gpu.cu:119-139 written with its own
    index arithmetic (flat tap 256*over + 16*b + a + 119, a and b in
    (-8, 8]) on tiny inputs, against the oracle's scatter."""
    rng = np.random.default_rng(3)
    N, n = 24, 20
    vis = np.column_stack([rng.uniform(0, N, n), rng.uniform(0, N, n), np.zeros(n),
                           rng.standard_normal(n), rng.standard_normal(n)])
    gcf = rng.standard_normal((64, 16, 16, 2))
    flat = (gcf[..., 0] + 1j * gcf[..., 1]).reshape(-1)
    exp = np.zeros((N, N), complex)
    for this_y in range(N):
        for this_x in range(N):
            for x, y, _, re, im in vis:
                ix = int(np.floor(x)) * 8 + int(np.floor(8 * (x - np.floor(x))))
                iy = int(np.floor(y)) * 8 + int(np.floor(8 * (y - np.floor(y))))
                main_x, sub_x = ix // 8, ix % 8
                main_y, sub_y = iy // 8, iy % 8
                a, b = this_x - main_x, this_y - main_y
                if not (-8 < a <= 8 and -8 < b <= 8):
                    continue
                exp[this_y, this_x] += (re + 1j * im) * flat[256 * (8 * sub_y + sub_x) +
                                                            16 * b + a + 119]
    got = H.nv_grid(vis, gcf, 1.0, N)
    assert np.allclose(got[..., 0] + 1j * got[..., 1], exp, rtol=0, atol=1e-12)


# --- MS6 imaging glue -------------------------------------------------------

def test_rotation_at_the_pole_is_a_z_rotation():
    # uvw2xyz(lon, pi/2) = Z(pi/2 + lon), xyz2uvw(lon', pi/2) = Z(-pi/2 - lon'),
    # so R = Z(lon - lon'): no phase, uv turned by lon - lon' under uvproj
    pi = np.pi
    R, mtx, pos = H.rotation(0.3, pi / 2, 0.1, pi / 2, 1)
    d = 0.3 - 0.1
    Z = np.array([[np.cos(d), -np.sin(d), 0], [np.sin(d), np.cos(d), 0], [0, 0, 1]])
    assert np.allclose(R, Z, atol=1e-15) and np.allclose(mtx, Z, atol=1e-15)
    assert np.allclose(pos, 0, atol=1e-15)
    _, ident, _ = H.rotation(0.3, pi / 2, 0.1, pi / 2, 0)
    assert np.array_equal(ident, np.eye(3))


def test_rotation_in_latitude_gives_the_phase_path():
    # at lon 0: R = X(lat' - pi/2) X(pi/2 - lat) = X(delta), pos = w - X(delta) w
    # = (0, sin delta, 1 - cos delta)
    lat, delta = -0.7, 0.02
    R, _, pos = H.rotation(0.0, lat, 0.0, lat + delta, 0)
    X = np.array([[1, 0, 0], [0, np.cos(delta), -np.sin(delta)], [0, np.sin(delta), np.cos(delta)]])
    assert np.allclose(R, X, atol=1e-15)
    assert np.allclose(pos, [0, np.sin(delta), 1 - np.cos(delta)], atol=1e-15)
    vis = np.array([[10.0, 20.0, 5.0, 1.0, 0.0]])
    out = H.rotate(vis, 0.0, lat, 0.0, lat, 0.0, delta, 0, 0, 1)  # facet m = 1
    phi = 2 * np.pi * (20.0 * np.sin(delta) + 5.0 * (1 - np.cos(delta)))
    assert np.allclose(out[0], [10.0, 20.0, 5.0, np.cos(phi), np.sin(phi)], atol=1e-12)


def test_rotation_is_orthogonal():
    rng = np.random.default_rng(1)
    for _ in range(5):
        a = rng.uniform(-1.5, 1.5, 4)
        R, mtx, _ = H.rotation(*a, 1)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14
        # the projected transform keeps the w axis of the rotated frame
        assert np.isfinite(mtx).all()


def test_detile_adds_the_overlap():
    a = np.arange(3 * 4 * 2, dtype=float).reshape(3, 4, 2)   # x 5..8, y 2..4
    out = np.ones((4, 3, 2))                                  # x 7..9, y 3..6
    H.detile(a, (5, 2), out, (7, 3))
    exp = np.ones((4, 3, 2))
    # overlap x 7..8, y 3..4
    exp[0:2, 0:2] += a[1:3, 2:4]
    assert np.array_equal(out, exp)


def test_defacet_last_write_wins_and_clamps_rows_with_the_x_bounds():
    # image x 2..9, y 0..11; facets of 2x2 at l in 0..1, m = 0, x, y from 8
    img = np.full((12, 8), -1.0)
    facet = np.arange(1 * 2 * 2 * 2, dtype=float).reshape(1, 2, 2, 2) + 1  # [m][l][y][x]
    H.defacet(facet, (8, 8, 0, 0), img, (2, 0))
    # x + l: 8, 9, 9, 10 -> clamp to 2..9: 8, 9, 9, 9; y + m: 8, 9 (inside 2..9)
    # the last writer of (9, y) is l = 1, x = 9
    assert img[8, 8 - 2] == facet[0, 0, 0, 0] and img[8, 9 - 2] == facet[0, 1, 0, 1]
    assert img[9, 9 - 2] == facet[0, 1, 1, 1]
    # rows outside 2..9 are never written, whatever y + m is
    facet2 = np.ones((1, 1, 3, 1))
    img2 = np.zeros((12, 8))
    H.defacet(facet2, (2, 9, 0, 1), img2, (2, 0))   # y + m = 10, 11, 12 -> row 9
    assert img2[9, 0] == 1.0 and np.count_nonzero(img2) == 1


def test_image_sum_accumulates_clamped_borders():
    img = np.ones((4, 4))                # x 0..3, y 0..3
    out = np.zeros((2, 2))               # x 1..2, y 1..2
    H.image_sum(img, (0, 0), out, (1, 1))
    # corner (1, 1) takes x in 0..1, y in 0..1: 4 cells; likewise every corner
    assert np.array_equal(out, np.full((2, 2), 4.0))
