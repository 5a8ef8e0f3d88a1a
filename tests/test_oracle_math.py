"""The oracle against what the mathematics fixes: known-answer tests,
invariants and brute force on tiny inputs (CPU only).  These pin the parts
the golden vectors leave open and would catch a dropped term, a sign or an
index error in the restatement."""
import ctypes as C

import numpy as np
import pytest

from oracle import fft_oracle as F
from oracle import oracle as O
from rc_amd import synth
from rc_amd.tables import OVER, GCFTable, VisChunks


def delta_table(S, pos, value=1.0 + 0j, planes=(0,), half=False):
    """Every layer of every plane: a single tap `value` at window position
    pos = (su, sv)."""
    out = {}
    for p in planes:
        t = np.zeros((OVER, OVER, S, S), np.complex128)
        t[:, :, pos[0], pos[1]] = value
        out[p] = t
    return GCFTable.from_planes(out, half=half)


def one_point(u, v, w=0.0, S=4, vis=1.0 + 0j, scale=1.0, wstep=1.0):
    return VisChunks(uvw=np.array([[[u, v, w]]]), vis=np.array([[vis]]),
                     supps=np.array([S], np.int32), scale=scale, wstep=wstep)


def test_kat_single_visibility_delta_tap():
    """Delta tap at window (1, 2): the grid holds vis * tap at
    (floor(u) + N/2 - S/2 + 1, floor(v) + N/2 - S/2 + 2) and nothing else."""
    N, S = 32, 4
    gcf = delta_table(S, (1, 2), 0.5 - 0.25j)
    vc = one_point(3.7, -5.2, vis=2.0 + 1.0j, S=S)
    g, taps = O.grid(vc, gcf, N)
    assert taps == S * S
    nz = np.argwhere(g != 0)
    assert nz.tolist() == [[3 + 16 - 2 + 1, -6 + 16 - 2 + 2]]
    assert g[nz[0][0], nz[0][1]] == (2.0 + 1.0j) * (0.5 - 0.25j)


def test_kat_two_visibilities_ms2_style():
    """The pattern of the MS2 conformance test: two visibilities with a
    constant support; the grid equals vis * tap over each window and the
    overlap holds the sum."""
    N, S = 16, 3
    t = np.full((OVER, OVER, S, S), 1.0 + 0j)
    gcf = GCFTable.from_planes({0: t}, half=False)
    vc = VisChunks(uvw=np.array([[[0.0, 0.0, 0.0], [1.0, 1.0, 0.0]]]),
                   vis=np.array([[1.0 + 0j, 2.0 + 0j]]), supps=np.array([S], np.int32),
                   scale=1.0, wstep=1.0)
    g, taps = O.grid(vc, gcf, N)
    assert taps == 2 * S * S
    exp = np.zeros((N, N), complex)
    u0 = N // 2 - S // 2
    exp[u0:u0 + S, u0:u0 + S] += 1
    exp[u0 + 1:u0 + 1 + S, u0 + 1:u0 + 1 + S] += 2
    assert np.array_equal(g, exp)


def test_oversampling_selects_layer():
    """over_u = floor(8 * frac(u * scale)): a table whose layer (ou, ov) holds
    the value ou*8 + ov at the window corner reports it back."""
    N, S = 32, 2
    t = np.zeros((OVER, OVER, S, S), np.complex128)
    for ou in range(OVER):
        for ov in range(OVER):
            t[ou, ov, 0, 0] = ou * 8 + ov + 1
    gcf = GCFTable.from_planes({0: t}, half=False)
    for fu, fv in [(0.0, 0.0), (0.125, 0.5), (0.99, 0.374), (0.5, 0.875)]:
        g, _ = O.grid(one_point(2 + fu, -3 + fv, S=S), gcf, N)
        val = g[2 + 16 - 1, -3 + 16 - 1].real
        assert val == int(np.floor(8 * fu)) * 8 + int(np.floor(8 * fv)) + 1


def test_w_plane_rounding_and_half_conjugation():
    """w_plane = round(w / wstep) (half away from zero); a half table uses
    plane |p| and conjugates for negative w."""
    N, S = 32, 2
    planes = {}
    for p in range(3):
        t = np.zeros((OVER, OVER, S, S), np.complex128)
        t[:, :, 0, 0] = (p + 1) + 1j
        planes[p] = t
    gcf = GCFTable.from_planes(planes, half=True)
    for w, p, conj in [(0.49, 0, False), (0.5, 1, False), (-0.5, 1, True), (-1.6, 2, True),
                       (2.4, 2, False)]:
        g, _ = O.grid(one_point(0.0, 0.0, w=w, S=S), gcf, N)
        val = g[15, 15]
        assert val == ((p + 1) - 1j if conj else (p + 1) + 1j), (w, p)


def test_out_of_bound_strict_upper_limit():
    """In-bound needs 0 <= u0 < N - S (strict): u0 = N - S is skipped."""
    N, S = 16, 4
    gcf = delta_table(S, (0, 0))
    # u0 = floor(u) + N/2 - S/2 = N - S  <=>  floor(u) = N/2 - S/2
    g, taps = O.grid(one_point(N / 2 - S / 2, 0.0, S=S), gcf, N)
    assert taps == 0 and not g.any()
    g, taps = O.grid(one_point(N / 2 - S / 2 - 1, 0.0, S=S), gcf, N)
    assert taps == S * S


def test_support_truncation_recentres_window():
    """A chunk support below the plane support reads the centred S_bl
    window: offset (gs - S)/2 * (gs + 1)."""
    N, gs, S = 32, 6, 2
    t = np.zeros((OVER, OVER, gs, gs), np.complex128)
    t[:, :, :, :] = np.arange(gs * gs).reshape(gs, gs)
    gcf = GCFTable.from_planes({0: t}, half=False)
    g, taps = O.grid(one_point(0.0, 0.0, S=S), gcf, N)
    win = g[16 - 1:16 + 1, 16 - 1:16 + 1].real
    off = (gs - S) // 2
    assert np.array_equal(win, np.arange(gs * gs).reshape(gs, gs)[off:off + S, off:off + S])
    assert taps == S * S


@pytest.mark.parametrize("half", [False, True])
def test_linearity(half):
    vc, gcf, N = synth.random_small(31, half=half)
    rng = np.random.default_rng(0)
    v2 = rng.standard_normal(vc.vis.shape) + 1j * rng.standard_normal(vc.vis.shape)
    a, b = 0.3 - 1.1j, -2.0 + 0.5j
    g1, _ = O.grid(vc, gcf, N)
    g2, _ = O.grid(VisChunks(vc.uvw, v2, vc.supps, vc.scale, vc.wstep), gcf, N)
    g3, _ = O.grid(VisChunks(vc.uvw, a * vc.vis + b * v2, vc.supps, vc.scale, vc.wstep), gcf, N)
    assert np.linalg.norm(g3 - (a * g1 + b * g2)) <= 1e-12 * np.linalg.norm(g3)


def test_sum_invariant():
    """sum(grid) = sum over in-bound points of vis * sum(window taps)."""
    vc, gcf, N = synth.random_small(32, half=True)
    g, _ = O.grid(vc, gcf, N)
    pg = O.pregrid(vc, N, True)
    S = np.repeat(vc.supps, vc.ts_ch)
    total = 0j
    for (u, v, ou, ov, wp, idx), s, vis in zip(pg, S, vc.vis.reshape(-1)):
        if not (0 <= u < N - s and 0 <= v < N - s):
            continue
        layer = gcf.layer(idx)
        gs = layer.shape[0]
        off = (gs - s) // 2
        win = layer[off:off + s, off:off + s]
        if idx < 0:
            win = np.conj(win)
        total += vis * win.sum()
    assert abs(g.sum() - total) <= 1e-12 * max(1.0, abs(total))


@pytest.mark.parametrize("half", [False, True])
def test_degrid_is_the_transpose(half):
    """Bilinear identity sum_p (G v)_p y_p = sum_i v_i (D y)_i (no conj)."""
    vc, gcf, N = synth.random_small(33, half=half)
    rng = np.random.default_rng(1)
    y = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    g, _ = O.grid(vc, gcf, N)
    d, _ = O.degrid(vc, gcf, y)
    lhs = np.sum(g * y)
    rhs = np.sum(vc.vis * d)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_full_and_half_agree_for_conjugate_symmetric_table():
    """A full table with g(-p) = conj(g(p)) grids like the half table."""
    rng = np.random.default_rng(2)
    sup = (4, 6)
    planes = {p: rng.standard_normal((OVER, OVER, sup[p], sup[p]))
              + 1j * rng.standard_normal((OVER, OVER, sup[p], sup[p])) for p in (0, 1)}
    full = dict(planes)
    full[-1] = np.conj(planes[1])
    half_t = GCFTable.from_planes(planes, half=True)
    full_t = GCFTable.from_planes(full, half=False)
    vc, _, N = synth.random_small(34, max_plane=1, supports=sup, chunk_supp="random")
    gh, th = O.grid(vc, half_t, N)
    gf, tf = O.grid(vc, full_t, N)
    assert th == tf
    # plane 0 of a full table is not conjugated; of a half table neither
    assert np.allclose(gh, gf, rtol=0, atol=1e-12)


def test_brute_force_tiny():
    """An independent brute-force restatement on 3 points (pure Python)."""
    vc, gcf, N = synth.random_small(35, grid_size=24, chunks=1, ts_ch=3, max_plane=1,
                                    supports=(3, 5), half=False, chunk_supp="random")
    g, _ = O.grid(vc, gcf, N)
    ref = np.zeros((N, N), complex)
    S = int(vc.supps[0])
    for (u, v, w), vis in zip(vc.uvw[0], vc.vis[0]):
        us, vs = u * vc.scale, v * vc.scale
        fu, fv = np.floor(us), np.floor(vs)
        ou, ov = int(np.floor(8 * (us - fu))), int(np.floor(8 * (vs - fv)))
        p = int(np.sign(w / vc.wstep) * np.floor(abs(w / vc.wstep) + 0.5))
        u0, v0 = int(fu) + N // 2 - S // 2, int(fv) + N // 2 - S // 2
        if not (0 <= u0 < N - S and 0 <= v0 < N - S):
            continue
        layer = gcf.layer((p * 8 + ou) * 8 + ov)
        gs = layer.shape[0]
        off = (gs - S) // 2
        ref[u0:u0 + S, v0:v0 + S] += vis * layer[off:off + S, off:off + S]
    assert np.allclose(g, ref, rtol=0, atol=1e-13)


# ---------------------------------------------------------- small kernels --

def test_normalize_rotate_closed_forms():
    h = O.oracle_lib()
    rng = np.random.default_rng(3)
    a = rng.standard_normal((8, 10)) + 1j * rng.standard_normal((8, 10))
    b = a.copy()
    h.orc_normalize(b.ctypes.data, 10, 8)
    assert np.allclose(b, a / 80, rtol=1e-15)
    uvw = rng.standard_normal((50, 3)) * 100
    vis = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    r = vis.copy()
    h.orc_rotate(uvw.ctypes.data, r.ctypes.data, 50, 0.37)
    assert np.allclose(r, vis * np.exp(2j * np.pi * uvw[:, 2] * 0.37), rtol=1e-13)


def test_reweight_closed_form():
    h = O.oracle_lib()
    N = 16
    uvw = np.array([[0.2, 0.1, 0], [0.4, -0.3, 0], [3.0, 1.0, 0], [50.0, 0.0, 0]])
    vis = np.ones(4, complex) * (2 + 2j)
    h.orc_reweight(uvw.ctypes.data, vis.ctypes.data, 1.0, 4, N)
    # first two share cell (0, 0); the third is alone; the fourth is outside
    assert np.allclose(vis, [1 + 1j, 1 + 1j, 2 + 2j, 2 + 2j])


def test_herm_closed_form():
    h = O.oracle_lib()
    rng = np.random.default_rng(4)
    N, P = 8, 11
    a = rng.standard_normal((N, P)) + 1j * rng.standard_normal((N, P))
    b = a.copy()
    h.orc_herm_padded(b.ctypes.data, N, P)
    exp = a.copy()
    for i in range(N):
        for j in range(N // 2):
            exp[i, j] += np.conj(a[N - 1 - i, N - 1 - j])
    assert np.array_equal(b, exp)


def test_fftshift_is_numpy_fftshift():
    h = O.oracle_lib()
    a = (np.arange(36) + 1j).reshape(6, 6).astype(np.complex128)
    b = np.zeros((6, 9), np.complex128)
    b[:, :6] = a
    h.orc_fftshift_even(b.ctypes.data, 16, 6, 9)
    assert np.array_equal(b[:, :6], np.fft.fftshift(a))
    assert np.array_equal(F.fftshift_even(a), np.fft.fftshift(a))


@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_oracle_c2c_matches_direct_dft(sign):
    rng = np.random.default_rng(5)
    n, p = 6, 7
    buf = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    out = F.fft_inplace_even(buf, sign, n)
    a = np.fft.fftshift(buf[:, :n])
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    exp = np.fft.fftshift(W @ a @ W)
    assert np.allclose(out[:, :n], exp, rtol=0, atol=1e-12)
    assert np.array_equal(out[:, n:], buf[:, n:])


def test_fft_oracle_c2r_of_hermitian_grid_is_real_dft():
    """For a grid that is the shifted DFT of a real image, C2R returns N^2
    times the (shifted) image: the FFTW c2r convention."""
    rng = np.random.default_rng(6)
    n, p = 8, 10
    img = rng.standard_normal((n, n))
    spec = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(img)))
    buf = np.zeros((n, p), np.complex128)
    buf[:, :n] = spec
    out = F.fft_inplace_even(buf, 2, n)
    rv = F.real_view(out, n, p)[:, :n]
    assert np.allclose(rv, img * n * n, rtol=0, atol=1e-10)


def test_fft_oracle_r2c_then_c2r_round_trip():
    rng = np.random.default_rng(7)
    n, p = 8, 8
    buf = np.zeros((n, p), np.complex128)
    F.real_view(buf, n, p)[:, :n] = rng.standard_normal((n, n))
    img = F.real_view(buf, n, p)[:, :n].copy()
    spec = F.fft_inplace_even(buf, 0, n)
    # the valid half spectrum sits in columns {0} U [n/2, n) after the shift;
    # rebuild the full spectrum from it by hermitian symmetry and invert
    s = np.fft.fftshift(spec[:, :n])        # back to unshifted columns 0..n/2
    full = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(img)))
    assert np.allclose(s[:, : n // 2 + 1], np.fft.fftshift(full)[:, : n // 2 + 1], atol=1e-10)


# ---------------------------------------------------------------- Hogbom --

def _hogbom(res, psf, niters, gain, thr):
    h = O.oracle_lib()
    siz, pitch = res.shape
    mod = np.zeros_like(res)
    it = h.orc_hogbom(mod.ctypes.data, res.ctypes.data, psf.ctypes.data, siz, pitch, niters,
                      gain, thr)
    return mod, it


def gaussian_psf(n, pitch, sigma=1.5, peak=(None, None)):
    c = n // 2 if peak[0] is None else peak[0]
    x = np.arange(n)
    g = np.exp(-((x[:, None] - c) ** 2 + (x[None, :] - c) ** 2) / (2 * sigma ** 2))
    out = np.zeros((n, pitch))
    out[:, :n] = g
    return out


def test_hogbom_recovers_point_source_exactly():
    n, p = 32, 35
    psf = gaussian_psf(n, p)
    res = np.zeros((n, p))
    # dirty image of one source of flux 3 at (10, 20): psf shifted so its
    # peak (16, 16) lands there
    for i in range(n):
        for j in range(n):
            si, sj = i - 10 + 16, j - 20 + 16
            if 0 <= si < n and 0 <= sj < n:
                res[i, j] = 3.0 * psf[si, sj]
    mod, it = _hogbom(res, psf, 5, 1.0, 1e-9)
    assert mod[10, 20] == pytest.approx(3.0, rel=1e-15)
    assert np.count_nonzero(mod) == 1
    assert np.abs(res[:, :n]).max() < 1e-12
    assert it == 1


def test_hogbom_residual_identity_and_negative_peaks():
    """dirty = residual + sum_k model[r_k] * psf shifted to r_k, for a mix of
    positive and negative sources (the peak is signed)."""
    n, p = 24, 24
    psf = gaussian_psf(n, p, sigma=1.2)
    rng = np.random.default_rng(8)
    dirty = np.zeros((n, p))
    for (r, c, f) in [(5, 5, 2.0), (12, 17, -1.5), (18, 8, 0.7)]:
        for i in range(n):
            for j in range(n):
                si, sj = i - r + n // 2, j - c + n // 2
                if 0 <= si < n and 0 <= sj < n:
                    dirty[i, j] += f * psf[si, sj]
    dirty[:, :n] += 1e-3 * rng.standard_normal((n, n))
    res = dirty.copy()
    mod, it = _hogbom(res, psf, 200, 0.1, 1e-3)
    assert 0 < it <= 200
    recon = res.copy()
    for r, c in zip(*np.nonzero(mod)):
        for i in range(n):
            for j in range(n):
                si, sj = i - r + n // 2, j - c + n // 2
                if 0 <= si < n and 0 <= sj < n:
                    recon[i, j] += mod[r, c] * psf[si, sj]
    assert np.allclose(recon, dirty, atol=1e-12)
    assert mod[12, 17] < 0 < mod[5, 5]
