"""Pinning of the GCF oracle (oracle/gcf_oracle.py, a restatement of
MS4/kernel/cpu/gcf/GCF.cpp, which needs FFTW and cannot be built here):
closed form of the w = 0 layer, a direct DFT of a tiny arena with explicit
index loops, the normalisation invariant, the pad quirk, and calc_accums
against a scalar loop with C rounding."""
import math

import numpy as np
import pytest

from oracle import gcf_oracle as G

OVER = G.OVER


def direct_layer(support, max_support, src_pad, t2, w, pad_quirk=False):
    """Independent restatement: explicit shifts, a DFT by matrix products,
    sums and extraction by scalar index arithmetic."""
    L = max_support * OVER
    pitch = L + src_pad
    arena = np.zeros((L, pitch), np.complex128)
    radius = max_support // 2
    normer = t2 / radius
    for a in range(max_support):
        for b in range(max_support):
            i, j = a - radius, b - radius
            x, y = i * normer, j * normer
            ph = w * (1 - math.sqrt(1 - x * x - y * y))
            arena[L // 2 - radius + a, L // 2 - radius + b] = complex(
                math.cos(2 * math.pi * ph), -math.sin(2 * math.pi * ph))
    h = L // 2
    shift = lambda m: np.roll(np.roll(m, h, 0), h, 1)  # noqa: E731 (even L)
    k = np.arange(L)
    E = np.exp(2j * math.pi * np.outer(k, k) / L)  # backward DFT matrix
    arena[:, :L] = shift(E @ shift(arena[:, :L]) @ E.T)
    flat = arena.reshape(-1)
    sums = np.zeros((OVER, OVER))
    for su in range(max_support):
        for ou in range(OVER):
            for sv in range(max_support):
                for ov in range(OVER):
                    if pad_quirk:
                        e = flat[((su * OVER + ou) * max_support + sv) * OVER + ov]
                    else:
                        e = arena[su * OVER + ou, sv * OVER + ov]
                    sums[ou, ov] += e.real
    st = (max_support // 2 - support // 2) * OVER
    out = np.zeros((OVER, OVER, support, support), np.complex128)
    for ou in range(OVER):
        for ov in range(OVER):
            for su in range(support):
                for sv in range(support):
                    out[ou, ov, su, sv] = arena[st + ou + su * OVER, st + ov + sv * OVER] / sums[ou, ov]
    return out


@pytest.mark.parametrize("support,max_support,pad,w,quirk", [
    (4, 4, 0, 0.0, False), (2, 4, 3, 1.7, False), (4, 4, 3, -2.3, True), (3, 5, 2, 0.9, False)])
def test_layer_matches_direct_dft(support, max_support, pad, w, quirk):
    got = G.mk_gcf_layer(support, max_support, pad, 0.3, w, pad_quirk=quirk)
    exp = direct_layer(support, max_support, pad, 0.3, w, pad_quirk=quirk)
    assert np.abs(got - exp).max() <= 1e-12 * np.abs(exp).max()


def test_w0_layer_is_a_normalised_dirichlet_kernel():
    """w = 0: the screen is a box of ones, so after the shifted backward DFT
    every arena cell is f(m1) f(m2), f(m) = sum_{j=-r}^{ms-r-1} (-1)^j
    e^{2 pi i j m / L}."""
    ms, S = 8, 6
    L = ms * OVER
    r = ms // 2
    m = np.arange(L)
    j = np.arange(-r, ms - r)
    f = ((-1.0) ** j[None, :] * np.exp(2j * math.pi * np.outer(m, j) / L)).sum(axis=1)
    arena = np.outer(f, f)
    sums = arena.real.reshape(ms, OVER, ms, OVER).sum(axis=(0, 2))
    st = (ms // 2 - S // 2) * OVER
    got = G.mk_gcf_layer(S, ms, 4, 0.2, 0.0)
    for ou in (0, 3, 7):
        for ov in (0, 5):
            exp = arena[st + ou::OVER, st + ov::OVER][:S, :S] / sums[ou, ov]
            assert np.abs(got[ou, ov] - exp).max() <= 1e-12 * np.abs(exp).max()


def test_full_support_layers_sum_to_one():
    lay = G.mk_gcf_layer(8, 8, 0, 0.25, 3.1)
    np.testing.assert_allclose(lay.real.sum(axis=(2, 3)), np.ones((OVER, OVER)), rtol=0,
                               atol=1e-12)


def test_pad_quirk_only_matters_with_padding():
    a = G.mk_gcf_layer(6, 8, 0, 0.25, 2.0, pad_quirk=False)
    b = G.mk_gcf_layer(6, 8, 0, 0.25, 2.0, pad_quirk=True)
    assert np.array_equal(a, b)
    c = G.mk_gcf_layer(6, 8, 4, 0.25, 2.0, pad_quirk=False)
    d = G.mk_gcf_layer(6, 8, 4, 0.25, 2.0, pad_quirk=True)
    assert np.array_equal(a, c)  # padding itself changes nothing
    assert np.abs(c - d).max() > 1e-6


def test_calc_accums_matches_scalar_loop():
    rng = np.random.default_rng(0)
    wstep = 0.5
    w = rng.uniform(-3.2, 3.2, 500)
    w[:20] = (np.arange(20) - 10) * 0.25  # exact halves of a step
    w[20] = np.nextafter(0.25, 0.0)
    uvw = np.zeros((w.size, 3))
    uvw[:, 2] = w
    n_planes = 15
    sums, npts = G.calc_accums(uvw, wstep, n_planes)
    es, en = np.zeros(n_planes), np.zeros(n_planes, np.int64)
    for x in w:
        q = x / wstep
        r = math.floor(abs(q))
        if abs(q) - r >= 0.5:
            r += 1
        p = int(math.copysign(r, q)) + n_planes // 2
        es[p] += x
        en[p] += 1
    assert np.array_equal(npts, en)
    np.testing.assert_allclose(sums, es, rtol=1e-13, atol=1e-13)


def test_kernel_uses_mean_w_and_nominal_w_for_empty_planes():
    uvw = np.zeros((4, 3))
    uvw[:, 2] = [0.9, 1.1, 1.2, -0.1]  # plane 1: mean 1.0667; plane 0: -0.1
    supps, layers, means = G.gcf_kernel(uvw, 1.0, 0.01, 4, 2, 8, 2, src_pad=4)
    assert supps == [8, 6, 4, 6, 8]
    assert means[2] == pytest.approx(-0.1) and means[3] == pytest.approx(3.2 / 3)
    assert means[0] == -2.0 and means[4] == 2.0  # no points: nominal p * wstep
    np.testing.assert_allclose(layers[3], G.mk_gcf_layer(6, 8, 4, 0.005, means[3] * 0.01),
                               rtol=0, atol=0)
