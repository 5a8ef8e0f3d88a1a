"""Synthetic workloads (seeded) for tests and the benchmark.

Shapes follow the benchmark configurations of this project:

  plain_c1      512 x 512 grid, 100,000 points, u, v ~ U(-240, 240) px,
                w = 0, one plane of 8 x 8 separable Gaussian taps (sigma 1.5 px)
                at 8x oversampling, 100 chunks x 1000 points, support 8.
  wproj         4096 x 4096 grid, 32 |w|-planes with support growing
                linearly 8 -> 64, points from 256 antennas (2-D Gaussian,
                sigma 800 m, clipped to a 3 km radius disc) tracked over 4 h
                at declination -45 deg, 1.4 GHz; half GCF table (conjugated
                taps for w < 0); one chunk per point so that every point uses
                the support of its own plane.
  random_small  small randomised cases exercising truncation, negative
                planes, out-of-bound points and odd supports.

Taps: the Gaussian of plain_c1 is the closed form of the configuration.
For w-planes a smooth w-dependent kernel is used: a separable Gaussian
envelope (sigma = S/6) times the chirp exp(i*pi*c_p*r^2), c_p = p/(S_p^2),
sampled at x = s - S/2 - o/8.  (The reference's own FFT-based w-kernel
generator is a separate component; the gridder's cost and parity do not
depend on the tap values.)
"""
from __future__ import annotations

import math

import numpy as np

from .tables import OVER, GCFTable, VisChunks

C_LIGHT = 299792458.0


def gaussian_layers(S: int, sigma: float, chirp: float = 0.0,
                    dtype=np.complex128) -> np.ndarray:
    """taps[ou, ov, su, sv] = g(su - S/2 - ou/8) g(sv - S/2 - ov/8)
    * exp(i pi chirp (x^2 + y^2))."""
    o = np.arange(OVER)[:, None] / OVER
    s = np.arange(S)[None, :]
    x = s - S / 2 - o                                   # [ou, su]
    gx = np.exp(-x * x / (2 * sigma * sigma))
    px = np.exp(1j * math.pi * chirp * x * x)
    kx = gx * px                                        # [o, s]
    t = kx[:, None, :, None] * kx[None, :, None, :]     # [ou, ov, su, sv]
    return t.astype(dtype)


def wplane_supports(n_planes: int, s_min: int, s_max: int) -> list[int]:
    """Even supports growing linearly from s_min to s_max over planes."""
    if n_planes == 1:
        return [s_min]
    return [s_min + 2 * int(round((s_max - s_min) / 2 * p / (n_planes - 1)))
            for p in range(n_planes)]


def wplane_gcf(n_planes: int, s_min: int, s_max: int, half: bool = True,
               dtype=np.complex128) -> GCFTable:
    sup = wplane_supports(n_planes, s_min, s_max)
    planes = {}
    for p in range(n_planes):
        S = sup[p]
        planes[p] = gaussian_layers(S, S / 6.0, chirp=p / float(S * S), dtype=dtype)
        if not half and p > 0:
            planes[-p] = np.conj(planes[p])
    return GCFTable.from_planes(planes, half=half, dtype=dtype)


def plain_c1(seed: int = 1, n: int = 100_000, chunks: int = 100,
             grid_size: int = 512) -> tuple[VisChunks, GCFTable, int]:
    rng = np.random.default_rng(seed)
    ts = n // chunks
    uvw = np.zeros((chunks, ts, 3))
    uvw[..., 0] = rng.uniform(-240.0, 240.0, (chunks, ts))
    uvw[..., 1] = rng.uniform(-240.0, 240.0, (chunks, ts))
    vis = (rng.standard_normal((chunks, ts)) + 1j * rng.standard_normal((chunks, ts)))
    gcf = GCFTable.from_planes({0: gaussian_layers(8, 1.5)}, half=False)
    vc = VisChunks(uvw=uvw, vis=vis.astype(np.complex128),
                   supps=np.full(chunks, 8, np.int32), scale=1.0, wstep=1.0)
    return vc, gcf, grid_size


def antenna_layout(rng, n_ant: int = 256, sigma: float = 800.0,
                   radius: float = 3000.0) -> np.ndarray:
    """East/North antenna positions [n_ant, 2] in metres: 2-D Gaussian,
    redrawn until inside the disc of the given radius."""
    out = np.zeros((0, 2))
    while out.shape[0] < n_ant:
        p = rng.normal(0.0, sigma, (2 * n_ant, 2))
        p = p[np.hypot(p[:, 0], p[:, 1]) <= radius]
        out = np.concatenate([out, p])
    return out[:n_ant]


def earth_rotation_uvw(ant_en: np.ndarray, n_times: int, hours: float = 4.0,
                       dec_deg: float = -45.0, lat_deg: float = -30.7,
                       freq_hz: float = 1.4e9) -> np.ndarray:
    """uvw [n_baselines, n_times, 3] in wavelengths for every antenna pair
    (i < j), hour angle spanning +-hours/2, standard (X, Y, Z) rotation."""
    lam = C_LIGHT / freq_hz
    lat = math.radians(lat_deg)
    dec = math.radians(dec_deg)
    e, nn = ant_en[:, 0], ant_en[:, 1]
    X = -math.sin(lat) * nn
    Y = e
    Z = math.cos(lat) * nn
    i, j = np.triu_indices(ant_en.shape[0], k=1)
    bx, by, bz = X[j] - X[i], Y[j] - Y[i], Z[j] - Z[i]
    H = np.linspace(-hours / 2, hours / 2, n_times) * (2 * math.pi / 24.0)
    sH, cH = np.sin(H)[None, :], np.cos(H)[None, :]
    bx, by, bz = bx[:, None], by[:, None], bz[:, None]
    u = sH * bx + cH * by
    v = -math.sin(dec) * cH * bx + math.sin(dec) * sH * by + math.cos(dec) * bz
    w = math.cos(dec) * cH * bx - math.cos(dec) * sH * by + math.sin(dec) * bz
    return np.stack([u, v, w], axis=-1) / lam


def wproj(seed: int = 2, n_ant: int = 256, n_times: int = 306,
          grid_size: int = 4096, n_planes: int = 32, s_min: int = 8,
          s_max: int = 64, half: bool = True, vis_dtype=np.complex64,
          gcf_dtype=np.complex128) -> tuple[VisChunks, GCFTable, int]:
    """The w-projection workload (n_ant*(n_ant-1)/2 * n_times points)."""
    rng = np.random.default_rng(seed)
    ant = antenna_layout(rng, n_ant)
    uvw = earth_rotation_uvw(ant, n_times).reshape(-1, 3)
    maxuv = float(np.max(np.abs(uvw[:, :2])))
    scale = (grid_size / 2 - 1 - s_max / 2) / maxuv
    maxw = float(np.max(np.abs(uvw[:, 2])))
    wstep = maxw / (n_planes - 1) if n_planes > 1 else 1.0
    sup = wplane_supports(n_planes, s_min, s_max)
    plane = np.rint(np.abs(uvw[:, 2]) / wstep).astype(np.int64)
    plane = np.minimum(plane, n_planes - 1)
    supps = np.asarray(sup, np.int32)[plane]
    n = uvw.shape[0]
    vis = (rng.standard_normal(n, dtype=np.float32)
           + 1j * rng.standard_normal(n, dtype=np.float32)).astype(vis_dtype)
    gcf = wplane_gcf(n_planes, s_min, s_max, half=half, dtype=gcf_dtype)
    vc = VisChunks(uvw=uvw.reshape(n, 1, 3), vis=vis.reshape(n, 1),
                   supps=supps, scale=scale, wstep=wstep)
    return vc, gcf, grid_size


def snapshots(seed: int = 4, n_inst: int = 64, grid_size: int = 2048, n_planes: int = 16,
              s_min: int = 8, s_max: int = 48, median: float = 400_000.0,
              sigma: float = 1.0, clip: tuple[int, int] = (20_000, 5_000_000),
              n_ant: int = 256, minutes: float = 10.0, hours: float = 4.0,
              vis_dtype=np.complex64, gcf_dtype=np.complex128
              ) -> tuple[list[VisChunks], GCFTable, int]:
    """Uneven batch of snapshot instances of the 256-antenna array: instance
    k holds n_k points (lognormal, median `median`, log-sigma `sigma`,
    clipped to `clip`) of random baselines at random times within a window
    of `minutes` centred on an hour angle uniform over +-hours/2 (dec -45 deg,
    1.4 GHz).  All instances share the grid size, scale, w step and the GCF
    table (n_planes |w|-planes, supports s_min..s_max, half table)."""
    rng = np.random.default_rng(seed)
    ant = antenna_layout(rng, n_ant)
    lam = C_LIGHT / 1.4e9
    lat, dec = math.radians(-30.7), math.radians(-45.0)
    e, nn = ant[:, 0], ant[:, 1]
    X, Y, Z = -math.sin(lat) * nn, e, math.cos(lat) * nn
    i, j = np.triu_indices(n_ant, k=1)
    bx, by, bz = X[j] - X[i], Y[j] - Y[i], Z[j] - Z[i]
    bmax = float(np.max(np.sqrt(bx ** 2 + by ** 2 + bz ** 2))) / lam
    scale = (grid_size / 2 - 1 - s_max / 2) / bmax
    wstep = bmax / (n_planes - 1)
    sup = np.asarray(wplane_supports(n_planes, s_min, s_max), np.int32)
    counts = np.clip(np.rint(median * np.exp(sigma * rng.standard_normal(n_inst))),
                     clip[0], clip[1]).astype(np.int64)
    out = []
    for k in range(n_inst):
        n = int(counts[k])
        bl = rng.integers(0, bx.size, n)
        h0 = rng.uniform(-hours / 2, hours / 2)
        H = (h0 + rng.uniform(-0.5, 0.5, n) * minutes / 60.0) * (2 * math.pi / 24.0)
        sH, cH = np.sin(H), np.cos(H)
        x, y, z = bx[bl], by[bl], bz[bl]
        uvw = np.empty((n, 3))
        uvw[:, 0] = (sH * x + cH * y) / lam
        uvw[:, 1] = (-math.sin(dec) * cH * x + math.sin(dec) * sH * y + math.cos(dec) * z) / lam
        uvw[:, 2] = (math.cos(dec) * cH * x - math.cos(dec) * sH * y + math.sin(dec) * z) / lam
        plane = np.minimum(np.rint(np.abs(uvw[:, 2]) / wstep).astype(np.int64), n_planes - 1)
        vis = (rng.standard_normal(n, dtype=np.float32)
               + 1j * rng.standard_normal(n, dtype=np.float32)).astype(vis_dtype)
        out.append(VisChunks(uvw=uvw.reshape(n, 1, 3), vis=vis.reshape(n, 1),
                             supps=sup[plane], scale=scale, wstep=wstep))
    gcf = wplane_gcf(n_planes, s_min, s_max, half=True, dtype=gcf_dtype)
    return out, gcf, grid_size


def random_small(seed: int, grid_size: int = 64, chunks: int = 12,
                 ts_ch: int = 9, max_plane: int = 2, half: bool = False,
                 supports: tuple[int, ...] | None = None,
                 chunk_supp: str = "min", oob_frac: float = 0.1
                 ) -> tuple[VisChunks, GCFTable, int]:
    """Small randomised case.

    supports[k] is the support of |plane| k (defaults to 4 + 2k); chunk
    supports are the support of the chunk's minimum-|w| plane
    (chunk_supp='min', the reference caller's rule, which truncates larger-w
    points) or a random support not above it ('random').  A fraction of the
    points is pushed outside the grid.
    """
    rng = np.random.default_rng(seed)
    if supports is None:
        supports = tuple(4 + 2 * k for k in range(max_plane + 1))
    planes = {}
    for p in range(0, max_plane + 1):
        S = supports[p]
        t = (rng.standard_normal((OVER, OVER, S, S))
             + 1j * rng.standard_normal((OVER, OVER, S, S)))
        planes[p] = t
        if not half and p > 0:
            planes[-p] = (rng.standard_normal((OVER, OVER, S, S))
                          + 1j * rng.standard_normal((OVER, OVER, S, S)))
    gcf = GCFTable.from_planes(planes, half=half)
    wstep = 0.75
    lim = grid_size / 2 - max(supports) / 2 - 1
    uvw = np.zeros((chunks, ts_ch, 3))
    uvw[..., 0] = rng.uniform(-lim, lim, (chunks, ts_ch))
    uvw[..., 1] = rng.uniform(-lim, lim, (chunks, ts_ch))
    wmax = (max_plane + 0.49) * wstep
    uvw[..., 2] = rng.uniform(-wmax, wmax, (chunks, ts_ch))
    oob = rng.random((chunks, ts_ch)) < oob_frac
    uvw[..., 0] = np.where(oob, uvw[..., 0] + rng.choice([-1, 1], oob.shape) * grid_size,
                           uvw[..., 0])
    # exact integer and half-integer coordinates hit floor() edge cases
    edge = rng.random((chunks, ts_ch)) < 0.1
    uvw[..., 1] = np.where(edge, np.round(uvw[..., 1] * 2) / 2, uvw[..., 1])
    planes_pt = np.rint(np.abs(uvw[..., 2]) / wstep).astype(int)
    sup_arr = np.asarray(supports)
    if chunk_supp == "min":
        supps = sup_arr[planes_pt.min(axis=1)]
    else:
        cap = sup_arr[planes_pt.min(axis=1)]
        supps = np.array([rng.integers(1, c + 1) for c in cap])
    vis = rng.standard_normal((chunks, ts_ch)) + 1j * rng.standard_normal((chunks, ts_ch))
    vc = VisChunks(uvw=uvw, vis=vis, supps=supps.astype(np.int32), scale=1.0,
                   wstep=wstep)
    return vc, gcf, grid_size
