"""Writes the golden fixtures under tests/golden/ from the reference's own CPU
gridder (oracle/_ref/libms4ref.so, compiled from the reference sources by
oracle/Makefile), run with one OpenMP thread so the FP64 summation order is
the sequential one.

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py

Fixtures (inputs and expected outputs, .npz):
  pregrid_probe.npz   one point per call with vis = 1 and a GCF table whose
                      every tap encodes its layer and position, so the single
                      non-zero window the reference writes reveals its pregrid
                      (cell, GCF layer, recentring) for edge-case coordinates.
  random_<k>.npz      small randomised cases (full/half tables, truncated
                      supports, out-of-bound points): full grid, degridded
                      visibilities of that grid, tap counts.
  c1_summary.npz      the 512^2 / 100k-point plain configuration: tap count,
                      grid L2 norm and 4096 sampled cells; degridded
                      visibilities of the normalised grid, L2 norm and 4096
                      sampled points; sha256 of the generated inputs.
  small_kernels.npz   normalizeCPU, rotateCPU, grid0 and reweight on small
                      random inputs.
  callers.npz         gridKernelCPUFullGCF and deGridKernelCPUFullGCF called
                      with the tables and arguments of the reference's two
                      callers (tests/callers.py): cppcycle's and the Haskell
                      gridWrapper's (table pointer 32 entries past the centre).

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py callers   # one fixture
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from rc_amd import synth  # noqa: E402
from rc_amd.tables import OVER, GCFTable, VisChunks  # noqa: E402

sys.path.insert(0, str(HERE.parent))
import callers as K  # noqa: E402


def input_hash(vc: VisChunks, gcf: GCFTable) -> str:
    h = hashlib.sha256()
    for a in (vc.uvw, vc.vis, vc.supps, gcf.data, gcf.layer_off, gcf.supps):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(np.asarray([vc.scale, vc.wstep]).tobytes())
    return h.hexdigest()


def encoded_table(supps_by_plane: dict[int, int], half: bool) -> GCFTable:
    """tap(L, r, c) = (L_slot + 1) + i*(r*1000 + c)"""
    planes = {}
    for p, S in supps_by_plane.items():
        t = np.zeros((OVER, OVER, S, S), np.complex128)
        r = np.arange(S)[:, None]
        c = np.arange(S)[None, :]
        for ou in range(OVER):
            for ov in range(OVER):
                L = (p * OVER + ou) * OVER + ov
                t[ou, ov] = (L + 100000) + 1j * (r * 1000 + c)
        planes[p] = t
    return GCFTable.from_planes(planes, half=half)


def probe_points(rng, n: int, grid_size: int, scale: float, wstep: float,
                 max_plane: int) -> np.ndarray:
    pts = []
    lim = grid_size / 2 - 40
    for k in range(n):
        kind = k % 6
        u = rng.uniform(-lim, lim) / scale
        v = rng.uniform(-lim, lim) / scale
        w = rng.uniform(-(max_plane + 0.49), max_plane + 0.49) * wstep
        if kind == 1:      # pixel/(1/8) boundaries in scaled coordinates
            u = np.round(u * scale * 8) / 8 / scale
            v = np.round(v * scale * 8) / 8 / scale
        elif kind == 2:    # just below a boundary
            u = np.nextafter(np.round(u * scale * 8) / 8 / scale, -np.inf)
            v = np.nextafter(np.round(v * scale * 8) / 8 / scale, -np.inf)
        elif kind == 3:    # w on half-integer planes (round half away)
            w = (np.round(w / wstep) + 0.5 * np.sign(w or 1.0)) * wstep
            if abs(w / wstep) > max_plane:
                w = np.sign(w) * (max_plane - 0.5) * wstep
        elif kind == 4:    # negative zero and exact integers
            u = float(np.floor(u * scale)) / scale
            w = -0.0
        pts.append((u, v, w))
    return np.asarray(pts)


def make_probe(out: Path) -> None:
    rng = np.random.default_rng(101)
    cases = []
    for half in (False, True):
        for scale, wstep in ((1.0, 1.0), (0.37, 2.5), (3.1, 0.3)):
            N = 128
            max_plane = 3
            sup = {p: 6 + 2 * abs(p) for p in range(-max_plane if not half else 0, max_plane + 1)}
            gcf = encoded_table(sup, half)
            pts = probe_points(rng, 120, N, scale, wstep, max_plane)
            # chunk support: the minimum plane support (no over-read)
            S_chunk = rng.choice([3, 4, 5, 6], len(pts))
            # every sixth point sits on an edge of the in-bound range of its
            # support: top-left pixel N - S - 1, N - S, 0 or -1 (u, then v)
            for k in range(5, len(pts), 6):
                S = int(S_chunk[k])
                target = (N - S - 1, N - S, 0, -1)[(k // 6) % 4]
                pix = target - N // 2 + S // 2 + 0.3
                if (k // 24) % 2 == 0:
                    pts[k, 0] = pix / scale
                else:
                    pts[k, 1] = pix / scale
            res = []
            for (u, v, w), S in zip(pts, S_chunk):
                vc = VisChunks(uvw=np.array([[[u, v, w]]]), vis=np.ones((1, 1), np.complex128),
                               supps=np.array([S], np.int32), scale=scale, wstep=wstep)
                g, taps = O.grid(vc, gcf, N, impl="ref")
                nz = np.argwhere(g != 0)
                if len(nz) == 0:
                    res.append((-1, -1, 0.0, 0.0, taps))
                else:
                    u0, v0 = nz.min(axis=0)
                    val = g[u0, v0]
                    res.append((u0, v0, val.real, val.imag, taps))
            cases.append(dict(half=half, scale=scale, wstep=wstep, N=N, pts=pts,
                              S=np.asarray(S_chunk, np.int32),
                              res=np.asarray(res, np.float64)))
    np.savez_compressed(out, n_cases=len(cases), **{
        f"c{i}_{k}": v for i, c in enumerate(cases) for k, v in c.items()})


def random_cases():
    """(name, kwargs) of the small randomised fixtures."""
    return [
        ("random_0", dict(seed=10, half=False, chunk_supp="min")),
        ("random_1", dict(seed=11, half=True, chunk_supp="min")),
        ("random_2", dict(seed=12, half=False, chunk_supp="random",
                          supports=(5, 7, 9))),
        ("random_3", dict(seed=13, half=True, chunk_supp="random", grid_size=48,
                          supports=(3, 6, 11))),
    ]


def make_random(out_dir: Path) -> None:
    for name, kw in random_cases():
        vc, gcf, N = synth.random_small(**kw)
        g, taps = O.grid(vc, gcf, N, impl="ref")
        vis, dtaps = O.degrid(vc, gcf, g, impl="ref")
        np.savez_compressed(out_dir / f"{name}.npz", uvw=vc.uvw, vis_in=vc.vis,
                            supps=vc.supps, scale=vc.scale, wstep=vc.wstep,
                            N=N, half=gcf.half, gcf_data=gcf.data,
                            gcf_off=gcf.layer_off, gcf_len=gcf.layer_len,
                            gcf_center=gcf.center, gcf_supps=gcf.supps,
                            gcf_M=gcf.max_w_plane, grid=g, taps=taps,
                            vis_out=vis, dtaps=dtaps)


def make_c1(out: Path) -> None:
    vc, gcf, N = synth.plain_c1()
    g, taps = O.grid(vc, gcf, N, impl="ref")
    rng = np.random.default_rng(7)
    cells = rng.integers(0, N * N, 4096)
    ref = O.ref_lib()
    gn = O.aligned_zeros(g.shape, np.complex128)
    gn[...] = g
    ref.normalizeCPU(gn.ctypes.data, N, N)
    vis, dtaps = O.degrid(vc, gcf, gn, impl="ref")
    pts = rng.integers(0, vc.n, 4096)
    np.savez_compressed(out, input_sha256=input_hash(vc, gcf), taps=taps,
                        grid_l2=np.linalg.norm(g), cells=cells,
                        cell_vals=g.reshape(-1)[cells], dtaps=dtaps,
                        vis_l2=np.linalg.norm(vis), pts=pts,
                        pt_vals=vis.reshape(-1)[pts], N=N)


def make_small(out: Path) -> None:
    rng = np.random.default_rng(5)
    ref = O.ref_lib()
    N, P, n = 64, 70, 500
    src = rng.standard_normal((N, P)) + 1j * rng.standard_normal((N, P))
    norm = src.copy()
    ref.normalizeCPU(norm.ctypes.data, P, N)
    uvw = np.zeros((n, 3))
    uvw[:, 0] = rng.uniform(-40, 40, n)
    uvw[:, 1] = rng.uniform(-40, 40, n)
    uvw[:, 2] = rng.uniform(-900, 900, n)
    uvw[:20, :2] = np.round(uvw[:20, :2] * 2) / 2  # round-half cases
    vis = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    scale = 0.77
    rot = vis.copy()
    ref.rotateCPU(uvw.ctypes.data, rot.ctypes.data, n, scale)
    g0 = np.zeros((N, P), np.complex128)
    centre = g0.ctypes.data + (N // 2 * (P + 1)) * 16
    ref.grid0(uvw.ctypes.data, vis.ctypes.data, C.c_void_p(centre), scale, n, P, N)
    rw = vis.copy()
    ref.reweight(uvw.ctypes.data, rw.ctypes.data, scale, n, N)
    np.savez_compressed(out, N=N, P=P, src=src, norm=norm, uvw=uvw, vis=vis,
                        scale=scale, rot=rot, grid0=g0, reweight=rw)


def caller_cases():
    """The calls of callers.npz: (name, Call) with inputs generated from
    fixed seeds (the fixture keeps inputs and outputs)."""
    # cppcycle: its constants (growth 16, min 3, max 128, pad 2, wstep 1e4),
    # a 128^2 grid and its scale rule (most far points fall outside)
    wstep = 10000.0
    supp = K.support_rule(3, 16, 128)
    uvw, vis = K.baselines(21, 24, 16, wstep, 2.4, lim=1.0, far=22.0)
    lay = K.make_layers(K.cppcycle_max_plane(uvw, wstep), supp, np.random.default_rng(31))
    cpp = K.cppcycle_call(uvw, vis, wstep, supp, 128, 2, lay)
    # gridWrapper: GCFPar min 8, growth 8, max 40, a square 96^2 grid
    wstep = 2.0
    supp = K.support_rule(8, 8, 40)
    uvw, vis = K.baselines(22, 24, 16, wstep, 2.4, lim=96 / 2 / 1.3 - 22)
    lay = K.make_layers(K.haskell_max_plane(uvw, wstep), supp, np.random.default_rng(32))
    hs = K.wrapper_call(uvw, vis, wstep, supp, 96, 96, 1.3, lay)
    return [("cpp", cpp), ("hs", hs)]


def make_callers(out: Path) -> None:
    ref = O.ref_lib()
    arrs = {}
    for name, call in caller_cases():
        g, taps = call.grid(ref.gridKernelCPUFullGCF)
        v, dtaps = call.degrid(ref.deGridKernelCPUFullGCF, g)
        arrs.update({f"{name}_uvw": call.uvw, f"{name}_vis": call.vis,
                     f"{name}_chunk_supps": call.chunk_supps, f"{name}_grid": g,
                     f"{name}_taps": taps, f"{name}_vis_out": v, f"{name}_dtaps": dtaps,
                     f"{name}_layers_sha256": hashlib.sha256(call.layers.data.tobytes()).hexdigest()})
    np.savez_compressed(out, **arrs)


def main(argv: list[str]) -> int:
    if os.environ.get("OMP_NUM_THREADS") != "1":
        print("run with OMP_NUM_THREADS=1 (sequential reference summation order)")
        return 2
    if not O.ref_available():
        print("oracle/_ref/libms4ref.so missing: make -C oracle ref")
        return 2
    makers = {"pregrid_probe": lambda: make_probe(HERE / "pregrid_probe.npz"),
              "random": lambda: make_random(HERE),
              "c1_summary": lambda: make_c1(HERE / "c1_summary.npz"),
              "small_kernels": lambda: make_small(HERE / "small_kernels.npz"),
              "callers": lambda: make_callers(HERE / "callers.npz")}
    for name in (argv or list(makers)):
        makers[name]()
    for p in sorted(HERE.glob("*.npz")):
        print(f"{p.name}: {p.stat().st_size} bytes")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
