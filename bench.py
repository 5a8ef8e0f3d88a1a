"""Benchmark: w-projection gridding + degridding on one MI355X.

Workload: 4096 x 4096 complex64 grid, 32 |w|-planes (supports 8..64, 8x
oversampling, half GCF table), 9,987,840 points from 256 antennas tracked
over 4 h at declination -45 deg, 1.4 GHz, complex64 visibilities N(0,1).
All inputs are resident in HBM before timing starts.

One step = pregrid (FP64) + uv-tile binning + FP32 gridding + FP32
degridding of every point (rc_amd tiled path).  value = points / step time
(Mvis/s gridded+degridded).  The FP64 plain path (pregrid + FP64-atomic
gridder + FP64 degridder, one launch each) is timed the same way as the
comparison leg, and the FP32 results are compared with it (the FP64 path is
within 1e-12 of the reference's CPU gridder).

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD published
# MI355X FP32 vector peak without packed FP32 instructions (which this
# library does not use): 256 CUs x 64 lanes x 2 flops (FMA) x 2.4 GHz.  AMD's
# 157.3 TFLOP/s counts packed FP32.
FP32_VALU_PEAK_TFLOPS = 78.6
CLOCK_GHZ = 2.4                # MI355X peak engine clock
N_CU = 256


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tile", type=int, default=32)
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--waves", type=int, default=0,
                    help="wavefronts per gridding workgroup, 1..4 (0: the plan's default)")
    ap.add_argument("--degrid-waves", type=int, default=0,
                    help="wavefronts per degridding workgroup, 2, 4 or 8 (0: the default 4)")
    ap.add_argument("--layout", choices=["pairs", "halo", "halo16"], default="pairs",
                    help="record layout of the tiled plan (see rcg_tiled_plan_create_ex)")
    ap.add_argument("--n-times", type=int, default=306)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--workload", choices=["wproj", "batched", "imaging", "precision", "binner",
                                           "flow", "small", "next", "ms4", "tune"],
                    default="wproj",
                    help="wproj: the 4096^2 w-projection workload (default); batched: 64 "
                         "uneven 2048^2 snapshot instances, one graph launch vs one launch "
                         "per instance; imaging: the differentiable measurement operator "
                         "on a 2048^2 sky, 20 gradient steps and 1000 Hogbom iterations; "
                         "precision: 50M points of the w-projection layout, FP32 with and "
                         "without compensated flushes against the FP64 path; binner: the "
                         "MS4 uv binner (binm) through its host-memory C ABI; flow: the "
                         "MS6 binSizer + binner kernels; small: normalize, rotate, grid0, "
                         "reweight, hermitian fold and FFT against their rooflines; next: "
                         "GCF generation, MS6 scatter/degrid and nvGridder; ms4: the MS4 "
                         "host-memory grid/degrid calls on 32640 chunks of 306 and the "
                         "device FP64 gridders; tune: the autotuner's (tile, chunk, waves) "
                         "sweep on the wproj workload")
    ap.add_argument("--instances", type=int, default=64)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled(kernel: str) -> dict:
    """Counter figures per dispatch of `kernel` at the default headline
    workload, from the rocprofv3 summary committed under profiles/ (lines
    '<key> <kernel> <value>', the kernel named without spaces, e.g.
    k_grid_tiles<32,3>: traffic_bytes = L2-to-memory bytes from the
    per-request-size counters, valu_instr, salu_instr)."""
    p = ROOT / "profiles" / "wproj_tiled_traffic.md"
    out = {}
    if p.exists():
        for line in p.read_text().splitlines():
            f = line.split()
            if len(f) == 3 and f[1] == kernel and f[0] in ("traffic_bytes", "valu_instr",
                                                            "salu_instr"):
                out[f[0]] = int(f[2])
    return out


def cpu_baseline(vc, gcf, N, budget_s: float) -> dict:
    """The reference's own CPU gridder + degridder (compiled from its sources
    with its release flags, oracle/_ref/libms4ref_fast.so), else the oracle
    restatement, on a prefix sample of the same workload; OpenMP threads =
    the host cores available to this process, at most 16: the share of the
    host that one GPU is given on the benchmark machines (which set
    OMP_NUM_THREADS=16 for it), so the baseline uses what a one-GPU job has."""
    from oracle import oracle as O
    from rc_amd.tables import VisChunks
    avail = len(os.sched_getaffinity(0))
    threads = min(16, avail)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    kind = "reference" if O.ref_available(fast=True) else "port"
    impl = "ref_fast" if kind == "reference" else "oracle"
    gcf128 = gcf.astype(np.complex128)

    def sample(m):
        return VisChunks(uvw=vc.uvw[:m], vis=vc.vis[:m].astype(np.complex128),
                         supps=vc.supps[:m], scale=vc.scale, wstep=vc.wstep)

    def run(m):
        s = sample(m)
        t0 = time.perf_counter()
        g, taps = O.grid(s, gcf128, N, impl=impl, nthreads=threads)
        t1 = time.perf_counter()
        O.degrid(s, gcf128, g, impl=impl, nthreads=threads)
        t2 = time.perf_counter()
        return t2 - t0, t1 - t0, t2 - t1, taps

    m = 100_000
    dt, _, _, _ = run(m)
    m2 = int(min(vc.n, max(m, m * budget_s / max(dt, 1e-3))))
    dt, tg, td, taps = run(m2)
    return {"value": round(m2 / dt / 1e6, 4), "unit": "Mvis/s", "cores": threads,
            "kind": kind, "host_cores": os.cpu_count(), "cores_available": avail,
            "cores_note": "threads capped at 16, the host share of one GPU on these machines",
            "sample": f"first {m2} points of the workload (taps {taps}), full 4096^2 "
                      f"complex128 grid, grid {tg:.2f} s + degrid {td:.2f} s, "
                      f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}"}


def bench_batched(args) -> int:
    """64 snapshot instances (2048^2 grids, 16 w-planes, supports 8..48,
    lognormal point counts, median 400k, clipped to 20k..5M): tile, chunk and
    gridding waves per workgroup from the autotuner, then bin + grid + degrid of all instances as one
    captured graph (one launch per kernel over all instances) against one
    plan and one launch sequence per instance, at the faster of the batch's
    wavefronts per workgroup and the plan defaults."""
    import torch
    from rc_amd import autotune
    from rc_amd import device as D
    from rc_amd import synth

    dev = torch.device("cuda:0")
    t0 = time.time()
    vcs, gcf, N = synth.snapshots(n_inst=args.instances)
    n = sum(v.n for v in vcs)
    log(f"batched workload: {len(vcs)} instances, {n} points, {time.time() - t0:.1f} s")
    geoms = [D.DeviceGeom.from_chunks(v, N) for v in vcs]
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    vis = torch.from_numpy(np.concatenate([v.vis.reshape(-1) for v in vcs])
                           .astype(np.complex64)).to(dev)
    tune = autotune.autotune(geoms, dg, vis, max_support=48, log=log)
    tile, chunk, waves, dwaves = tune["tile"], tune["chunk"], tune["waves"], tune["degrid_waves"]
    log(f"autotune: tile {tile} chunk {chunk} waves {waves} degrid waves {dwaves} "
        f"({tune['tune_s']:.1f} s)")
    plan = D.TiledPlan(geoms, tile=tile, chunk=chunk, max_support=48, waves=waves,
                       degrid_waves=dwaves)
    grid = torch.empty(plan.grid_shape, dtype=torch.complex64, device=dev)
    out = torch.empty(n, dtype=torch.complex64, device=dev)
    graph = plan.capture(dg, D.STAGE_BIN | D.STAGE_GRID | D.STAGE_DEGRID, vis=vis, grid=grid,
                         vis_out=out)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps * 1e3

    ms_batch = timed(graph.launch, args.steps, args.warmup)
    st = plan.stats()
    if st["err"]:
        print(json.dumps({"error": f"binning error word {st['err']}"}))
        return 1
    # one plan and one launch sequence per instance, same tile and chunk.  The
    # wavefronts per workgroup are timed both as tuned for the batch and at the
    # plan's defaults (more wavefronts share each work item, which suits the
    # small instances); the faster is the baseline.
    offs = np.cumsum([0] + [g.n for g in geoms])

    def per_instance(plans):
        def run():
            for k, p in enumerate(plans):
                v, o = vis[offs[k]:offs[k + 1]], out[offs[k]:offs[k + 1]]
                p.bin(dg)
                p.grid(dg, v, grid[k])
                p.degrid(dg, grid[k], o)
        return run

    single, cands = {}, {}
    for label, w, dw in (("batch_tuned", waves, dwaves), ("plan_default", 0, 0)):
        cands[label] = [D.TiledPlan(g, tile=tile, chunk=chunk, max_support=48, waves=w,
                                    degrid_waves=dw) for g in geoms]
        single[label] = timed(per_instance(cands[label]), max(2, args.steps // 2), 1)
    single_cfg = min(single, key=single.get)
    for label, ps in cands.items():
        if label != single_cfg:
            for p in ps:
                p.close()
    plans, ms_single = cands[single_cfg], single[single_cfg]
    graphs = [p.capture(dg, D.STAGE_BIN | D.STAGE_GRID | D.STAGE_DEGRID,
                        vis=vis[offs[k]:offs[k + 1]], grid=grid[k],
                        vis_out=out[offs[k]:offs[k + 1]]) for k, p in enumerate(plans)]
    ms_single_graph = timed(lambda: [g.launch() for g in graphs], max(2, args.steps // 2), 1)
    for g in graphs:
        g.close()
    for p in plans:
        p.close()
    # errors against the FP64 plain path, instance by instance
    graph.launch()
    dg64 = D.DeviceGCF.from_table(gcf, dtype=np.complex128)
    num_g = den_g = num_v = den_v = 0.0
    g32in = grid.clone()
    plan.degrid(dg, g32in, out)
    for k, g in enumerate(geoms):
        pg, _, _ = D.pregrid(g, dg64)
        v64 = vis[offs[k]:offs[k + 1]].to(torch.complex128)
        g64 = D.grid_plain_f64(g, dg64, pg, v64)
        o64 = D.degrid_plain_f64(g, dg64, pg, g32in[k].to(torch.complex128))
        num_g += float(torch.linalg.vector_norm(grid[k].to(torch.complex128) - g64)) ** 2
        den_g += float(torch.linalg.vector_norm(g64)) ** 2
        num_v += float(torch.linalg.vector_norm(out[offs[k]:offs[k + 1]].to(torch.complex128)
                                                - o64)) ** 2
        den_v += float(torch.linalg.vector_norm(o64)) ** 2
    torch.cuda.synchronize()
    g_err, d_err = (num_g / den_g) ** 0.5, (num_v / den_v) ** 0.5
    value = n / (ms_batch / 1e3) / 1e6
    log(f"batched graph {ms_batch:.3f} ms/step -> {value:.1f} Mvis/s; per instance "
        f"{ms_single:.3f} ms (graphs {ms_single_graph:.3f} ms); errors {g_err:.2e}/{d_err:.2e}")
    line = {
        "metric": "Mvis/s gridded+degridded, 64 uneven 2048² instances, 16 w-planes: one "
                  "graph launch vs one launch per instance",
        "value": round(value, 3), "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_batch, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64/f32 (f64 pregrid)",
        "data": "synthetic",
        "config": {"workload": f"{len(vcs)} snapshot instances, 2048x2048, 16 w-planes, "
                               "supports 8-48, lognormal counts (median 400k, sigma 1, "
                               "20k-5M)",
                   "points": n, "taps_per_direction": st["taps"], "tile": tile,
                   "chunk": chunk, "waves": waves, "degrid_waves": dwaves},
        "per_instance_ms_per_step": round(ms_single, 4),
        "per_instance_waves": {k: round(v, 4) for k, v in single.items()},
        "per_instance_baseline": single_cfg,
        "per_instance_graphs_ms_per_step": round(ms_single_graph, 4),
        "speedup_vs_per_instance": round(ms_single / ms_batch, 3),
        "error_vs_fp64": {"grid_rel_l2": g_err, "degrid_rel_l2": d_err},
        "autotune": [{k: r.get(k) for k in ("tile", "chunk", "waves", "ms_step", "ms_grid",
                                            "ms_degrid", "skipped")}
                     for r in tune["table"]],
        "autotune_degrid_waves": [{k: r.get(k) for k in ("degrid_waves", "ms_degrid")}
                                  for r in tune["degrid_table"]],
    }
    print(json.dumps(line), flush=True)
    graph.close()
    plan.close()
    return 0


def bench_imaging(args) -> int:
    """Measurement operator A = degrid o FFT on a 2048^2 FP32 sky of 200
    point sources (uniform positions in the inner half, lognormal fluxes
    mu 0 sigma 1) plus N(0, 1e-3) noise, 4,993,920 points of the 256-antenna
    array (16 w-planes, supports 8..48): forward and adjoint rates, 20
    gradient-descent steps on chi^2 through the autograd operator, then 1000
    Hogbom iterations (gain 0.1) on the dirty image with its PSF."""
    import torch
    from rc_amd import device as D
    from rc_amd import operator as M
    from rc_amd import synth

    dev = torch.device("cuda:0")
    N = 2048
    vc, gcf, _ = synth.wproj(seed=5, n_times=153, grid_size=N, n_planes=16, s_min=8, s_max=48)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    op = M.MeasurementOp(geom, dg, tile=32, chunk=1024, max_support=48)
    rng = np.random.default_rng(11)
    sky = np.zeros((N, N), np.float32)
    pos = rng.integers(N // 4, 3 * N // 4, (200, 2))
    np.add.at(sky, (pos[:, 0], pos[:, 1]), rng.lognormal(0.0, 1.0, 200).astype(np.float32))
    sky += rng.normal(0.0, 1e-3, (N, N)).astype(np.float32)
    x_true = torch.from_numpy(sky).to(dev)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps * 1e3, out

    ms_fwd, d = timed(lambda: op.forward(x_true), args.steps)
    ms_adj, _ = timed(lambda: op.adjoint(d), args.steps)
    # step 1/(2 L) with L from a power iteration on Re(A^H A)
    v = torch.from_numpy(rng.standard_normal((N, N)).astype(np.float32)).to(dev)
    lip = 1.0
    for _ in range(10):
        v = op.adjoint(op.forward(v)).real
        lip = float(v.norm())
        v = v / lip
    step = 1.0 / (2.0 * lip * 1.05)
    x = torch.zeros((N, N), device=dev, requires_grad=True)
    chi0 = float(M.chi2(op, x, d).detach())
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(20):
        loss = M.chi2(op, x, d)
        x.grad = None
        loss.backward()
        with torch.no_grad():
            x -= step * x.grad
    torch.cuda.synchronize()
    ms_steps = (time.perf_counter() - t) * 1e3
    chi20 = float(M.chi2(op, x, d).detach())
    # dirty image and PSF (vis = 1), normalised by the PSF peak
    psf = op.grid_image(torch.ones_like(d)).real.contiguous()
    peak = float(psf.abs().max())
    psf /= peak
    dirty = (op.grid_image(d).real / peak).contiguous()
    res = dirty.clone()
    torch.cuda.synchronize()
    t = time.perf_counter()
    model, iters = D.hogbom(res, psf, 1000, 0.1, 0.0)
    torch.cuda.synchronize()
    ms_clean = (time.perf_counter() - t) * 1e3
    value = 2 * op.n / ((ms_fwd + ms_adj) / 1e3) / 1e6
    log(f"imaging: forward {ms_fwd:.3f} ms adjoint {ms_adj:.3f} ms; 20 steps {ms_steps:.1f} ms "
        f"chi2 {chi0:.4g} -> {chi20:.4g}; hogbom {iters} iterations {ms_clean:.1f} ms, "
        f"residual peak {float(res.abs().max()):.4g} of {float(dirty.abs().max()):.4g}")
    line = {
        "metric": "Mvis/s through the measurement operator (forward degrid+FFT and adjoint "
                  "grid+FFT), 2048² 16 w-planes",
        "value": round(value, 3), "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(ms_fwd + ms_adj, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64/f32 (f64 pregrid)",
        "data": "synthetic",
        "config": {"workload": "200 point sources on a 2048x2048 FP32 sky + N(0,1e-3) noise, "
                               "16 w-planes, supports 8-48",
                   "points": op.n, "taps_per_direction": op.taps},
        "forward_ms": round(ms_fwd, 4), "adjoint_ms": round(ms_adj, 4),
        "gradient_steps": {"steps": 20, "ms_total": round(ms_steps, 2), "chi2_start": chi0,
                           "chi2_end": chi20},
        "hogbom": {"iterations": iters, "gain": 0.1, "ms_total": round(ms_clean, 2),
                   "residual_peak": float(res.abs().max()),
                   "dirty_peak": float(dirty.abs().max())},
    }
    print(json.dumps(line), flush=True)
    op.close()
    return 0


def bench_precision(args) -> int:
    """The w-projection layout at 1530 time samples (49,939,200 points):
    FP32 gridding with plain and with compensated (TwoSum) tile flushes, and
    FP32 degridding, against the FP64 GPU path on the same complex64 inputs
    (the FP64 path itself is within 1e-12 of the reference's CPU gridder).
    Reports relative L2 and max-abs errors, kernel times, algorithmic HBM
    GB/s and FP32 VALU utilisation."""
    import torch
    from rc_amd import device as D
    from rc_amd import synth

    dev = torch.device("cuda:0")
    t0 = time.time()
    vc, gcf, N = synth.wproj(n_times=args.n_times if args.n_times != 306 else 1530)
    log(f"precision workload: {vc.n} points, {time.time() - t0:.1f} s to generate")
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg32 = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    dg64 = D.DeviceGCF.from_table(gcf, dtype=np.complex128)
    vis32 = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).to(dev)
    plan = D.TiledPlan(geom, tile=args.tile, chunk=args.chunk, max_support=64)
    plan.bin(dg32)
    taps = plan.stats()["taps"]
    # FP64 references on the same inputs
    pg, _, _ = D.pregrid(geom, dg64)
    g64 = D.grid_plain_f64(geom, dg64, pg, vis32.to(torch.complex128))
    torch.cuda.synchronize()

    def errs(a, b):
        d = (a.to(torch.complex128) - b)
        return {"rel_l2": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(b)),
                "max_abs": float(d.abs().max()), "max_ref": float(b.abs().max())}

    def kernel_ms(which, fn, reps=5):
        fn()
        plan.timing(True)
        for _ in range(reps):
            out = fn()
        t = float(np.mean(plan.kernel_times(which)))
        plan.timing(False)
        return t, out

    b_dir = vc.n * (24 + 8) + 8 * taps + 8 * N * N
    res = {}
    for comp in (False, True):
        plan.compensated(comp)
        tk, g32 = kernel_ms("grid", lambda: plan.grid(dg32, vis32))
        torch.cuda.synchronize()
        res["grid_compensated" if comp else "grid_plain"] = {
            **errs(g32, g64), "kernel_ms": round(tk, 4),
            "hbm_gbs_algorithmic": round(b_dir / (tk / 1e3) / 1e9, 1),
            "valu_frac": round(8 * taps / (tk / 1e3) / 1e12 / FP32_VALU_PEAK_TFLOPS, 4)}
        del g32
    plan.compensated(False)
    g_in = g64.to(torch.complex64)
    td, v32 = kernel_ms("degrid", lambda: plan.degrid(dg32, g_in))
    v64 = D.degrid_plain_f64(geom, dg64, pg, g_in.to(torch.complex128))
    res["degrid"] = {**errs(v32, v64), "kernel_ms": round(td, 4),
                     "hbm_gbs_algorithmic": round(b_dir / (td / 1e3) / 1e9, 1),
                     "valu_frac": round(8 * taps / (td / 1e3) / 1e12 / FP32_VALU_PEAK_TFLOPS, 4)}
    torch.cuda.synchronize()
    log(json.dumps(res))
    gp, gc = res["grid_plain"], res["grid_compensated"]
    value = vc.n / ((gp["kernel_ms"] + td) / 1e3) / 1e6
    line = {
        "metric": "precision study: FP32 grid/degrid error vs FP64 at 50M points, 4096² 32 "
                  "w-planes (Mvis/s of the plain-flush kernels)",
        "value": round(value, 3), "unit": "Mvis/s", "n_gpus": 1, "steps": 5, "warmup": 1,
        "ms_per_step": round(gp["kernel_ms"] + td, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64/f32 (f64 pregrid)",
        "data": "synthetic",
        "config": {"workload": "w-projection layout, 256 antennas x 1530 times, 4096x4096, "
                               "32 w-planes, supports 8-64", "points": vc.n,
                   "taps_per_direction": taps, "tile": args.tile, "chunk": args.chunk},
        "errors_vs_fp64_gpu": res,
        "compensation_gain": round(gp["rel_l2"] / max(gc["rel_l2"], 1e-30), 3),
    }
    print(json.dumps(line), flush=True)
    plan.close()
    return 0


def bench_binner(args) -> int:
    """The MS4 uv binner (binm: 32 x 32 bins, margin copies, half GCF
    indices) through its host-memory C ABI on the points of the w-projection
    workload, moved to non-negative cells (the reference bins the
    untranslated cell, MS4/kernel/cpu/binner/binner.cpp:133-139).  One step =
    one binm call: upload of uvw and amplitudes, the device passes, and the
    copy of the binned block into pinned host memory.  value = points per
    second through the call (PCIe included).  The CPU baseline is the
    oracle's single-threaded restatement on the same points."""
    import ctypes as C
    from rc_amd import _lib, synth

    vc, gcf, N = synth.wproj(n_times=args.n_times)
    uvw = np.ascontiguousarray(vc.uvw.reshape(-1, 3), np.float64)
    uvw[:, :2] += (N / 2) / vc.scale
    vis = np.ascontiguousarray(vc.vis.reshape(-1), np.complex128)
    sup = np.ascontiguousarray(gcf.supps, np.int32)
    M = gcf.max_w_plane
    n = len(vis)
    h = _lib.lib()
    centre = sup.ctypes.data + 4 * M

    def call():
        p = h.binm(n, vc.scale, vc.wstep, vis.ctypes.data, uvw.ctypes.data, centre, N)
        if not p:
            raise _lib.RcgError(f"binm: {_lib.last_error()[1]}")
        size = C.c_int.from_address(p - 12320).value
        h.rcg_bin_free(p)
        return size

    for _ in range(args.warmup):
        size = call()
    ts = []
    for _ in range(args.steps):
        t = time.perf_counter()
        call()
        ts.append(time.perf_counter() - t)
    ms = float(np.median(ts)) * 1e3
    copies = (size - 12320) // 24
    line = {
        "metric": "Mpoints/s binned by binm (32x32 uv bins with margin copies), host memory in "
                  "and out",
        "value": round(n / (ms / 1e3) / 1e6, 3), "unit": "Mpoints/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (int16 records)", "data": "synthetic",
        "config": {"workload": "w-projection points (256 antennas x 306 times) on 4096x4096, "
                               "32 w-planes, supports 8-64, cells moved to [0, 4096)",
                   "points": n, "copies": int(copies), "block_bytes": int(size)},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        t = time.perf_counter()
        O.bin_points(uvw, vis, vc.scale, vc.wstep, sup, M, N, True)
        cpu_s = time.perf_counter() - t
        line["cpu_baseline"] = {"value": round(n / cpu_s / 1e6, 3), "unit": "Mpoints/s",
                                "cores": 1, "kind": "port",
                                "sample": f"all {n} points, oracle orc_bin (C, one thread), "
                                          f"{cpu_s:.2f} s including the copy out"}
    print(json.dumps(line), flush=True)
    return 0


def bench_ms4(args) -> int:
    """The MS4 host-memory entry points (gridKernelCPUHalfGCF +
    deGridKernelCPUHalfGCF, MS4/kernel/cpu/scatter_grid/
    scatter_gridder_w_dependent_dyn_1p.cpp:245-266, 368-431) on the
    w-projection points in the reference's chunk layout: 32640 baselines x
    306 samples, each chunk's support that of its smallest-|w| plane,
    complex128 throughout.  One step = one grid call + one degrid call of the
    gridded result, host memory in and out (PCIe copies included).  Also
    reported: the device-resident FP64 gridders and degridders on the same
    points, timed with HIP events on the stream they run on:
    rcg_grid_tiled_f64 / rcg_degrid_tiled_f64 (binned pairs, what the calls
    run) and the plain FP64 kernels."""
    import torch
    from rc_amd import device as D
    from rc_amd import ms4, synth
    from rc_amd.tables import VisChunks

    vc, gcf, N = synth.wproj(n_times=args.n_times)
    ts = args.n_times
    nb = vc.n // ts
    uvw = vc.uvw.reshape(nb, ts, 3)
    vis = vc.vis.reshape(nb, ts).astype(np.complex128)
    sup = vc.supps.reshape(nb, ts).min(axis=1).astype(np.int32)
    v2 = VisChunks(uvw=uvw, vis=vis, supps=sup, scale=vc.scale, wstep=vc.wstep)
    g128 = gcf.astype(np.complex128)
    n = vc.n

    # device-resident FP64 gridders
    stream = torch.cuda.current_stream()
    geom = D.DeviceGeom.from_chunks(v2, N)
    dg = D.DeviceGCF.from_table(g128)
    dvis = torch.from_numpy(vis.reshape(-1)).cuda()
    pg, taps_d, _ = D.pregrid(geom, dg)
    plan = D.TiledPlan(geom, tile=32, chunk=1024)
    plan.bin(dg)
    gt = plan.grid_f64(dg, dvis)
    gp = D.grid_plain_f64(geom, dg, pg, dvis)
    torch.cuda.synchronize()

    def timed(fn, reps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in ev]))

    ms_tiled = timed(lambda: plan.grid_f64(dg, dvis, gt), max(3, args.steps))
    ms_plain = timed(lambda: D.grid_plain_f64(geom, dg, pg, dvis, gp), 3)
    rel = float((gt - gp).abs().pow(2).sum().sqrt() / gp.abs().pow(2).sum().sqrt())
    # the FP64 degridders on the gridded result
    vt = plan.degrid_f64(dg, gp)
    vpl = D.degrid_plain_f64(geom, dg, pg, gp)
    ms_dtiled = timed(lambda: plan.degrid_f64(dg, gp, vt), max(3, args.steps))
    ms_dplain = timed(lambda: D.degrid_plain_f64(geom, dg, pg, gp, vpl), 3)
    drel = float((vt - vpl).abs().pow(2).sum().sqrt() / vpl.abs().pow(2).sum().sqrt())
    taps = int(taps_d.item())
    plan.close()
    del gt, gp, pg, dvis, dg, geom, vt, vpl
    torch.cuda.empty_cache()

    # the host-memory calls
    # output buffers allocated (and touched) before the calls, as the
    # reference's driver passes a grid vector it has already built
    # (MS4/misc/cppcycle.cpp:208-211)
    g_out = np.zeros((N, N), np.complex128)
    v_out = np.zeros((nb, ts), np.complex128)

    def step():
        t0 = time.perf_counter()
        g, tg = ms4.grid(v2, g128, N, out=g_out)
        t1 = time.perf_counter()
        g /= N * N
        t2 = time.perf_counter()
        _, td = ms4.degrid(v2, g128, g, out=v_out)
        t3 = time.perf_counter()
        assert tg == td == taps
        return t1 - t0, t3 - t2

    for _ in range(max(1, args.warmup)):
        step()
    ts_g, ts_d = zip(*[step() for _ in range(args.steps)])
    mg, md = float(np.median(ts_g)) * 1e3, float(np.median(ts_d)) * 1e3
    ms = mg + md
    pcie = (24 + 16) * n * 2 + 16 * N * N * 2  # uvw twice, vis in + out, grid out + in
    line = {
        "metric": "Mvis/s gridded+degridded through the MS4 host-memory calls "
                  "(gridKernelCPUHalfGCF + deGridKernelCPUHalfGCF, PCIe copies included)",
        "value": round(n / (ms / 1e3) / 1e6, 3), "unit": "Mvis/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128/f64", "data": "synthetic",
        "config": {"workload": "w-projection points in 32640 chunks of 306 (256 antennas), "
                               "4096x4096, 32 w-planes, chunk supports 8-64, 8x oversampling",
                   "points": n, "taps_per_direction": taps},
        "calls_ms": {"grid": round(mg, 3), "degrid": round(md, 3)},
        "host_copy_bytes_per_step": pcie,
        "device_fp64_grid_ms": {"tiled": round(ms_tiled, 4), "plain": round(ms_plain, 4),
                                "speedup": round(ms_plain / ms_tiled, 2),
                                "tiled_gtaps_per_s": round(taps / ms_tiled / 1e6, 1),
                                "rel_l2_tiled_vs_plain": rel},
        "device_fp64_degrid_ms": {"tiled": round(ms_dtiled, 4), "plain": round(ms_dplain, 4),
                                  "speedup": round(ms_dplain / ms_dtiled, 2),
                                  "rel_l2_tiled_vs_plain": drel},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(16, len(os.sched_getaffinity(0)))
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        kind = "reference" if O.ref_available(fast=True) else "port"
        impl = "ref_fast" if kind == "reference" else "oracle"

        def run(m):
            s = VisChunks(uvw=uvw[:m], vis=vis[:m], supps=sup[:m], scale=vc.scale,
                          wstep=vc.wstep)
            t0 = time.perf_counter()
            g, tp = O.grid(s, g128, N, impl=impl, nthreads=threads)
            t1 = time.perf_counter()
            O.degrid(s, g128, g, impl=impl, nthreads=threads)
            t2 = time.perf_counter()
            return t2 - t0, t1 - t0, t2 - t1, tp

        m = 300
        dt, _, _, _ = run(m)
        m2 = int(min(nb, max(m, m * args.cpu_seconds / max(dt, 1e-3))))
        dt, tg, td, tp = run(m2)
        line["cpu_baseline"] = {
            "value": round(m2 * ts / dt / 1e6, 4), "unit": "Mvis/s", "cores": threads,
            "kind": kind,
            "sample": f"first {m2} chunks ({m2 * ts} points, taps {tp}), full 4096^2 "
                      f"complex128 grid, grid {tg:.2f} s + degrid {td:.2f} s, "
                      f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}"}
    print(json.dumps(line), flush=True)
    return 0


def bench_flow(args) -> int:
    """The MS6 visibility binning kernels (binSizer + binner) through their
    host-memory C ABI on the points of the w-projection workload as 5-double
    rows, 32 x 32 uv regions over the grid, 10 w bins.  One step = one
    rcg_flow_bin_sizes + one rcg_flow_binner call.  value = points per second
    through both calls (PCIe included).  The CPU baseline is the oracle's
    restatement (pure Python, as the reference's Haskell loops are scalar) on
    a sample."""
    import ctypes as C
    from rc_amd import _lib, flow, synth

    vc, _, N = synth.wproj(n_times=args.n_times)
    n = vc.n
    rows = np.zeros((n, 5))
    rows[:, :3] = vc.uvw.reshape(-1, 3)
    rows[:, 3] = vc.vis.reshape(-1).real
    rows[:, 4] = vc.vis.reshape(-1).imag
    half = (N / 2) / vc.scale                 # gridXY2UV of the grid edges
    e = np.linspace(-half, half, 33)
    regions = np.array([[e[i], e[i + 1], e[j], e[j + 1]] for i in range(32) for j in range(32)])
    bins = 10
    h = _lib.lib()
    sizes = np.zeros((len(regions), bins, 3), np.int64)
    _, _, se, counts = flow.bin_sizes(rows, regions, bins)
    off = np.arange(0, bins * (len(regions) + 1), bins, dtype=np.int32)
    rng_ = np.ascontiguousarray(se.reshape(-1, 2))
    szs = np.ascontiguousarray(counts.reshape(-1))
    cap = counts.sum(axis=1)
    outs = [np.empty((int(c), 5)) for c in cap]
    ptrs = (C.c_void_p * len(outs))(*[o.ctypes.data if o.size else None for o in outs])

    def step():
        if h.rcg_flow_bin_sizes(rows.ctypes.data, n, 5, regions.ctypes.data, len(regions), bins,
                                sizes.ctypes.data, None):
            raise _lib.RcgError(_lib.last_error()[1])
        if h.rcg_flow_binner(rows.ctypes.data, n, 5, regions.ctypes.data, len(regions),
                             off.ctypes.data, rng_.ctypes.data, szs.ctypes.data, ptrs, None):
            raise _lib.RcgError(_lib.last_error()[1])

    for _ in range(args.warmup):
        step()
    ts = []
    for _ in range(args.steps):
        t = time.perf_counter()
        step()
        ts.append(time.perf_counter() - t)
    ms = float(np.median(ts)) * 1e3
    line = {
        "metric": "Mpoints/s through binSizer + binner (32x32 uv regions, 10 w bins), host "
                  "memory in and out",
        "value": round(n / (ms / 1e3) / 1e6, 3), "unit": "Mpoints/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "w-projection points (256 antennas x 306 times) as 5-double "
                               "rows, 32x32 uv regions over 4096x4096, 10 w bins",
                   "points": n, "binned": int(cap.sum())},
    }
    if not args.no_cpu:
        from oracle import flow_oracle as F
        m = 200_000
        t = time.perf_counter()
        F.bin_sizes(rows[:m], regions, bins)
        t1 = time.perf_counter()
        line["cpu_baseline"] = {"value": round(m / (t1 - t) / 1e6, 4), "unit": "Mpoints/s",
                                "cores": 1, "kind": "port",
                                "sample": f"binSizer over the first {m} points, oracle "
                                          f"(pure Python), {t1 - t:.2f} s"}
    print(json.dumps(line), flush=True)
    return 0


def bench_small(args) -> int:
    """The small kernels next to the gridder on the headline sizes (4096^2
    complex128 grid, the 9,987,840 points of the w-projection workload):
    normalize, rotate, grid0, reweight, the hermitian fold and the
    fft_inplace_even transform, each timed with HIP events on the stream it
    runs on, against its algorithmic bytes (HBM roofline, 8 TB/s).  The CPU
    baselines are the reference's own functions compiled with its release
    flags (normalizeCPU with OpenMP, the others serial as in the reference),
    the oracle's hermitian fold and numpy's FFT."""
    import ctypes as C
    import torch
    from rc_amd import _lib, synth
    from rc_amd import device as D

    vc, _, N = synth.wproj(n_times=args.n_times)
    n, P = vc.n, N
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    uvw_h = np.ascontiguousarray(vc.uvw.reshape(-1, 3), np.float64)
    vis_h = vc.vis.reshape(-1).astype(np.complex128)
    grid_h = (rng.standard_normal((N, P)) + 1j * rng.standard_normal((N, P))).astype(np.complex128)
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.from_numpy(vis_h).to(dev)
    grid = torch.from_numpy(grid_h).to(dev)
    counts = torch.empty(N * N, dtype=torch.int32, device=dev)
    h = _lib.lib()
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)
    fft = D.FFTPlan(-1, N, P, f64=True)
    scale = vc.scale

    def run(name, fn):
        _lib.check(fn(), name)

    kernels = {
        # name: (call, algorithmic bytes per launch)
        "normalize": (lambda: h.rcg_normalize(grid.data_ptr(), 1, P, N, sp), 32 * N * P),
        "rotate": (lambda: h.rcg_rotate(uvw.data_ptr(), vis.data_ptr(), 1, n, scale, sp),
                   (24 + 32) * n),
        "grid0": (lambda: h.rcg_grid0(uvw.data_ptr(), vis.data_ptr(),
                                      grid.data_ptr() + (N // 2 * (P + 1)) * 16, scale, n, P, N,
                                      sp), (24 + 16 + 32) * n),
        "reweight": (lambda: h.rcg_reweight(uvw.data_ptr(), vis.data_ptr(), scale, n, N,
                                            counts.data_ptr(), sp),
                     (24 + 8) * n + (24 + 4 + 32) * n + 4 * N * N),
        "herm_padded": (lambda: h.rcg_herm_padded(grid.data_ptr(), N, P, sp), 24 * N * P),
        "fft_c2c": (lambda: h.rcg_fft_execute(fft._h, grid.data_ptr(), sp), 3 * 32 * N * P),
    }
    res = {}
    for name, (fn, nbytes) in kernels.items():
        for _ in range(args.warmup):
            run(name, fn)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(args.steps):
            run(name, fn)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        gbs = nbytes / (ms / 1e3) / 1e9
        res[name] = {"ms": round(ms, 4), "bytes": int(nbytes), "gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 3)}
    fft.close()
    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(16, len(os.sched_getaffinity(0)))
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        ref = O.ref_lib(fast=True) if O.ref_available(fast=True) else None
        g = grid_h.copy()
        v = vis_h.copy()

        def cpu(name, fn, kind, cores):
            t = time.perf_counter()
            fn()
            res[name]["cpu_ms"] = round((time.perf_counter() - t) * 1e3, 2)
            res[name]["cpu_kind"] = kind
            res[name]["cpu_cores"] = cores
            res[name]["speedup"] = round(res[name]["cpu_ms"] / res[name]["ms"], 1)

        if ref is not None:
            cpu("normalize", lambda: ref.normalizeCPU(g.ctypes.data, P, N), "reference", threads)
            cpu("rotate", lambda: ref.rotateCPU(uvw_h.ctypes.data, v.ctypes.data, n, scale),
                "reference", 1)
            cpu("grid0", lambda: ref.grid0(uvw_h.ctypes.data, v.ctypes.data,
                                           g.ctypes.data + (N // 2 * (P + 1)) * 16, scale, n, P,
                                           N), "reference", 1)
            cpu("reweight", lambda: ref.reweight(uvw_h.ctypes.data, v.ctypes.data, scale, n, N),
                "reference", 1)
        cpu("herm_padded", lambda: O.oracle_lib().orc_herm_padded(g.ctypes.data, N, P), "port",
            1)
        cpu("fft_c2c", lambda: np.fft.fft2(g), "port (numpy)", 1)
    fr = [r["frac"] for r in res.values()]
    line = {
        "metric": "small kernels on the headline sizes: geometric-mean fraction of the HBM "
                  "roofline", "value": round(float(np.exp(np.mean(np.log(fr)))), 3),
        "unit": "fraction of 8 TB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sum(r["ms"] for r in res.values()), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "4096x4096 complex128 grid, 9,987,840 w-projection points",
                   "points": n, "grid": N},
        "kernels": res,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_next(args) -> int:
    """The components beside the path, on the headline points:
    - GCF generation (the reference's GCF kernel: per-plane mean w, then one
      mkGCFLayer per plane; planes -31..31, supports min(64, 8 + 2|p|))
      on the device, against the oracle's numpy restatement (the reference's
      GCF.cpp needs FFTW, which is absent);
    - the MS6 buffer_t scatter and degrid (FP64, S = 16, the full 4096^2
      grid) and nvGridder (2048^2 grid, 16x16 GCF, with its in-place sort)
      through their host-memory ABIs, against the oracle's per-point
      restatements on a sample (Halide and CUDA are absent)."""
    import ctypes as C
    import torch
    from oracle import gcf_oracle as GO
    from oracle import halide_oracle as HO
    from rc_amd import gcf as G
    from rc_amd import halide as H
    from rc_amd import synth

    vc, _, N = synth.wproj(n_times=args.n_times)
    n = vc.n
    uvw_h = np.ascontiguousarray(vc.uvw.reshape(-1, 3), np.float64)
    res = {}

    def timed(fn, steps, warmup, sync=True):
        for _ in range(warmup):
            fn()
        if sync:
            torch.cuda.synchronize()
        ts = []
        for _ in range(steps):
            t = time.perf_counter()
            fn()
            if sync:
                torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)) * 1e3

    # GCF generation
    uvw_d = torch.from_numpy(uvw_h).cuda()
    M, theta = 31, 0.1
    ms = timed(lambda: G.make_gcf(uvw_d, vc.wstep, theta, 8, 2, 64, max_w_plane=M),
               max(3, args.steps // 2), 1)
    r = {"ms": round(ms, 3), "planes": 2 * M + 1}
    if not args.no_cpu:
        t = time.perf_counter()
        GO.gcf_kernel(uvw_h, vc.wstep, theta, 8, 2, 64, M)
        r["cpu_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        r["cpu_kind"] = "port (oracle, numpy FFT)"
        r["speedup"] = round(r["cpu_ms"] / ms, 1)
    res["gcf_generation"] = r

    # MS6 scatter / degrid, buffer_t ABI, FP64
    rng = np.random.default_rng(7)
    S = 16
    vis5 = np.zeros((n, 5))
    vis5[:, :3] = uvw_h
    vis5[:, 3] = vc.vis.reshape(-1).real
    vis5[:, 4] = vc.vis.reshape(-1).imag
    gcf = np.ascontiguousarray(rng.standard_normal((64, S, S, 2)))
    uvg = np.zeros((N, N, 2))
    vb, gb, ub = H.vis_buffer(vis5), H.gcf_buffer(gcf), H.uvg_buffer(uvg)
    out5 = np.zeros_like(vis5)
    ob = H.vis_buffer(out5)
    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} returned {rc}")

    ms_s = timed(lambda: ok(H.scatter(vc.scale, N, 0, vb, gb, ub), "kern_scatter"), 3, 1,
                 sync=False)
    ms_d = timed(lambda: ok(H.degrid(vc.scale, N, 0, gb, ub, vb, ob), "kern_degrid"), 3, 1,
                 sync=False)
    r = {"points": n, "scatter_ms": round(ms_s, 1), "degrid_ms": round(ms_d, 1),
         "scatter_mpts_s": round(n / ms_s / 1e3, 1), "degrid_mpts_s": round(n / ms_d / 1e3, 1)}
    if not args.no_cpu:
        m = 20_000
        g2 = np.zeros((N, N, 2))
        t = time.perf_counter()
        HO.scatter(vis5[:m], gcf, g2, vc.scale, N, 0)
        r["cpu_scatter_mpts_s"] = round(m / (time.perf_counter() - t) / 1e6, 4)
        r["cpu_kind"] = f"port (oracle, Python loop over the first {m} points)"
    res["ms6_scatter_degrid"] = r
    del uvg

    # nvGridder, 2048^2, 16x16 GCF
    NG = 2048
    m_nv = min(n, 2_000_000)
    nv = np.zeros((m_nv, 5))
    nv[:, 0] = uvw_h[:m_nv, 0] * vc.scale * (NG - 48) / N + NG / 2
    nv[:, 1] = uvw_h[:m_nv, 1] * vc.scale * (NG - 48) / N + NG / 2
    nv[:, 3:] = vis5[:m_nv, 3:]
    gnv = np.ascontiguousarray(rng.standard_normal((64, 16, 16, 2)))
    out_nv = np.zeros((NG, NG, 2))
    nvb, ngb, nob = H.vis_buffer(nv), H.nv_gcf_buffer(gnv), H.uvg_buffer(out_nv)
    ms_nv = timed(lambda: ok(H.nv_gridder(1.0, NG, nvb, ngb, nob), "nvGridder"), 3, 1,
                  sync=False)
    r = {"points": m_nv, "ms": round(ms_nv, 1), "mpts_s": round(m_nv / ms_nv / 1e3, 1)}
    if not args.no_cpu:
        m = 1000
        t = time.perf_counter()
        HO.nv_grid(nv[:m], gnv, 1.0, NG)
        r["cpu_mpts_s"] = round(m / (time.perf_counter() - t) / 1e6, 5)
        r["cpu_kind"] = f"port (oracle, Python loop over {m} points)"
    res["nv_gridder"] = r
    line = {
        "metric": "next-row components on the headline points: GCF generation, MS6 buffer_t "
                  "scatter/degrid, nvGridder (host-memory ABIs, PCIe included)",
        "value": res["ms6_scatter_degrid"]["scatter_mpts_s"], "unit": "Mpoints/s (MS6 scatter)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "w-projection points (256 antennas x 306 times)", "points": n},
        "components": res,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_tune(args) -> int:
    """The autotuner's sweep on the wproj workload (4096^2, 32 w-planes,
    9,987,840 points): every (tile, chunk, waves) candidate timed over bin +
    grid + degrid with the library's kernel events; prints the table and the
    chosen point."""
    import torch
    from rc_amd import autotune
    from rc_amd import device as D
    from rc_amd import synth

    vc, gcf, N = synth.wproj(n_times=args.n_times)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    vis = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).cuda()
    tune = autotune.autotune(geom, dg, vis, max_support=64, reps=max(2, args.steps), log=log)
    line = {"workload": "wproj 4096x4096, 32 w-planes, supports 8-64", "points": vc.n,
            "best": {k: tune[k] for k in ("tile", "chunk", "waves", "degrid_waves", "ms_step")},
            "degrid_table": [{k: round(r[k], 4) if isinstance(r[k], float) else r[k]
                              for k in ("degrid_waves", "ms_degrid")}
                             for r in tune["degrid_table"]],
            "tune_s": round(tune["tune_s"], 1),
            "table": [{k: (round(r[k], 4) if isinstance(r.get(k), float) else r.get(k))
                       for k in ("tile", "chunk", "waves", "ms_step", "ms_grid", "ms_degrid",
                                 "skipped")} for r in tune["table"]]}
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the path "
                          "is designed to shard across GPUs, but no multi-GPU "
                          "measurement is made", "n_gpus": args.gpus}))
        return 2
    import torch
    from rc_amd import device as D
    from rc_amd import synth

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device visible"}))
        return 1
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "imaging":
        return bench_imaging(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "binner":
        return bench_binner(args)
    if args.workload == "flow":
        return bench_flow(args)
    if args.workload == "small":
        return bench_small(args)
    if args.workload == "next":
        return bench_next(args)
    if args.workload == "ms4":
        return bench_ms4(args)
    if args.workload == "tune":
        return bench_tune(args)
    dev = torch.device("cuda:0")
    t0 = time.time()
    vc, gcf, N = synth.wproj(n_times=args.n_times)
    log(f"workload: {vc.n} points, {time.time() - t0:.1f} s to generate")
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg32 = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    vis32 = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).to(dev)
    plan = D.TiledPlan(geom, tile=args.tile, chunk=args.chunk, max_support=64,
                       layout=args.layout, waves=args.waves, degrid_waves=args.degrid_waves)
    grid32 = torch.empty((N, N), dtype=torch.complex64, device=dev)
    out32 = torch.empty(vc.n, dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream()

    def step32(ev=None):
        if ev is not None:
            ev[0].record(stream)
        plan.bin(dg32)
        if ev is not None:
            ev[1].record(stream)
        plan.grid(dg32, vis32, grid32)
        if ev is not None:
            ev[2].record(stream)
        plan.degrid(dg32, grid32, out32)
        if ev is not None:
            ev[3].record(stream)

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()
    st = plan.stats()
    if st["err"]:
        print(json.dumps({"error": f"binning error word {st['err']}"}))
        return 1
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    plan.timing(True)  # HIP events around the grid and degrid kernels themselves
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step32(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k_grid = float(np.mean(plan.kernel_times("grid")))
    k_degrid = float(np.mean(plan.kernel_times("degrid")))
    plan.timing(False)
    ms_step = wall / args.steps * 1e3
    t_bin = np.mean([e[0].elapsed_time(e[1]) for e in evs])
    t_grid = np.mean([e[1].elapsed_time(e[2]) for e in evs])
    t_degrid = np.mean([e[2].elapsed_time(e[3]) for e in evs])
    value = vc.n / (ms_step / 1e3) / 1e6
    log(f"fp32 tiled: {ms_step:.3f} ms/step (bin {t_bin:.3f} grid {t_grid:.3f} "
        f"degrid {t_degrid:.3f} ms; kernels grid {k_grid:.3f} degrid {k_degrid:.3f} ms) "
        f"-> {value:.1f} Mvis/s; stats {st}")

    extra = {}
    taps = st["taps"]
    if not args.no_fp64:
        dg64 = D.DeviceGCF.from_table(gcf, dtype=np.complex128)
        vis64 = vis32.to(torch.complex128)
        grid64 = torch.empty((N, N), dtype=torch.complex128, device=dev)
        out64 = torch.empty(vc.n, dtype=torch.complex128, device=dev)

        def step64():
            pg, _, _ = D.pregrid(geom, dg64)
            D.grid_plain_f64(geom, dg64, pg, vis64, grid64)
            D.degrid_plain_f64(geom, dg64, pg, grid64, out64)

        step64()
        torch.cuda.synchronize()
        n64 = max(2, args.steps // 3)
        t0 = time.perf_counter()
        for _ in range(n64):
            step64()
        torch.cuda.synchronize()
        ms64 = (time.perf_counter() - t0) / n64 * 1e3
        v64 = vc.n / (ms64 / 1e3) / 1e6
        # errors of the FP32 path against FP64 on the same complex64 inputs
        plan.grid(dg32, vis32, grid32)
        g_err = float(torch.linalg.vector_norm(grid32.to(torch.complex128) - grid64)
                      / torch.linalg.vector_norm(grid64))
        g32in = grid64.to(torch.complex64)
        plan.degrid(dg32, g32in, out32)
        D.degrid_plain_f64(geom, dg64, D.pregrid(geom, dg64)[0], g32in.to(torch.complex128), out64)
        d_err = float(torch.linalg.vector_norm(out32.to(torch.complex128) - out64)
                      / torch.linalg.vector_norm(out64))
        torch.cuda.synchronize()
        log(f"fp64 plain: {ms64:.3f} ms/step -> {v64:.1f} Mvis/s; fp32 grid rel L2 "
            f"{g_err:.2e}, degrid rel L2 {d_err:.2e}")
        # the fastest FP64 path of this library on the same inputs: binned
        # pairs + tiled FP64 gridder, plain FP64 degridder (faster than the
        # tiled one)
        plan64 = D.TiledPlan(geom, tile=32, chunk=1024, max_support=64)

        def step64t():
            plan64.bin(dg64)
            plan64.grid_f64(dg64, vis64, grid64)
            pg, _, _ = D.pregrid(geom, dg64)
            D.degrid_plain_f64(geom, dg64, pg, grid64, out64)

        step64t()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n64):
            step64t()
        torch.cuda.synchronize()
        ms64t = (time.perf_counter() - t0) / n64 * 1e3
        plan64.close()
        v64t = vc.n / (ms64t / 1e3) / 1e6
        log(f"fp64 tiled grid + plain degrid: {ms64t:.3f} ms/step -> {v64t:.1f} Mvis/s")
        extra = {"fp64_plain_mvis_s": round(v64, 3), "fp64_plain_ms_per_step": round(ms64, 3),
                 "speedup_vs_fp64_plain": round(value / v64, 3),
                 "fp64_tiled_mvis_s": round(v64t, 3), "fp64_tiled_ms_per_step": round(ms64t, 3),
                 "speedup_vs_fp64_tiled": round(value / v64t, 3),
                 "error_vs_fp64": {"grid_rel_l2": g_err, "degrid_rel_l2": d_err}}

    # Roofline of the dominant kernel, timed by the HIP events the library
    # records around it on its stream.  Its bound is the larger of two
    # fractions: bytes between L2 and memory (rocprofv3 per-request-size
    # counters, profiled at the default configuration) against HBM peak, and
    # FP32 work (8 flops per tap: the reference's accounting) against the
    # unpacked FP32 peak.  Also reported: VALU issue (the useful FMA
    # instructions, 4 per 64 taps, and all VALU/SALU instructions the
    # counters saw, against one instruction per CU per cycle at peak clock),
    # and the tap stream (8 B per tap read, mostly served from L2 and the
    # MALL, so never an HBM figure).
    t_dom = max(k_grid, k_degrid)
    # the kernels are templates on tile and wavefronts per workgroup (the
    # plan's defaults: 3 gridding wavefronts at tile 32, else 4; 4 degridding)
    if k_grid >= k_degrid:
        dom, nw = "k_grid_tiles", args.waves or (3 if args.tile == 32 else 4)
    else:
        dom, nw = "k_degrid_tiles", args.degrid_waves or 4
    kname = f"{dom}<{args.tile}, {nw}>"
    sec = t_dom / 1e3
    slots = sec * CLOCK_GHZ * 1e9 * N_CU
    fp32_tf = 8 * taps / sec / 1e12
    b_dir = vc.n * (24 + 8) + 8 * taps + 8 * N * N
    default_cfg = (args.tile == 32 and args.chunk == 1024 and args.n_times == 306 and
                   args.layout == "pairs" and args.waves == 0 and args.degrid_waves == 0)
    prof = profiled(kname.replace(" ", "")) if default_cfg else {}
    tb = prof.get("traffic_bytes")
    hbm_gbs = tb / sec / 1e9 if tb else None
    fracs = {"hbm": hbm_gbs / HBM_PEAK_GBS if tb else -1.0,
             "fp32": fp32_tf / FP32_VALU_PEAK_TFLOPS}
    bound = max(fracs, key=fracs.get)
    if bound == "hbm":
        roof = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(fracs["hbm"], 4)}
    else:
        roof = {"bound": "fp32", "achieved": round(fp32_tf, 3), "peak": FP32_VALU_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(fracs["fp32"], 4)}
    roof.update({
        "traffic": tb,
        "traffic_source": ("profiles/wproj_tiled_traffic.md: rocprofv3 TCC_EA0_RDREQ_32B/64B/128B "
                           "x 32/64/128 B + TCC_EA0_WRREQ(_64B) x 32/64 B per dispatch "
                           "(L2 to memory, an upper bound of HBM bytes)") if tb else None,
        "kernel": kname, "kernel_ms": round(t_dom, 4),
        "fp32": {"achieved": round(fp32_tf, 3), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(fracs["fp32"], 4), "flops_per_tap": 8},
        "valu_issue": {"useful_instr": int(taps * 4 // 64), "slots": int(slots),
                       "frac_useful": round(taps * 4 / 64 / slots, 4),
                       "valu_instr": prof.get("valu_instr"), "salu_instr": prof.get("salu_instr"),
                       "frac_valu": round(prof["valu_instr"] / slots, 4)
                       if "valu_instr" in prof else None,
                       "frac_salu": round(prof["salu_instr"] / slots, 4)
                       if "salu_instr" in prof else None,
                       "clock_ghz": CLOCK_GHZ},
        "tap_stream": {"bytes": b_dir, "gbs": round(b_dir / sec / 1e9, 1),
                       "note": "uvw + vis + 8 B per tap + grid per direction; the taps are "
                               "re-read from L2/MALL, so this is not an HBM figure"},
    })
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(vc, gcf, N, args.cpu_seconds)
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # reported, not fatal
            cpu = {"error": repr(e)}

    line = {
        "metric": "Mvis/s gridded+degridded, 4096² 32 w-planes: FP32 batched vs FP64 plain @1e-5",
        "value": round(value, 3), "unit": "Mvis/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64/f32 (f64 pregrid)",
        "data": "synthetic",
        "config": {"workload": "w-projection grid+degrid, 4096x4096, 32 w-planes, "
                               "supports 8-64, 8x oversampling",
                   "points": vc.n, "taps_per_direction": taps, "tile": args.tile,
                   "chunk": args.chunk, "waves": plan.waves or "default",
                   "degrid_waves": plan.degrid_waves or "default",
                   "layout": args.layout},
        "phases_ms": {"pregrid_bin": round(float(t_bin), 4), "grid": round(float(t_grid), 4),
                      "degrid": round(float(t_degrid), 4)},
        "roofline": roof, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(line), flush=True)
    plan.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
