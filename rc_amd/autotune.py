"""Autotuner of the tiled FP32 path: tile size, pairs per work item and
wavefronts per gridding workgroup.

For a given workload (one geometry or a batch of instances, a GCF table and
visibilities on the device) every candidate (tile, chunk, waves) is built as
a plan, binned, warmed up once and timed over `reps` bin + grid + degrid
passes with the kernel events the library records; the fastest candidate
wins.  Candidates the plan rejects (too many tiles per point for the
support) are reported and skipped.  Results are cached per workload
signature (instances, points, grid shape, support histogram, tap count) in
memory and, if `cache_path` is given, in a JSON file.

Tile size sets how many tiles a window spans and the LDS footprint of one
wave-private FP32 copy of (T+1) x T cells; waves (1..4) sets how many such
copies a gridding workgroup holds, i.e. the split of a CU's 160 KB of LDS
between copies per workgroup and workgroups per CU; chunk sets the pairs
per work item, i.e. how many visibilities each wavefront walks between tile
flushes.  The plan for a candidate is binned once per tile and chunk and
re-timed for each wave count (rcg_tiled_set_grid_waves).
"""
from __future__ import annotations

import hashlib
import json
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib
from . import device as D

DEFAULT_TILES = (16, 32, 64)
DEFAULT_CHUNKS = (256, 512, 1024, 2048, 4096)
DEFAULT_WAVES = (1, 2, 3, 4)
DEFAULT_DEGRID_WAVES = (2, 4, 8)

_memo: dict[str, dict] = {}


def signature(geoms, gcf: D.DeviceGCF) -> str:
    geoms = list(geoms) if isinstance(geoms, (list, tuple)) else [geoms]
    h = hashlib.sha1()
    h.update(np.asarray([len(geoms), geoms[0].grid_size, geoms[0].grid_pitch,
                         gcf.data.numel()], np.int64).tobytes())
    for g in geoms:
        hist = torch.bincount(g.chunk_supps.to(torch.int64).cpu(), minlength=1) \
            if g.n_chunks else torch.zeros(1, dtype=torch.int64)
        h.update(np.asarray([g.n_chunks, g.ts_ch], np.int64).tobytes())
        h.update(hist.numpy().astype(np.int64).tobytes())
    return h.hexdigest()


def _time_plan(plan, gcf, vis, grid, out, reps):
    plan.bin(gcf)
    plan.grid(gcf, vis, grid)
    plan.degrid(gcf, grid, out)
    st = plan.stats()
    if st["err"]:
        raise _lib.RcgError(f"binning error word {st['err']}")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    plan.timing(True)
    ev[0].record()
    for _ in range(reps):
        plan.bin(gcf)
        plan.grid(gcf, vis, grid)
        plan.degrid(gcf, grid, out)
    ev[1].record()
    torch.cuda.synchronize()
    tg, td = plan.kernel_times("grid"), plan.kernel_times("degrid")
    plan.timing(False)
    return {"ms_step": ev[0].elapsed_time(ev[1]) / reps, "ms_grid": float(np.mean(tg)),
            "ms_degrid": float(np.mean(td)), "pairs": st["pairs"], "items": st["items"]}


def _time_candidates(geoms, gcf, vis, tile, chunk, waves, max_support, reps):
    """One row per wave count for a (tile, chunk) plan."""
    plan = D.TiledPlan(geoms, tile=tile, chunk=chunk, max_support=max_support)
    rows = []
    try:
        grid = torch.empty(plan.grid_shape, dtype=torch.complex64, device=vis.device)
        out = torch.empty_like(vis)
        for w in waves:
            plan.set_grid_waves(w)
            rows.append({"tile": tile, "chunk": chunk, "waves": w,
                         **_time_plan(plan, gcf, vis, grid, out, reps)})
    finally:
        plan.close()
    return rows


def _time_degrid_waves(geoms, gcf, vis, best, dwaves, max_support, reps):
    """The chosen (tile, chunk, waves) plan re-timed per degrid wave count."""
    plan = D.TiledPlan(geoms, tile=best["tile"], chunk=best["chunk"], max_support=max_support,
                       waves=best["waves"])
    rows = []
    try:
        grid = torch.empty(plan.grid_shape, dtype=torch.complex64, device=vis.device)
        out = torch.empty_like(vis)
        for dw in dwaves:
            plan.set_degrid_waves(dw)
            rows.append({"degrid_waves": dw, **_time_plan(plan, gcf, vis, grid, out, reps)})
    finally:
        plan.close()
    return rows


def autotune(geoms, gcf: D.DeviceGCF, vis: torch.Tensor, tiles=DEFAULT_TILES,
             chunks=DEFAULT_CHUNKS, waves=DEFAULT_WAVES, degrid_waves=DEFAULT_DEGRID_WAVES,
             reps: int = 2, max_support: int | None = None,
             cache_path: str | Path | None = None, log=None) -> dict:
    """Returns {"tile", "chunk", "waves", "degrid_waves", "ms_step", "table":
    [...], "degrid_table": [...], "signature"}.  The degridder's wave count
    is searched after the others, on the chosen plan (the degridder's time
    does not depend on the gridder's wave count)."""
    sig = (signature(geoms, gcf)
           + f"/{tuple(tiles)}/{tuple(chunks)}/{tuple(waves)}/{tuple(degrid_waves)}")
    if sig in _memo:
        return _memo[sig]
    if cache_path is not None and Path(cache_path).exists():
        disk = json.loads(Path(cache_path).read_text())
        if sig in disk:
            _memo[sig] = disk[sig]
            return disk[sig]
    table = []
    t0 = time.perf_counter()
    for tile in tiles:
        for chunk in chunks:
            try:
                rows = _time_candidates(geoms, gcf, vis, tile, chunk, waves, max_support, reps)
            except _lib.RcgError as e:
                rows = [{"tile": tile, "chunk": chunk, "waves": None, "ms_step": None,
                         "skipped": str(e)}]
            table += rows
            if log:
                for r in rows:
                    log(f"autotune tile {tile} chunk {chunk} waves {r.get('waves')}: "
                        f"{r.get('ms_step')}")
    ok = [r for r in table if r.get("ms_step") is not None]
    if not ok:
        raise _lib.RcgError("autotune: no candidate could be planned")
    best = min(ok, key=lambda r: r["ms_step"])
    dtable = _time_degrid_waves(geoms, gcf, vis, best, degrid_waves, max_support, reps) \
        if degrid_waves else []
    dbest = min(dtable, key=lambda r: r["ms_degrid"]) if dtable else {"degrid_waves": 0}
    if log:
        for r in dtable:
            log(f"autotune degrid waves {r['degrid_waves']}: degrid {r['ms_degrid']:.4f} ms")
    res = {"tile": best["tile"], "chunk": best["chunk"], "waves": best["waves"],
           "degrid_waves": dbest["degrid_waves"], "ms_step": best["ms_step"], "table": table,
           "degrid_table": dtable, "signature": sig, "tune_s": time.perf_counter() - t0}
    _memo[sig] = res
    if cache_path is not None:
        p = Path(cache_path)
        disk = json.loads(p.read_text()) if p.exists() else {}
        disk[sig] = res
        p.write_text(json.dumps(disk, indent=1))
    return res
