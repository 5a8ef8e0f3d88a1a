"""Autotuner of the tiled FP32 path: tile size and pairs per work item.

For a given workload (one geometry or a batch of instances, a GCF table and
visibilities on the device) every candidate (tile, chunk) is built as a
plan, binned, warmed up once and timed over `reps` bin + grid + degrid
passes with the kernel events the library records; the fastest candidate
wins.  Candidates the plan rejects (too many tiles for the binning
histograms, too many tiles per point) are reported and skipped.  Results
are cached per workload signature (instances, points, grid shape, support
histogram, tap count) in memory and, if `cache_path` is given, in a JSON
file.

Tile size sets the LDS footprint per workgroup (4 wave-private FP32 copies
of (T+1) x T cells for gridding) and with it the occupancy and how many
tiles a window spans; chunk sets the pairs per work item, i.e. how many
visibilities each wavefront walks between tile flushes.
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
DEFAULT_CHUNKS = (128, 256, 512, 1024)

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


def _time_candidate(geoms, gcf, vis, tile, chunk, max_support, reps):
    plan = D.TiledPlan(geoms, tile=tile, chunk=chunk, max_support=max_support)
    try:
        grid = torch.empty(plan.grid_shape, dtype=torch.complex64, device=vis.device)
        out = torch.empty_like(vis)
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
        return {"tile": tile, "chunk": chunk, "ms_step": ev[0].elapsed_time(ev[1]) / reps,
                "ms_grid": float(np.mean(tg)), "ms_degrid": float(np.mean(td)),
                "pairs": st["pairs"], "items": st["items"]}
    finally:
        plan.close()


def autotune(geoms, gcf: D.DeviceGCF, vis: torch.Tensor, tiles=DEFAULT_TILES,
             chunks=DEFAULT_CHUNKS, reps: int = 2, max_support: int | None = None,
             cache_path: str | Path | None = None, log=None) -> dict:
    """Returns {"tile", "chunk", "ms_step", "table": [...], "signature"}."""
    sig = signature(geoms, gcf)
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
                r = _time_candidate(geoms, gcf, vis, tile, chunk, max_support, reps)
            except _lib.RcgError as e:
                r = {"tile": tile, "chunk": chunk, "ms_step": None, "skipped": str(e)}
            table.append(r)
            if log:
                log(f"autotune tile {tile} chunk {chunk}: {r.get('ms_step')}")
    ok = [r for r in table if r.get("ms_step") is not None]
    if not ok:
        raise _lib.RcgError("autotune: no candidate could be planned")
    best = min(ok, key=lambda r: r["ms_step"])
    res = {"tile": best["tile"], "chunk": best["chunk"], "ms_step": best["ms_step"],
           "table": table, "signature": sig, "tune_s": time.perf_counter() - t0}
    _memo[sig] = res
    if cache_path is not None:
        p = Path(cache_path)
        disk = json.loads(p.read_text()) if p.exists() else {}
        disk[sig] = res
        p.write_text(json.dumps(disk, indent=1))
    return res
