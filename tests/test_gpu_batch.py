"""GPU tests of batched plans (several instances in one launch), HIP graph
capture and replay, and kernel timing.

Parity bar as for the one-instance tiled path: every instance's grid and
visibilities within 1e-5 relative L2 of the FP64 oracle run on the same
complex64-rounded inputs, tap counts equal."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rc_amd import device as D
from rc_amd import synth
from rc_amd.tables import VisChunks

pytestmark = pytest.mark.gpu

TOL = 1e-5


def small_batch(seed=7, n_inst=5, grid_size=256, empty=True):
    vcs, gcf, N = synth.snapshots(seed=seed, n_inst=n_inst, grid_size=grid_size, n_planes=4,
                                  s_min=8, s_max=24, median=3000, sigma=1.0,
                                  clip=(200, 20000), n_ant=48, vis_dtype=np.complex128)
    if empty:  # an instance without points
        vcs[2] = VisChunks(uvw=np.zeros((0, 1, 3)), vis=np.zeros((0, 1), np.complex128),
                           supps=np.zeros(0, np.int32), scale=vcs[0].scale, wstep=vcs[0].wstep)
    r = lambda v: VisChunks(uvw=v.uvw, vis=v.vis.astype(np.complex64).astype(np.complex128),
                            supps=v.supps, scale=v.scale, wstep=v.wstep)
    return [r(v) for v in vcs], gcf.astype(np.complex64).astype(np.complex128), N


def upload(vcs, gcf, N):
    geoms = [D.DeviceGeom.from_chunks(v, N) for v in vcs]
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    vis = torch.from_numpy(np.concatenate([v.vis.reshape(-1) for v in vcs])
                           .astype(np.complex64)).cuda()
    return geoms, dg, vis


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_batch_vs_oracle(hip_device, tile):
    vcs, gcf, N = small_batch()
    geoms, dg, vis = upload(vcs, gcf, N)
    plan = D.TiledPlan(geoms, tile=tile, chunk=37)
    plan.bin(dg)
    g = plan.grid(dg, vis)
    # degrid from the oracle grids rounded to complex64
    refs, vrefs, taps = [], [], 0
    for v in vcs:
        if v.n:
            ref, t = O.grid(v, gcf, N, nthreads=8)
            taps += t
        else:
            ref = np.zeros((N, N), np.complex128)
        ref32 = ref.astype(np.complex64).astype(np.complex128)
        refs.append(ref)
        vrefs.append(O.degrid(v, gcf, ref32, nthreads=8)[0].reshape(-1) if v.n
                     else np.zeros(0, np.complex128))
    gin = torch.from_numpy(np.stack(refs).astype(np.complex64)).cuda()
    out = plan.degrid(dg, gin)
    st = plan.stats()
    assert st["err"] == 0 and st["taps"] == taps
    g_np, out_np = g.cpu().numpy(), out.cpu().numpy()
    off = 0
    for k, v in enumerate(vcs):
        if v.n:
            assert rel_l2(g_np[k], refs[k]) <= TOL, k
            assert rel_l2(out_np[off:off + v.n], vrefs[k]) <= TOL, k
        else:
            assert not np.any(g_np[k])
        off += v.n
    plan.close()


def test_batch_matches_one_plan_per_instance(hip_device):
    vcs, gcf, N = small_batch(seed=8, n_inst=4, empty=False)
    geoms, dg, vis = upload(vcs, gcf, N)
    plan = D.TiledPlan(geoms, tile=32, chunk=64)
    plan.bin(dg)
    g = plan.grid(dg, vis)
    v = plan.degrid(dg, g)
    off = 0
    for k, geom in enumerate(geoms):
        p1 = D.TiledPlan(geom, tile=32, chunk=64)
        p1.bin(dg)
        g1 = p1.grid(dg, vis[off:off + geom.n].contiguous())
        v1 = p1.degrid(dg, g[k].contiguous())
        assert rel_l2(g[k].cpu().numpy(), g1.cpu().numpy()) <= 1e-6
        assert rel_l2(v[off:off + geom.n].cpu().numpy(), v1.cpu().numpy()) <= 1e-6
        off += geom.n
        p1.close()
    plan.close()


def test_graph_replays_match_eager(hip_device):
    """bin + grid + degrid captured once, replayed several times with eager
    work of the program in between; every replay equals the eager result
    (FP32 atomics make the order of additions differ, hence 1e-6), and a
    replay after the visibilities change sees the new values."""
    vcs, gcf, N = small_batch(seed=9, n_inst=3, empty=False)
    geoms, dg, vis = upload(vcs, gcf, N)
    plan = D.TiledPlan(geoms, tile=32, chunk=50)
    grid = torch.empty(plan.grid_shape, dtype=torch.complex64, device="cuda")
    out = torch.empty(plan.n, dtype=torch.complex64, device="cuda")
    graph = plan.capture(dg, D.STAGE_BIN | D.STAGE_GRID | D.STAGE_DEGRID, vis=vis, grid=grid,
                         vis_out=out)
    plan.bin(dg)
    g_eager = plan.grid(dg, vis).cpu().numpy()
    v_eager = plan.degrid(dg, torch.from_numpy(g_eager).cuda()).cpu().numpy()
    for rep in range(3):
        grid.fill_(1.0)
        out.fill_(1.0)
        graph.launch()
        torch.cuda.synchronize()
        assert rel_l2(grid.cpu().numpy(), g_eager) <= 1e-6, rep
        assert rel_l2(out.cpu().numpy(), v_eager) <= 1e-5, rep
        # eager work between replays
        tmp = plan.grid(dg, vis)
        assert rel_l2(tmp.cpu().numpy(), g_eager) <= 1e-6
    vis.mul_(2.0)
    graph.launch()
    torch.cuda.synchronize()
    assert rel_l2(grid.cpu().numpy(), 2 * g_eager) <= 1e-6
    graph.close()
    plan.close()


def test_kernel_timing(hip_device):
    vcs, gcf, N = small_batch(seed=10, n_inst=2, empty=False)
    geoms, dg, vis = upload(vcs, gcf, N)
    plan = D.TiledPlan(geoms, tile=32, chunk=64)
    plan.bin(dg)
    plan.timing(True)
    for _ in range(3):
        g = plan.grid(dg, vis)
        plan.degrid(dg, g)
    tg, td = plan.kernel_times("grid"), plan.kernel_times("degrid")
    assert len(tg) == 3 and len(td) == 3
    assert all(t > 0 for t in tg + td)
    assert plan.kernel_times("grid") == []
    plan.close()


def test_plan_rejects_mixed_grid_shapes(hip_device):
    vcs, gcf, N = small_batch(seed=11, n_inst=2, empty=False)
    g0 = D.DeviceGeom.from_chunks(vcs[0], N)
    g1 = D.DeviceGeom.from_chunks(vcs[1], N // 2)
    from rc_amd import _lib
    with pytest.raises(_lib.RcgError, match="grid shape"):
        D.TiledPlan([g0, g1], tile=32, chunk=64, max_support=24)


def test_autotune_picks_a_timed_candidate(hip_device):
    from rc_amd import autotune
    vcs, gcf, N = small_batch(seed=12, n_inst=3, empty=False)
    geoms, dg, vis = upload(vcs, gcf, N)
    kw = dict(tiles=(16, 32, 64), chunks=(64, 256), waves=(1, 3), reps=1)
    res = autotune.autotune(geoms, dg, vis, **kw)
    timed = [r for r in res["table"] if r["ms_step"] is not None]
    assert len(res["table"]) == 12 and timed
    assert {(r["tile"], r["chunk"], r["waves"]) for r in timed} <= \
        {(t, c, w) for t in (16, 32, 64) for c in (64, 256) for w in (1, 3)}
    assert (res["tile"], res["chunk"], res["waves"]) in \
        [(r["tile"], r["chunk"], r["waves"]) for r in timed]
    assert res["ms_step"] == min(r["ms_step"] for r in timed)
    assert res["degrid_waves"] in (2, 4, 8) and len(res["degrid_table"]) == 3
    # memoised per workload signature and sweep
    assert autotune.autotune(geoms, dg, vis, **kw) is res
