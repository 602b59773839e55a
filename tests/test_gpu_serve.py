"""Serving (continuous batching) on the GPU: requests over three resident
shapes of 1,280, 5,120 and 20,480 facets, with and without craters, at
heliocentric distances other than the shapes' own (flux scale), submitted
while earlier ones run, under an HBM budget tight enough that admissions wait
for memory, and with injected out-of-memory set builds (BTPM_INJECT_NOMEM)
so the scheduler's deferral path runs.  Every result equals the same request
run alone, bit for bit; sampled results match the FP64 oracle.  Plus the
device memory cap (btpm_ctx_set_mem_limit)."""
import hashlib
import time

import numpy as np
import pytest

from binarytpm_amd import _lib, inputs
from binarytpm_amd import thermal as th
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL_K = 0.05

# (subdivision, seed) -> 20 * 4^subdiv facets
SHAPES = ((3, 31), (4, 41), (5, 51))
# (subdivision, TI, craters, r in AU)
REQUESTS = [(5, 300.0, True, 1.3), (3, 50.0, True, 0.95), (4, 1200.0, False, 2.2),
            (5, 80.0, True, 2.7), (3, 2000.0, True, 1.0), (4, 400.0, True, 1.6),
            (5, 30.0, False, 0.9), (4, 15.0, True, 2.9), (3, 700.0, False, 1.8)]


def _digest(a):
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()


def _pool(ctx):
    cg = inputs.crater_cap(45.0)
    table = inputs.CraterTable(cg)
    pool = {}
    for sd, seed in SHAPES:
        case, sun_local = inputs.serving_shape(sd, seed=seed)
        p = th.params(case.ntau, case.rot_period, tacc=0.01)
        cm = th.CraterModel(vf=cg.vf, sun_local=sun_local, vis_table=table.vis,
                            normals=cg.normals, solar=inputs.solar_at(1.0), multiple=True,
                            scatter_acc=1e-4)
        pool[sd] = (case, th.Shape(ctx, p, case.illum, 0.1, crater=cm), cg)
    return pool


def _result(s, craters):
    out = (_digest(s.surface_temps()), _digest(s.crater_temps()) if craters else None)
    inf = s.info()
    return out, (inf.rotations, inf.crater_rotations)


def test_server_mixed_shapes_memory_pressure(monkeypatch):
    with th.Context(0) as ctx:
        pool = _pool(ctx)
        alone = []
        for sd, ti, cr, r in REQUESTS:
            s = th.Solve(ctx, pool[sd][1], ti, with_craters=cr, flux_scale=1.0 / (r * r)).run()
            alone.append(_result(s, cr))
            s.close()
        # admissions budgeted to 6 % of the free HBM: one 20,480-facet crater
        # solve (11.5 GB) fits, a second has to wait; every third set build
        # fails as if the device were full and is retried
        monkeypatch.setenv("BTPM_INJECT_NOMEM", "3")
        srv = th.Server(ctx, th.batch_opts(rotations_per_sync=0, mem_fraction=0.06))
        solves = []
        for sd, ti, cr, r in REQUESTS:
            s = th.Solve(ctx, pool[sd][1], ti, with_craters=cr, flux_scale=1.0 / (r * r))
            srv.submit(s)
            solves.append(s)
            time.sleep(0.02)
        mixed = {}
        left = list(range(len(solves)))
        while left:
            k = srv.wait_any([solves[i] for i in left], timeout_s=300.0)
            i = left.pop(k)
            mixed[i] = _result(solves[i], REQUESTS[i][2])
            if not (REQUESTS[i][0] == 3 and REQUESTS[i][2]):
                solves[i].close()   # the 1,280-facet crater results stay for the oracle
        st = srv.stop()
        monkeypatch.delenv("BTPM_INJECT_NOMEM")
        print(f"[serve] {st.solves} solves, {st.converged} converged, max resident "
              f"{st.max_resident}, admissions that waited for memory {st.mem_waits}, "
              f"deferred set builds {st.deferrals}, graph launches {st.graph_launches}",
              flush=True)
        assert st.solves == len(REQUESTS) and st.converged == len(REQUESTS)
        assert st.mem_waits > 0 and st.deferrals > 0 and st.max_resident >= 2
        for i, req in enumerate(REQUESTS):
            assert mixed[i] == alone[i], (req, mixed[i], alone[i])
        # oracle: the 1,280-facet crater requests, smooth in full and 20 crater
        # parents (no shape self-heating on the serving shapes: RGH = 0)
        case, shape, cg = pool[3]
        il_all = shape.crater_illum()
        sel = np.linspace(0, case.illum.shape[0] - 1, 20).astype(int)
        p = O.params(case.ntau, case.rot_period, tacc=0.01)
        for i, (sd, ti, cr, r) in enumerate(REQUESTS):
            if sd != 3 or not cr:
                continue
            g = 1.0 / (r * r)
            s = solves[i]
            sm = O.smooth_solve(p, case.illum.astype(np.float64) * g, 0.1, ti)
            d = np.abs(s.surface_temps() - sm.stemp).max()
            il = il_all[sel].astype(np.float64) * g
            crs = O.crater_solve(p, cg.vf, 0.1, ti, il, np.zeros_like(il))
            dc = np.abs(s.crater_temps()[sel] - crs.stemp).max()
            inf = s.info()
            print(f"[serve] oracle TI {ti:g} r {r} AU: smooth max|dT| {d:.4f} K ({inf.rotations} "
                  f"rot, oracle {sm.rotations}), craters {dc:.4f} K ({inf.crater_rotations} rot, "
                  f"oracle {crs.rotations})", flush=True)
            assert d <= TOL_K and dc <= TOL_K
            assert abs(inf.rotations - sm.rotations) <= 1
            s.close()
        for _, shp, _ in pool.values():
            shp.close()


def test_mem_limit_caps_allocations(ctx):
    """A device cap below what a solve needs makes its set build fail with
    an out-of-memory error, not a crash; lifting the cap lets it run.  (Setting
    the cap returns the cached blocks, so every buffer of the solve is a new
    allocation.)"""
    case, sun_local = inputs.serving_shape(3, seed=5)
    cg = inputs.crater_cap(45.0)
    table = inputs.CraterTable(cg)
    p = th.params(case.ntau, case.rot_period, tacc=0.01)
    shape = th.Shape(ctx, p, case.illum, 0.1,
                     crater=th.CraterModel(vf=cg.vf, sun_local=sun_local, vis_table=table.vis,
                                           normals=cg.normals, solar=inputs.solar_at(1.0)))
    try:
        ctx.set_mem_limit(1 << 20)     # far below what the shape already holds
        s = th.Solve(ctx, shape, 200.0, with_craters=True)
        with pytest.raises(_lib.BtpmError, match="device limit|out of memory") as e:
            s.run()
        assert e.value.code == _lib.BTPM_ERR_NOMEM
        s.close()
    finally:
        ctx.set_mem_limit(0)
    s = th.Solve(ctx, shape, 200.0, with_craters=True).run()
    assert s.info().crater_converged
    s.close()
    shape.close()
