#!/usr/bin/env python3
"""Benchmark: the 128-solve thermal-inertia x roughness sweep on one MI355X.

Workload (one "step"): 32 thermal inertias log-spaced in 10..2500 SI x
roughness fractions {0, 0.25, 0.5, 1.0} = 128 solves over the rough primary:
a 5120-facet icosphere-ellipsoid (1.2 x 1.0 x 0.9 km, P = 4.3 h, 650 steps,
r = 1.1 AU, A = 0.1) with 3 %-dense facet view factors and a 45-degree
spherical-cap crater of 100 sub-facets on every facet.  f = 0 solves are the
smooth facets only (5120 columns); f > 0 solves are smooth facets, then RGH
and 512,000 crater columns, each to convergence (TACC = 0.01 K, Newton
TACC/10, freeze after rev > 5, cap 301).  All 128 run through the batched,
graph-captured scheduler; inputs are resident in HBM before timing, and the
results stay on the device (end_to_end adds copying them out).

Beside the headline: the same sweep warm-started along each roughness
class's TI chain as the reference runs it (warm_sweep), the reference-loop
CPU baseline on a bounded sample (cpu_baseline, warm as the reference, and
cold), the plain single-solve GPU path, a subset of the sweep on the
config-1 binary primary (9750-step 15:4 cycle with eclipses), serving, and
the producers and consumer around the stage.

Prints ONE JSON line.  value = converged solves per hour (whole job).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "conduct_crater_traffic.md")
VALU_FILE = os.path.join(ROOT, "profiles", "conduct_crater_valu.md")
MFMA_PEAK_TFLOPS = 2500.0  # dense FP16
NX = 56
BYTES_PER_COLSTEP = 12.0   # read absorbed flux + previous-rotation T, write T


def summary_value(path, key):
    """A `key: value` line of a committed counter summary under profiles/,
    or None when absent."""
    try:
        with open(path) as fh:
            for ln in fh:
                if ln.startswith(key + ":"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic_per_colstep():
    """HBM bytes per integrated crater column-step from the FETCH_SIZE /
    WRITE_SIZE counter summary (scaled by a copy of known size)."""
    return summary_value(TRAFFIC_FILE, "traffic_bytes_per_column_step")


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def build_workload(subdiv, gamma):
    from binarytpm_amd import inputs
    t0 = time.time()
    case, crater = inputs.rough_primary(subdiv=subdiv, gamma_deg=gamma)
    log(f"inputs generated in {time.time() - t0:.1f} s: nf={case.illum.shape[0]}, "
        f"nnz={len(case.csr[1])}")
    return case, crater


def make_shape(ctx, case, crater):
    from binarytpm_amd import thermal as th
    p = th.params(case.ntau, case.rot_period, tacc=0.01)
    cm = th.CraterModel(vf=crater["vf"], shadow=crater["shadow"], cosang=crater["cosang"],
                        solar=crater["solar"], multiple=True, scatter_acc=1e-4,
                        drf=crater["drf"])
    return th.Shape(ctx, p, case.illum, 0.1, csr=case.csr, scatter=2, scatter_acc=1e-4, crater=cm)


def sweep_plan(n_ti=32):
    from binarytpm_amd import inputs
    tis = inputs.sweep_inertias(n_ti)
    return [(float(t), f) for t in tis for f in (0.0, 0.25, 0.5, 1.0)]


def run_sweep(ctx, shape, plan, opts):
    from binarytpm_amd import thermal as th
    solves = [th.Solve(ctx, shape, ti, with_craters=f > 0) for ti, f in plan]
    t0 = time.perf_counter()
    stats = th.batch_run(ctx, solves, opts)
    wall = time.perf_counter() - t0
    infos = [s.info() for s in solves]
    return solves, stats, wall, infos


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ctx, shape, case, crater, plan, solves, infos, n_craters=256, n_check=16,
                 warm_ti_idx=(6, 25), cold_ti_idx=(6,)):
    """Times the float OpenMP restatement of the reference loop shape (refcpu)
    on a bounded sample of the sweep and checks the GPU results of the same
    solves against the FP64 oracle.

    Sample: the full 5120-facet smooth solve (self-heating couples every
    facet) and `n_craters` crater parents (their view-factor rows kept, so RGH
    and RGI are the full shape's).  Warm, as the reference runs its sweep
    (each TI from the previous TI's converged state, thermal.cpp:3676-3760):
    TI k = 6 and 25 of the 32 started from the GPU's converged state of TI
    k-1 (its smooth profiles and series, and the sampled craters' state).
    Cold: TI k = 6 from the initial state.  Each timed solve gives the CPU
    cost per rotation (smooth) and per rotation and crater parent (craters);
    a sweep's CPU time is that cost times the rotations every one of its
    solves took (the GPU runs' counts, within one of the CPU's)."""
    from binarytpm_amd import inputs
    from oracle import oracle as O
    nf = case.illum.shape[0]
    es = 0.95 * 5.67e-8
    threads = len(os.sched_getaffinity(0))
    tis = inputs.sweep_inertias(32)
    t0 = time.time()
    F = O.facet_scatter(case.illum, 0.1, case.csr, multiple=True, acc=1e-4)
    rp, ci, vv = case.csr

    def parents(sel):
        sub_rp = np.concatenate([[0], np.cumsum(rp[sel + 1] - rp[sel])]).astype(np.int32)
        rows = np.concatenate([np.arange(rp[f], rp[f + 1]) for f in sel])
        sub_csr = (sub_rp, ci[rows], vv[rows])
        sub_drf = crater["drf"][rows]
        rgi = O.rgi(sub_csr, sub_drf, F, np.full(nf, 0.1), len(sel))
        il = O.crater_scatter(crater["vf"], 0.1, crater["shadow"][sel], crater["cosang"][sel],
                              rgi, crater["solar"], multiple=True, acc=1e-4)
        return sub_csr, sub_drf, il

    sel = np.linspace(0, nf - 1, n_craters).astype(int)
    csr_t, drf_t, il_t = parents(sel)
    chk = np.linspace(0, nf - 1, n_check).astype(int)
    csr_c, drf_c, il_c = parents(chk)
    log(f"CPU sample inputs (oracle scattering of {n_craters} + {n_check} parents) "
        f"{time.time() - t0:.1f} s")

    def rc_solves(ti, state_s=None, state_c=None):
        t = time.time()
        st_rc, rev_s, cs_s = O.refcpu_smooth_solve(F, 0.1, ti, case.ntau, case.rot_period,
                                                   tacc=0.01, csr=case.csr, nthreads=threads,
                                                   state=state_s)
        ts = time.time() - t
        rgh_rc = O.rgh(csr_t, drf_t, st_rc, es, n_craters)
        t = time.time()
        _, rev_c, cs_c = O.refcpu_crater_solve(crater["vf"], 0.1, ti, il_t, rgh_rc,
                                               case.rot_period, tacc=0.01, nthreads=threads,
                                               state=state_c)
        tc = time.time() - t
        return dict(ts=ts, tc=tc, rs=rev_s, rc=rev_c, cs=cs_s + cs_c)

    runs = {"warm": [], "cold": []}
    for k in warm_ti_idx:
        g = solves[plan.index((float(tis[k - 1]), 0.5))]
        st_s = {"temp": g.profile(), "stemp": g.surface_temps()}
        st_c = {"temp": g.crater_profile()[sel], "stemp": g.crater_temps()[sel]}
        runs["warm"].append(rc_solves(float(tis[k]), st_s, st_c))
    for k in cold_ti_idx:
        runs["cold"].append(rc_solves(float(tis[k])))
    for kind, ks in (("warm", warm_ti_idx), ("cold", cold_ti_idx)):
        for k, r in zip(ks, runs[kind]):
            log(f"refcpu TI={tis[k]:.1f} {kind}: smooth {r['ts']:.1f} s ({r['rs']} rot), "
                f"{n_craters} crater parents {r['tc']:.1f} s ({r['rc']} rot)")
    # cost per rotation: smooth facets, and one crater parent
    all_runs = runs["warm"] + runs["cold"]
    t_rot_s = sum(r["ts"] for r in all_runs) / sum(r["rs"] for r in all_runs)
    t_rot_c = sum(r["tc"] for r in all_runs) / sum(r["rc"] for r in all_runs) / n_craters

    def sweep_s(inf_list, fr):
        return sum(t_rot_s * i.rotations + (t_rot_c * nf * i.crater_rotations if f > 0 else 0.0)
                   for i, f in zip(inf_list, fr))

    # FP64 oracle check of the GPU results at the warm TIs (cold f = 0.5 solves)
    dmax, sq, cnt = 0.0, 0.0, 0
    p = O.params(case.ntau, case.rot_period, tacc=0.01)
    for k in warm_ti_idx:
        ti = float(tis[k])
        sm = O.smooth_solve(p, F, 0.1, ti, csr=case.csr)
        rgh = O.rgh(csr_c, drf_c, sm.stemp, es, n_check)
        cr = O.crater_solve(p, crater["vf"], 0.1, ti, il_c, rgh)
        idx = plan.index((ti, 0.5))
        g_s = solves[idx].surface_temps().astype(np.float64)
        g_c = solves[idx].crater_temps()[chk].astype(np.float64)
        for a, b in ((g_s, sm.stemp), (g_c, cr.stemp)):
            d = np.abs(a - b)
            dmax = max(dmax, float(d.max()))
            sq += float((d ** 2).sum())
            cnt += d.size
        log(f"oracle check TI={ti:.1f}: running max|dT| {dmax:.4f} K")
    sample_s = sum(r["ts"] + r["tc"] for r in all_runs)
    colsteps = sum(r["cs"] for r in all_runs)
    cold_s = sweep_s(infos, [f for _, f in plan])
    tis_w = [round(float(tis[k]), 1) for k in warm_ti_idx]
    return {
        "value": None,            # the warm sweep's, filled in by the caller
        "unit": "converged solves/hour",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": (f"refcpu (float OpenMP restatement of the reference loop shape): the full "
                   f"5120-facet smooth solve and {n_craters} of 5120 crater parents at TI "
                   f"{tis_w}, each from the converged state of the previous TI of the sweep "
                   f"(as the reference runs it), and at TI "
                   f"{[round(float(tis[k]), 1) for k in cold_ti_idx]} from the initial state; "
                   f"{sample_s:.1f} s of timed CPU work, {t_rot_s:.3f} s per smooth rotation "
                   f"and {t_rot_c * 1e3:.2f} ms per crater parent and rotation, times every "
                   f"solve's rotations in the sweep"),
        "facet_layer_steps_per_s": colsteps * NX / sample_s,
        "cold": {"value": 3600.0 * len(plan) / cold_s, "sweep_s": cold_s,
                 "note": "the same per-rotation costs times the cold sweep's rotations"},
        "t_rot_smooth_s": t_rot_s,
        "t_rot_crater_parent_s": t_rot_c,
        "sweep_cost": sweep_s,       # for the caller (removed before printing)
        "check_max_dT_K": dmax,
        "check_rms_dT_K": float(np.sqrt(sq / max(cnt, 1))),
        "check_sample": f"GPU f=0.5 solves at TI {tis_w}: smooth facets in full, "
                        f"{n_check} crater parents, vs the FP64 oracle",
    }


def warm_sweep(ctx, shape, plan, opts):
    """The sweep as the reference runs it: per roughness class one chain over
    the 32 TIs in ascending order, each solve warm-started from the previous
    one's final state (btpm_solve_warm_start); the four chains run as one
    batch (the scheduler admits a solve once its predecessor's phase is done),
    every f > 0 chain computed (no deduplication)."""
    from binarytpm_amd import thermal as th
    fracs = sorted({f for _, f in plan})
    solves, order = [], []
    for f in fracs:
        prev = None
        for ti in sorted({t for t, g in plan if g == f}):
            s = th.Solve(ctx, shape, ti, with_craters=f > 0, warm_from=prev)
            solves.append(s)
            order.append((ti, f))
            prev = s
    t0 = time.perf_counter()
    st = th.batch_run(ctx, solves, opts)
    wall = time.perf_counter() - t0
    infos = [s.info() for s in solves]
    out = {
        "solves_per_hour": 3600.0 * st.converged / wall,
        "wall_s": wall,
        "converged": int(st.converged),
        "rotations_smooth": int(sum(i.rotations for i in infos)),
        "rotations_crater": int(sum(i.crater_rotations for i in infos)),
        "facet_layer_steps_per_s": st.column_steps * NX / wall,
    }
    return out, solves, order


def fetch_all(solves):
    """Copies every solve's results to host memory (surface temperatures,
    and crater temperatures of the rough solves), as a caller of the stage
    takes them; reused host buffers."""
    t0 = time.perf_counter()
    nbytes = 0
    buf_s = buf_c = None
    for s in solves:
        buf_s = s.surface_temps(buf_s)
        nbytes += buf_s.nbytes
        if s.with_craters:
            buf_c = s.crater_temps(buf_c)
            nbytes += buf_c.nbytes
    return time.perf_counter() - t0, nbytes


def binary_phase(ctx, n_ti=4):
    """The sweep on the config-1 binary's primary: 15:4 super-cycle over 9750
    steps (numrevs 15), eclipse masks from the secondary, the primary's facet
    view factors and a cap crater on every facet (512,000 crater columns, 40 GB
    of crater state per solve): a stratified subset of the 128 solves -- TIs
    k = 0, 10, 21, 31 of the 32 at f = 0 and f = 0.5 -- in one batch, admitted
    by HBM footprint."""
    from binarytpm_amd import inputs
    from binarytpm_amd import thermal as th
    t0 = time.time()
    case, crater = inputs.rough_binary_primary()
    log(f"binary primary inputs in {time.time() - t0:.1f} s ({case.illum.shape[1]} steps, "
        f"{(1 - case.mask.mean()) * 100:.1f} % of facet-steps eclipsed)")
    cg, tab = crater["geom"], crater["table"]
    p = th.params(case.ntau, case.rot_period, tacc=0.01)
    t0 = time.time()
    shape = th.Shape(ctx, p, case.illum, 0.1, numrevs=case.numrevs, csr=case.csr, scatter=2,
                     scatter_acc=1e-4,
                     crater=th.CraterModel(vf=cg.vf, sun_local=crater["sun_local"],
                                           vis_table=tab.vis, normals=cg.normals,
                                           solar=crater["solar"], multiple=True,
                                           scatter_acc=1e-4, drf=crater["drf"]))
    log(f"binary primary shape resident in {time.time() - t0:.1f} s")
    del crater
    tis = inputs.sweep_inertias(32)
    ks = [0, 10, 21, 31][:n_ti]
    plan = [(float(tis[k]), f) for k in ks for f in (0.0, 0.5)]
    solves = [th.Solve(ctx, shape, ti, with_craters=f > 0) for ti, f in plan]
    opts = th.batch_opts(rotations_per_sync=4, use_graph=True, profile=True)
    t0 = time.perf_counter()
    st = th.batch_run(ctx, solves, opts)
    wall = time.perf_counter() - t0
    infos = [s.info() for s in solves]
    out = {
        "workload": (f"config-1 binary primary: 5120 facets x {case.ntau} steps (15:4 cycle, "
                     "eclipses), 3 %-dense view factors, 45-deg cap crater x100 per facet; "
                     f"TI k={ks} of the 32 x f in {{0, 0.5}} ({len(plan)} solves)"),
        "wall_s": wall,
        "solves_per_hour": 3600.0 * st.converged / wall,
        "converged": int(st.converged),
        "max_resident": int(st.max_resident),
        "mem_waits": int(st.mem_waits),
        "facet_layer_steps_per_s": st.column_steps * NX / wall,
        "crater_avg_launch_ms": 1000.0 * st.crater_kernel_s / max(st.crater_launches, 1),
        "crater_kernel_gbs": (BYTES_PER_COLSTEP * st.crater_column_steps / st.crater_kernel_s
                              / 1e9 if st.crater_kernel_s > 0 else 0.0),
        "rotations": [(i.rotations, i.crater_rotations) for i in infos],
    }
    log(f"binary primary: {len(plan)} solves in {wall:.2f} s, {st.max_resident} resident at "
        f"most, crater kernel {out['crater_avg_launch_ms']:.1f} ms per launch")
    for s in solves:
        s.close()
    shape.close()
    return out


def plain_stepwise(ctx, shape, plan, infos):
    """The plain single-solve GPU kernel: the reference's GPU execution
    structure (an FD and a boundary launch plus a host sync per time step,
    profile in global memory, per-lane crater loop, host convergence check
    per rotation).  Two complete solves at the middle TI (f = 0 and f > 0)
    are timed; the sweep's plain time is extrapolated per phase with every
    solve's own rotation counts from the batched run (same algorithm)."""
    from binarytpm_amd import thermal as th
    mid = plan[len(plan) // 2][0]
    i0 = plan.index((mid, 0.0))
    i1 = plan.index((mid, 0.5))
    runs = {}
    for key, with_c in (("smooth", False), ("crater", True)):
        s = th.Solve(ctx, shape, mid, with_craters=with_c)
        t0 = time.perf_counter()
        s.run_stepwise()
        runs[key] = (time.perf_counter() - t0, s.info())
        s.close()
    t_s, inf_s = runs["smooth"]
    t_c = runs["crater"][0] - t_s
    inf_c = runs["crater"][1]
    rs = max(inf_s.rotations, 1)
    rc = max(inf_c.crater_rotations, 1)
    total = 0.0
    for (ti, f), inf in zip(plan, infos):
        total += t_s * inf.rotations / rs
        if f > 0:
            total += t_c * inf.crater_rotations / rc
    log(f"stepwise plain path at TI {mid:.1f}: smooth {t_s:.2f} s ({inf_s.rotations} rot), "
        f"crater {t_c:.2f} s ({inf_c.crater_rotations} rot); batched run had "
        f"{infos[i0].rotations}/{infos[i1].crater_rotations} rotations there")
    return {
        "kind": "stepwise (reference GPU structure: 2 launches + host sync per step)",
        "solves_per_hour": 3600.0 * len(plan) / total,
        "sample": (f"complete solves at TI {mid:.1f}: smooth {t_s:.2f} s / {inf_s.rotations} "
                   f"rotations, crater phase {t_c:.2f} s / {inf_c.crater_rotations} rotations; "
                   "extrapolated per phase by each sweep solve's rotation counts"),
        "sweep_s_extrapolated": total,
    }


def single_solve_persistent(ctx, shape, plan):
    """The persistent kernels on one solve at a time (no batching, no graphs,
    host check per rotation), at the middle TI, extrapolated as above."""
    from binarytpm_amd import thermal as th
    mid = plan[len(plan) // 2][0]
    t = {}
    for f in (0.0, 0.5):
        s = th.Solve(ctx, shape, mid, with_craters=f > 0)
        t0 = time.perf_counter()
        s.run()
        t[f] = time.perf_counter() - t0
        s.close()
    n0 = sum(1 for _, f in plan if f == 0)
    total = n0 * t[0.0] + (len(plan) - n0) * t[0.5]
    return 3600.0 * len(plan) / total


def serve_phase(ctx, n_req, rate, seed=7, subdivs=(3, 4, 5, 6), seeds_per_size=2, rps=4):
    """Serving workload: Poisson arrivals (fixed rate, seeded) of solve
    requests over a pool of resident seeded icosphere-ellipsoid shapes of
    1,280..81,920 facets, each request with TI ~ U[10, 2500], roughness
    f ~ U[0, 1] (craters iff f > 0) and r ~ U[0.9, 3.0] AU (flux scale
    1/r^2 on the 1 AU shape).  The crater sunlight is evaluated on the GPU
    from per-facet Sun directions.  Latency = submit (at arrival) to done."""
    import threading
    from binarytpm_amd import inputs
    from binarytpm_amd import thermal as th
    t0 = time.time()
    cg = inputs.crater_cap(45.0)
    table = inputs.CraterTable(cg)
    pool = {}
    for sd in subdivs:
        for k in range(seeds_per_size):
            case, sun_local = inputs.serving_shape(sd, seed=1000 * sd + k)
            p = th.params(case.ntau, case.rot_period, tacc=0.01)
            cm = th.CraterModel(vf=cg.vf, sun_local=sun_local, vis_table=table.vis,
                                normals=cg.normals, solar=inputs.solar_at(1.0), multiple=True,
                                scatter_acc=1e-4)
            pool[(sd, k)] = th.Shape(ctx, p, case.illum, 0.1, crater=cm)
            del sun_local, case
    log(f"serving pool: {len(pool)} shapes resident in {time.time() - t0:.1f} s, "
        f"{ctx.mem_info()[0] / 1e9:.0f} GB of HBM free")
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_req)
    arrive = np.cumsum(gaps) - gaps[0]
    reqs = [(int(rng.choice(subdivs)), int(rng.integers(seeds_per_size)),
             float(rng.uniform(10.0, 2500.0)), float(rng.uniform(0.0, 1.0)),
             float(rng.uniform(0.9, 3.0))) for _ in range(n_req)]
    srv = th.Server(ctx, th.batch_opts(rotations_per_sync=rps, use_graph=True))
    submitted, results, errors = [], [], []
    lock = threading.Condition()

    def waiter():
        # consume results in completion order (btpm_server_wait_any), so a
        # finished request never holds device memory while an earlier one runs
        outstanding, taken, finished = [], 0, 0
        while finished < n_req:
            with lock:
                while taken < len(submitted):
                    outstanding.append(submitted[taken] + (taken,))
                    taken += 1
                if not outstanding:
                    lock.wait()
                    continue
            try:
                k = srv.wait_any([o[0] for o in outstanding], timeout_s=0.5)
            except Exception as e:  # noqa: BLE001 -- timeouts poll for new arrivals
                if "timeout" in str(e):
                    continue
                errors.append(repr(e))
                return
            s, req, i = outstanding.pop(k)
            inf = s.info()
            results.append((req, inf.latency_s, inf.converged and
                            (not s.with_craters or inf.crater_converged),
                            inf.column_steps + inf.crater_column_steps, time.perf_counter()))
            s.close()
            finished += 1
            fr, _ = ctx.mem_info()
            log(f"serving: request {i} done ({20 * 4 ** req[0]} facets, TI {req[2]:.0f}, "
                f"latency {inf.latency_s:.2f} s; {fr / 1e9:.0f} GB of HBM free)")

    th_w = threading.Thread(target=waiter, daemon=True)
    th_w.start()
    t_start = time.perf_counter()
    for t_arr, req in zip(arrive, reqs):
        dt = t_start + t_arr - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        sd, k, ti, f, r = req
        s = th.Solve(ctx, pool[(sd, k)], ti, with_craters=f > 0, flux_scale=1.0 / (r * r))
        srv.submit(s)
        with lock:
            submitted.append((s, req))
            lock.notify_all()
    th_w.join()
    wall = max(x[4] for x in results) - t_start if results else float("nan")
    st = srv.stop()
    for sh in pool.values():
        sh.close()
    if errors:
        raise RuntimeError(f"serving errors: {errors[:3]}")
    lat = np.array([x[1] for x in results])
    conv = sum(1 for x in results if x[2])
    cols = sum(x[3] for x in results)
    sizes = [20 * 4 ** x[0][0] for x in results]
    log(f"serving: {n_req} requests at {rate}/s in {wall:.1f} s, p50 {np.percentile(lat, 50):.2f} s,"
        f" p99 {np.percentile(lat, 99):.2f} s, {conv} converged")
    return {
        "workload": (f"{n_req} Poisson arrivals at {rate}/s (seed {seed}) over {len(pool)} "
                     f"resident shapes of {min(sizes)}..{max(sizes)} facets, TI~U[10,2500], "
                     "f~U[0,1], r~U[0.9,3.0] AU, craters (x100 sub-facets) iff f>0, "
                     "650 steps, TACC 0.01 K"),
        "requests": n_req,
        "rate_per_s": rate,
        "wall_s": wall,
        "solves_per_hour": 3600.0 * conv / wall,
        "converged": conv,
        "facet_layer_steps_per_s": cols * NX / wall,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "graph_launches": int(st.graph_launches),
        "max_resident": int(st.max_resident),
        "deferred_set_builds": int(st.deferrals),
    }


def geometry_phase(ctx, cpu=True):
    """Illumination geometry and shadow producer at the config-1 size: the
    5120-facet primary and a 1280-facet secondary 3 km away, 15:4 cycle over
    9750 steps, Sun in the orbit plane (eclipses every half orbit).  GPU wall
    time of the whole call (inputs uploaded, [nf][ntau] outputs copied back);
    the FP64 oracle (one core) timed on sampled steps and checked bit for bit
    against the GPU on them."""
    import math

    from binarytpm_amd import inputs
    from binarytpm_amd import thermal as th
    equ, fv, npv = inputs.binary_shape(4, 3, axes_p=(0.6, 0.5, 0.45), radius_s=0.3, sep=3.0)
    ntau = 9750
    kw = dict(nv_primary=npv, rev_primary=15.0, rev_secondary=4.0)
    pos, pole = (1.1, 0.0, 0.0), (math.pi / 2, 0.0)
    th.geometry_illumination(ctx, equ, fv[:64], 16, pos, pole, sun_facet=False, **kw)  # warm-up
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        g = th.geometry_illumination(ctx, equ, fv, ntau, pos, pole, sun_facet=False, **kw)
        walls.append(time.perf_counter() - t0)
    gpu_s = min(walls)
    nf = len(fv)
    out = {
        "workload": (f"{nf} facets (5120 + 1280, 3 km apart) x {ntau} steps of a 15:4 cycle, "
                     "self-shadowing and mutual eclipses"),
        "gpu_s": gpu_s,
        "facet_steps_per_s": nf * ntau / gpu_s,
        "facet_pairs_per_s": nf * nf / 2 * ntau / gpu_s,
        "unlit_fraction": float(1.0 - g.shadow.mean()),
    }
    log(f"geometry: {nf} x {ntau} in {gpu_s:.3f} s (best of 3)")
    if cpu:
        from oracle import oracle as O
        sample = list(range(0, ntau, ntau // 40))[:40]
        t0 = time.perf_counter()
        mism = 0
        for s in sample:
            ref = O.geometry(equ, fv, ntau, pos, pole, steps=(s, 1), **kw)
            mism += int((ref.shadow[:, 0] != g.shadow[:, s]).sum())
        per_step = (time.perf_counter() - t0) / len(sample)
        out["cpu_baseline"] = {
            "value": per_step * ntau, "unit": "s per cycle", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (oracle/geometry.c, one core) on {len(sample)} of {ntau} "
                      f"steps, {per_step * 1e3:.1f} ms per step, extrapolated to the cycle",
        }
        out["vs_cpu"] = per_step * ntau / gpu_s
        out["shadow_mismatches_on_sample"] = mism
        log(f"geometry: oracle {per_step * 1e3:.1f} ms per step, GPU {per_step * ntau / gpu_s:.0f}x; "
            f"{mism} mismatches on {len(sample)} steps")
    return out


def valu_ceiling(device=0):
    """Wave instructions per second the device issues for unpacked 32-bit
    VALU instructions of one kind (calibration kernel of
    binarytpm_amd/bench_tools/probe.hip, built into libbtpm_probe.so beside
    the library but not part of it): the ceiling against which the
    compute-bound kernels' issue rates are read."""
    import ctypes as C
    path = os.path.join(ROOT, "binarytpm_amd", "libbtpm_probe.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build with `python -m binarytpm_amd.build`")
    L = C.CDLL(path)
    L.btpm_probe_valu.restype = C.c_int
    L.btpm_probe_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    out = {}
    for code, op in enumerate(("fma", "add", "mul", "sqrt", "iadd")):
        r = C.c_double()
        rc = L.btpm_probe_valu(device, code, 16384, C.byref(r))
        if rc != 0:
            raise RuntimeError(f"btpm_probe_valu({op}) failed with {rc}")
        out[op] = r.value
    return out


def viewfactor_phase(ctx, cpu=True):
    """Facet view-factor producer (Shape_Selfheating) on its own input, the
    6400-facet Doppler binary mesh of the reference (a committed fixture)
    with n = 10, then the shape->crater view factors of the resulting lists
    against the shipped 100-facet crater model.  GPU wall times of the whole
    calls (inputs uploaded, lists / [nnz][100] factors copied back), best of
    3; the float restatement (OpenMP over rows) timed on the whole shape for
    the lists and on sampled rows for the crater factors, and checked against
    the GPU."""
    from binarytpm_amd import thermal as th
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    d = np.load(os.path.join(gold, "doppler_binary.npz"))
    V, F = d["V"], d["F"]
    c = np.load(os.path.join(gold, "new_crater2.npz"))
    cv, cf = c["V"].astype(np.float32), c["F"].astype(np.int32)
    th.viewfactors(ctx, V, F[:256])   # warm-up
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        g = th.viewfactors(ctx, V, F)
        walls.append(time.perf_counter() - t0)
    gpu_s = min(walls)
    nf = len(F)
    # the cull tests every (G, Z, S) triple of a row's facing partners; counted
    # here from the kept lists (the cull removes 4 of 4.1 M pairs on this mesh)
    triples = float((np.diff(g.rowptr).astype(np.float64) ** 2).sum())
    out = {
        "workload": (f"Doppler binary mesh, {nf} facets (two 3200-facet bodies), n = 10 "
                     "sub-facet split, occlusion cull"),
        "gpu_s": gpu_s,
        "facing_pairs": int(g.n_facing),
        "entries": int(len(g.col)),
        "refined": int(g.n_refined),
        "pair_tests_per_s": nf * (nf - 1) / gpu_s,
        "cull_triples": triples,
        "cull_triples_per_s": triples / gpu_s,
    }
    log(f"viewfactors: {nf} facets, {g.n_facing} facing, {len(g.col)} kept in {gpu_s:.3f} s")
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        drf = th.crater_viewfactors(ctx, V, F, g.rowptr, g.vf, g.dir, cv, cf)
        walls.append(time.perf_counter() - t0)
    drf_s = min(walls)
    out["crater_factors"] = {
        "workload": f"{len(g.col)} list entries x {len(cf)} crater sub-facets x 100 rim samples",
        "gpu_s": drf_s,
        "factors_per_s": drf.size / drf_s,
        "output_gb": drf.nbytes / 1e9,
    }
    log(f"crater factors: {drf.shape[0]} x {drf.shape[1]} in {drf_s:.3f} s")
    if cpu:
        from oracle import oracle as O
        threads = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = int(os.environ["OMP_NUM_THREADS"])
        t0 = time.perf_counter()
        o = O.viewfactors(V, F)
        cpu_s = time.perf_counter() - t0
        same = bool(np.array_equal(o.rowptr, g.rowptr) and np.array_equal(o.col, g.col))
        out["cpu_baseline"] = {
            "value": cpu_s, "unit": "s per shape", "cores": threads, "kind": "port",
            "sample": "the whole shape: float restatement (oracle/viewfactor.c), OpenMP over rows",
        }
        out["vs_cpu"] = cpu_s / gpu_s
        out["lists_identical"] = same
        rows = np.flatnonzero(np.diff(g.rowptr) > 0)[::25]
        cnt = np.diff(g.rowptr)[rows]
        rp = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
        sel = np.concatenate([np.arange(g.rowptr[f], g.rowptr[f + 1]) for f in rows])
        t0 = time.perf_counter()
        od = O.crater_drf(V, F[rows], rp, g.vf[sel], g.dir[sel], cv, cf)
        per_entry = (time.perf_counter() - t0) / len(sel)
        out["crater_factors"]["cpu_baseline"] = {
            "value": per_entry * len(g.col), "unit": "s per shape", "cores": threads,
            "kind": "port",
            "sample": f"float restatement on {len(rows)} rows ({len(sel)} entries), "
                      "extrapolated to every entry",
        }
        out["crater_factors"]["vs_cpu"] = per_entry * len(g.col) / drf_s
        out["crater_factors"]["mismatches_on_sample"] = int((od != drf[sel]).sum())
        log(f"viewfactors: restatement {cpu_s:.2f} s on {threads} threads ({cpu_s / gpu_s:.0f}x), "
            f"lists identical {same}; crater factors {per_entry * len(g.col):.1f} s "
            f"extrapolated ({per_entry * len(g.col) / drf_s:.0f}x), "
            f"{out['crater_factors']['mismatches_on_sample']} mismatches on the sample")
    return out


def observe_phase(ctx, cpu=True):
    """Observe flux synthesis at the reference's configuration
    (175706/Run-04/observe_generator.cpp:18-35): the 4096-facet 1996 FG3 mesh
    (a committed fixture), 9:2 revolutions over 5850 steps, the rough model
    with the shipped 100-facet crater.  The inputs come from this library:
    facet visibility from the geometry producer with the observer as the
    source, instantaneous-equilibrium temperatures from its Sun illumination
    (A = 0.01), crater temperatures and visible fractions from a fixed
    sub-facet pattern.  GPU wall time of the call (19.2 GB of crater inputs
    uploaded, included); the FP64 restatement on sampled steps."""
    import math

    from binarytpm_amd import thermal as th
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    d = np.load(os.path.join(gold, "fg3_observe.npz"))
    c = np.load(os.path.join(gold, "new_crater2.npz"))
    V, F, cc = d["V"], d["F"], d["colour"]
    nf, ntau = len(F), 5850
    hel = (1.160088, -1.9521 / 180.0 * math.pi, 198.3142 / 180.0 * math.pi)
    obs = (0.583905, -3.882500 / 180.0 * math.pi, 257.492321 / 180.0 * math.pi)
    rot = (-83.0 / 180.0 * math.pi, 266.0 / 180.0 * math.pi)
    kw = dict(nv_primary=1026, rev_primary=9.0, rev_secondary=2.0)
    sun = th.geometry_illumination(ctx, V, F, ntau, hel, rot, sun_facet=False, **kw)
    seen = th.geometry_illumination(ctx, V, F, ntau, obs, rot, sun_facet=False, **kw)
    t_eq = np.power(np.maximum(0.99 * sun.illum, 0.0) / (0.95 * 5.67e-8), 0.25)
    st = np.maximum(t_eq, 100.0).astype(np.float32)
    pat = (0.85 + 0.3 * (np.arange(100) % 7) / 6.0).astype(np.float32)
    ct = (st[:, :, None] * pat[None, None, :]).astype(np.float32)
    cvis = (seen.shadow[:, :, None] * (0.2 + 0.8 * (np.arange(100) % 5) / 4.0)).astype(np.float32)
    crater = (c["V"], c["F"], cvis, ct)
    th.observe(ctx, V, F[:64], 16, hel, obs, rot, seen.shadow[:64, :16], st[:64, :16], cc,
               crater=(c["V"], c["F"], cvis[:64, :16], ct[:64, :16]), **kw)   # warm-up
    t0 = time.perf_counter()
    sm, ro = th.observe(ctx, V, F, ntau, hel, obs, rot, seen.shadow, st, cc, crater=crater, **kw)
    gpu_s = time.perf_counter() - t0
    out = {
        "workload": (f"1996 FG3 mesh, {nf} facets x {ntau} steps (9:2 revolutions) x 100 crater "
                     "sub-facets, 3 WISE bands, the reference's Sun / observer / pole"),
        "gpu_s": gpu_s,
        "subfacet_steps_per_s": nf * ntau * 100 / gpu_s,
        "visible_fraction": float(seen.shadow.mean()),
        "rough_flux_mean": [float(x) for x in ro.mean(axis=0)],
    }
    log(f"observe: {nf} x {ntau} x 100 in {gpu_s:.2f} s (crater inputs {2 * ct.nbytes / 1e9:.1f} GB)")
    if cpu:
        from oracle import oracle as O
        sample = list(range(0, ntau, ntau // 24))[:24]
        t0 = time.perf_counter()
        worst = 0.0
        for s in sample:
            o_sm, o_ro = O.observe(V, F, ntau, hel, obs, rot, seen.shadow, st, cc, crater=crater,
                                   steps=(s, 1), **kw)
            for a, b in ((sm[s], o_sm[0]), (ro[s], o_ro[0])):
                worst = max(worst, float((np.abs(a - b) / np.abs(b)).max()))
        per_step = (time.perf_counter() - t0) / len(sample)
        out["cpu_baseline"] = {
            "value": per_step * ntau, "unit": "s per cycle", "cores": 1, "kind": "port",
            "sample": f"FP64 restatement (oracle/observe.c, one core) on {len(sample)} of {ntau} "
                      f"steps, {per_step * 1e3:.1f} ms per step, extrapolated to the cycle",
        }
        out["vs_cpu"] = per_step * ntau / gpu_s
        out["max_rel_diff_on_sample"] = worst
        log(f"observe: restatement {per_step * 1e3:.1f} ms per step ({per_step * ntau / gpu_s:.0f}x), "
            f"max rel diff {worst:.1e} on {len(sample)} steps")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subdiv", type=int, default=4)
    ap.add_argument("--ti", type=int, default=32, help="thermal inertias in the sweep")
    ap.add_argument("--gamma", type=float, default=45.0)
    ap.add_argument("--rps", type=int, default=4,
                    help="rotations per host sync of the sweep (0: the scheduler's choice)")
    ap.add_argument("--serve-rps", type=int, default=0,
                    help="rotations per host sync of the server (0: the scheduler's choice)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-geometry", action="store_true",
                    help="skip the illumination geometry and shadow producer measurement")
    ap.add_argument("--no-viewfactors", action="store_true",
                    help="skip the view-factor producer measurement")
    ap.add_argument("--no-observe", action="store_true",
                    help="skip the Observe flux synthesis measurement")
    ap.add_argument("--no-warm", action="store_true",
                    help="skip the warm-started (chained) sweep")
    ap.add_argument("--no-binary", action="store_true",
                    help="skip the sweep subset on the config-1 binary primary")
    ap.add_argument("--no-fetch", action="store_true",
                    help="skip copying every result to the host after the timed sweep")
    ap.add_argument("--producers-only", action="store_true",
                    help="measure only the illumination and view-factor producers (one JSON "
                         "line with their results; for profiling)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (the known-size "
                         "kernel that scales HBM traffic counters)")
    ap.add_argument("--serve-n", type=int, default=40, help="serving requests")
    ap.add_argument("--serve-rate", type=float, default=1.3, help="serving arrivals per second")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (solves are "
              "independent and would be dealt to one process per GPU)", file=sys.stderr)
        sys.exit(2)

    from binarytpm_amd import thermal as th
    if args.copy_mib > 0:
        import torch
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written")
        del src, dst
    if args.producers_only:
        with th.Context(0) as ctx:
            line = {"valu_ceiling_wave_instr_per_s": valu_ceiling(),
                    "illumination_producer": geometry_phase(ctx, cpu=not args.no_cpu),
                    "viewfactor_producer": viewfactor_phase(ctx, cpu=not args.no_cpu),
                    "observe_fluxes": observe_phase(ctx, cpu=not args.no_cpu)}
        print(json.dumps(line), flush=True)
        return
    case, crater = build_workload(args.subdiv, args.gamma)
    plan = sweep_plan(args.ti)
    with th.Context(0) as ctx:
        t0 = time.time()
        shape = make_shape(ctx, case, crater)
        log(f"shape resident (upload + facet/crater scattering) in {time.time() - t0:.1f} s")
        opts = th.batch_opts(rotations_per_sync=args.rps, use_graph=True, profile=True)
        for w in range(args.warmup):
            solves, st, wall, _ = run_sweep(ctx, shape, plan, opts)
            log(f"warmup {w}: {wall:.2f} s")
            for s in solves:
                s.close()
        walls, stats, last = [], [], None
        for k in range(args.steps):
            if last is not None:
                for s in last:
                    s.close()
            solves, st, wall, infos = run_sweep(ctx, shape, plan, opts)
            walls.append(wall)
            stats.append(st)
            last = solves
            log(f"step {k}: {wall:.2f} s, {st.solves} solves, {st.converged} converged, "
                f"{st.syncs} syncs")
            log("rotations (smooth/crater) per solve: " + " ".join(
                f"{i.rotations}/{i.crater_rotations}" for i in infos))
        T = float(np.sum(walls))
        nsol = len(plan) * args.steps
        conv = sum(s.converged for s in stats)
        colsteps = sum(s.column_steps for s in stats)
        colsteps_all = sum(s.column_steps_all for s in stats)
        ck_s = sum(s.crater_kernel_s for s in stats)
        ck_n = sum(s.crater_launches for s in stats)
        ck_cs = sum(s.crater_column_steps for s in stats)
        sk_s = sum(s.smooth_kernel_s for s in stats)
        sk_n = sum(s.smooth_launches for s in stats)
        sk_cs = sum(s.smooth_column_steps for s in stats)
        hk_s = sum(s.heat_kernel_s for s in stats)
        hk_n = sum(s.heat_launches for s in stats)
        hk_sr = sum(s.heat_set_rotations for s in stats)
        nnz, nfac = len(case.csr[1]), int(case.illum.shape[0])
        # facet self-heating: per set-rotation the view-factor entries (index +
        # value), the emission read and the source read plus total written
        heat_bytes = 8.0 * nnz + 3.0 * 4.0 * nfac * case.ntau
        lat = np.array([i.latency_s for i in infos])
        achieved = BYTES_PER_COLSTEP * ck_cs / ck_s / 1e9 if ck_s > 0 else 0.0
        tpc = measured_traffic_per_colstep()
        traffic = tpc * ck_cs / ck_n if (tpc is not None and ck_n > 0) else None
        line = {
            "metric": ("converged solves/hour & facet\u00b7layer\u00b7steps/s, p50/p99 latency, "
                       "at |\u0394T|\u22640.05 K vs FP64"),
            "value": 3600.0 * conv / T,
            "unit": "converged solves/hour",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": 1000.0 * T / args.steps,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "precision": ("f32 state and conduction, f64 Newton surface balance, crater product "
                          "fp16 hi/lo x3 on MFMA with f32 accumulation"),
            "data": "synthetic (seeded shapes, illumination, view factors, crater model)",
            "config": {
                "workload": "128-solve sweep: 32 TI x f{0,.25,.5,1}, 5120-facet primary, "
                            "45-deg cap crater x100 sub-facets per facet, 650 steps, TACC 0.01 K",
                "solves": len(plan), "facets": int(case.illum.shape[0]),
                "crater_columns_per_rough_solve": int(case.illum.shape[0]) * 100,
                "ntau": case.ntau, "nx": NX, "parallelism": "1 GPU, batched solves",
            },
            "facet_layer_steps_per_s": colsteps * NX / T,
            "facet_layer_steps_per_s_incl_frozen": colsteps_all * NX / T,
            "solves_converged": int(conv),
            "timed_region": ("the sweep's solves from submission to convergence, inputs "
                             "resident in HBM; every result stays on the device (see "
                             "end_to_end for copying them out)"),
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "roofline": {
                "kernel": "btpm_conduct_crater",
                "bound": "hbm",
                "achieved": achieved,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic,
                "traffic_source": ("profiles/conduct_crater_traffic.md: FETCH_SIZE + WRITE_SIZE "
                                   "per integrated column-step, copy-calibrated, times this "
                                   "run's column-steps per launch") if traffic else None,
                "avg_launch_ms": 1000.0 * ck_s / max(ck_n, 1),
                "launches": ck_n,
                "bytes_per_column_step": BYTES_PER_COLSTEP,
                "smooth_kernel_gbs": BYTES_PER_COLSTEP * sk_cs / sk_s / 1e9 if sk_s > 0 else 0.0,
                "smooth_avg_launch_ms": 1000.0 * sk_s / max(sk_n, 1),
                # the kernel's actual bound: VALU issue (SQ counters, committed summary)
                "valu_instr_per_column_step": summary_value(VALU_FILE,
                                                            "valu_instr_per_column_step"),
                "valu_issue_frac": summary_value(VALU_FILE, "valu_issue_frac"),
                # the crater contraction on the matrix cores (executed FLOP incl. split)
                "mfma_crater_product_tflops": (
                    summary_value(VALU_FILE, "mfma_flop_per_column_step") * ck_cs / ck_s / 1e12
                    if (ck_s > 0 and summary_value(VALU_FILE, "mfma_flop_per_column_step"))
                    else None),
            },
            "selfheat_roofline": {
                "kernel": "btpm_selfheat_ell_h32 / btpm_selfheat_ell (one launch per shape)",
                "bound": "hbm",
                "bytes_per_set_rotation": heat_bytes,
                "achieved": heat_bytes * hk_sr / hk_s / 1e9 if hk_s > 0 else 0.0,
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": heat_bytes * hk_sr / hk_s / 1e9 / HBM_PEAK_GBS if hk_s > 0 else 0.0,
                "avg_launch_ms": 1000.0 * hk_s / max(hk_n, 1),
                "launches": hk_n,
                "set_rotations_per_launch": hk_sr / max(hk_n, 1),
            },
        }
        if not args.no_fetch:
            fs, nb = fetch_all(last)
            line["end_to_end"] = {
                "fetch_s": fs,
                "fetch_gb": nb / 1e9,
                "fetch_gbs": nb / fs / 1e9,
                "solves_per_hour": 3600.0 * len(plan) / (T / args.steps + fs),
                "note": ("the last timed sweep plus copying every solve's surface temperatures "
                         "and every rough solve's crater temperatures [nf][ntau][100] to host "
                         "memory"),
            }
            log(f"end to end: results {nb / 1e9:.1f} GB copied out in {fs:.2f} s "
                f"({nb / fs / 1e9:.1f} GB/s)")
        if not args.no_plain:
            pl = plain_stepwise(ctx, shape, plan, infos)
            line["plain"] = pl
            line["batched_over_plain"] = line["value"] / pl["solves_per_hour"]
            v1 = single_solve_persistent(ctx, shape, plan)
            line["single_solve_persistent_solves_per_hour"] = v1
            line["batched_over_single_solve_persistent"] = line["value"] / v1
            log(f"plain stepwise {pl['solves_per_hour']:.0f} solves/h, persistent single-solve "
                f"{v1:.0f} solves/h")
        cb = None
        if not args.no_cpu:
            cb = cpu_baseline(ctx, shape, case, crater, plan, last, infos)
        for s in last:
            s.close()
        ws = None
        if not args.no_warm:
            ws, wsolves, worder = warm_sweep(ctx, shape, plan, opts)
            line["warm_sweep"] = ws
            log(f"warm-started sweep: {ws['wall_s']:.2f} s ({ws['solves_per_hour']:.0f} solves/h), "
                f"rotations smooth {ws['rotations_smooth']} crater {ws['rotations_crater']}")
            winfos = [x.info() for x in wsolves]
            for s in reversed(wsolves):
                s.close()
        if cb is not None:
            cost = cb.pop("sweep_cost")
            if ws is not None:
                # the reference's own sweep: warm, its rotations those of the warm chains
                warm_s = cost(winfos, [f for _, f in worder])
                cb["value"] = 3600.0 * len(plan) / warm_s
                cb["sweep_s"] = warm_s
                ws["vs_cpu"] = ws["solves_per_hour"] / cb["value"]
            else:
                cb["value"] = cb["cold"]["value"]
            line["cpu_baseline"] = cb
            line["vs_cpu"] = line["value"] / cb["value"]
            line["vs_cpu_cold"] = line["value"] / cb["cold"]["value"]
            line["max_dT_K"] = cb["check_max_dT_K"]
            line["rms_dT_K"] = cb["check_rms_dT_K"]
        shape.close()
        if not args.no_serve:
            line["serving"] = serve_phase(ctx, args.serve_n, args.serve_rate, rps=args.serve_rps)
        if not args.no_geometry:
            line["illumination_producer"] = geometry_phase(ctx, cpu=not args.no_cpu)
        if not args.no_viewfactors:
            line["viewfactor_producer"] = viewfactor_phase(ctx, cpu=not args.no_cpu)
        if not args.no_observe:
            line["observe_fluxes"] = observe_phase(ctx, cpu=not args.no_cpu)
        line["roofline"]["valu_ceiling_wave_instr_per_s"] = valu_ceiling()
    if not args.no_binary:
        # a context of its own: its 40 GB crater sets are returned to the
        # driver when it closes instead of staying in the device cache
        with th.Context(0) as ctx2:
            line["binary_primary"] = binary_phase(ctx2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
