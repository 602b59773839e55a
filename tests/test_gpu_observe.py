"""GPU parity of the Observe flux synthesis (btpm_observe, csrc/observe.hip)
against the FP64 restatement (oracle/observe.c).  The device evaluates exp,
acos and cos with its own libm and sums each facet's crater sub-facets across
lanes: fluxes agree within 1e-12 relative (and so does the colour temperature
each step's fit picks, since the outputs carry its correction)."""
import math
import os
import time

import numpy as np
import pytest

from binarytpm_amd import inputs
from binarytpm_amd import thermal as th
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-12

# the reference's configuration (175706/Run-04/observe_generator.cpp:18-35)
HEL = (1.160088, -1.9521 / 180.0 * math.pi, 198.3142 / 180.0 * math.pi)
OBS = (0.583905, -3.882500 / 180.0 * math.pi, 257.492321 / 180.0 * math.pi)
ROT = (-83.0 / 180.0 * math.pi, 266.0 / 180.0 * math.pi)


def _inputs(nf, ntau, ncf, seed):
    rng = np.random.default_rng(seed)
    vis = (rng.random((nf, ntau)) > 0.2).astype(np.int32)
    st = (150.0 + 250.0 * rng.random((nf, ntau))).astype(np.float32)
    cvis = rng.random((nf, ntau, ncf)).astype(np.float32)
    cvis[cvis < 0.3] = 0.0
    ct = (120.0 + 300.0 * rng.random((nf, ntau, ncf))).astype(np.float32)
    return vis, st, cvis, ct


def _check(name, g, o):
    for k, (a, b) in enumerate(zip(g, o)):
        rel = float((np.abs(a - b) / np.maximum(np.abs(b), 1e-300)).max())
        print(f"[parity] {name} {'smooth rough'.split()[k]}: max rel {rel:.1e}", flush=True)
        assert rel <= RTOL


def test_observe_binary_all_steps(ctx):
    """A dented binary, 650 steps, random visibility and temperatures, crater
    model with partially visible sub-facets: every step."""
    equ, fv, npv = inputs.binary_shape(2, 1, sep=1.3, dent=0.3)
    c = np.load(os.path.join(GOLD, "new_crater2.npz"))
    cc = np.load(os.path.join(GOLD, "fg3_observe.npz"))["colour"]
    ntau = 650
    vis, st, cvis, ct = _inputs(len(fv), ntau, 100, 1)
    kw = dict(nv_primary=npv, rev_primary=3.0, rev_secondary=1.0)
    crater = (c["V"], c["F"], cvis, ct)
    g = th.observe(ctx, equ, fv, ntau, HEL, OBS, ROT, vis, st, cc, crater=crater, **kw)
    o = O.observe(equ, fv, ntau, HEL, OBS, ROT, vis, st, cc, crater=crater, **kw)
    _check("binary 650 steps", g, o)
    # without the crater model: the rough output is zero flux
    g2 = th.observe(ctx, equ, fv, ntau, HEL, OBS, ROT, vis, st, cc, **kw)
    assert np.array_equal(g2[0], g[0]) and not g2[1].any()


def test_observe_fg3_full_cycle(ctx):
    """The reference's observation: the 4096-facet 1996 FG3 mesh, 9:2
    revolutions over 5850 steps (smooth model), and with the crater model
    over 650 steps; the restatement recomputes sampled steps."""
    d = np.load(os.path.join(GOLD, "fg3_observe.npz"))
    V, F, cc = d["V"], d["F"], d["colour"]
    kw = dict(nv_primary=1026, rev_primary=9.0, rev_secondary=2.0)
    ntau = 5850
    vis, st, _, _ = _inputs(len(F), ntau, 1, 2)
    t0 = time.time()
    g = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, st, cc, **kw)
    print(f"[observe] fg3 smooth, {len(F)} facets x {ntau} steps: {time.time() - t0:.2f} s",
          flush=True)
    for s in (0, 1111, 2925, 5849):
        o = O.observe(V, F, ntau, HEL, OBS, ROT, vis, st, cc, steps=(s, 1), **kw)
        _check(f"fg3 smooth step {s}", (g[0][s:s + 1], g[1][s:s + 1]), o)
    ntau = 650
    c = np.load(os.path.join(GOLD, "new_crater2.npz"))
    vis, st, cvis, ct = _inputs(len(F), ntau, 100, 3)
    crater = (c["V"], c["F"], cvis, ct)
    t0 = time.time()
    g = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, st, cc, crater=crater, **kw)
    print(f"[observe] fg3 rough, {len(F)} facets x {ntau} steps x 100: {time.time() - t0:.2f} s",
          flush=True)
    for s in (0, 333, 649):
        o = O.observe(V, F, ntau, HEL, OBS, ROT, vis, st, cc, crater=crater, steps=(s, 1), **kw)
        _check(f"fg3 rough step {s}", (g[0][s:s + 1], g[1][s:s + 1]), o)


def test_observe_from_solve_on_device(ctx):
    """btpm_observe_solve: the Observe synthesis reading a finished crater
    solve's temperatures in device memory equals btpm_observe on that solve's
    readbacks bit for bit (both floored at +0 K, same kernels), with and
    without the crater model, and with the crater visibility computed on the
    device from the observer's directions; an unfinished solve, a solve of
    another step count and a crater model that is not the solver's 100
    sub-facets are refused."""
    case, crater = inputs.rough_primary(subdiv=1)      # 80 facets x 100 sub-facets
    V, F = inputs.ellipsoid(1, (0.6, 0.5, 0.45))
    ntau = case.ntau
    c = np.load(os.path.join(GOLD, "new_crater2.npz"))
    cc = np.load(os.path.join(GOLD, "fg3_observe.npz"))["colour"]
    p = th.params(ntau, case.rot_period, tacc=0.01)
    cm = th.CraterModel(vf=crater["vf"], shadow=crater["shadow"], cosang=crater["cosang"],
                        solar=crater["solar"], multiple=True, scatter_acc=1e-4,
                        drf=crater["drf"])
    shape = th.Shape(ctx, p, case.illum, 0.1, csr=case.csr, scatter=2, scatter_acc=1e-4,
                     crater=cm)
    s = th.Solve(ctx, shape, 150.0, with_craters=True).run()
    vis, _, cvis, _ = _inputs(len(F), ntau, 100, 3)
    st, ct = s.surface_temps(), s.crater_temps()
    kw = dict(nv_primary=len(V))
    t0 = time.perf_counter()
    host = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, st, cc,
                      crater=(c["V"], c["F"], cvis, ct), **kw)
    t1 = time.perf_counter()
    dev = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, None, cc,
                     crater=(c["V"], c["F"], cvis, None), solve=s, **kw)
    t2 = time.perf_counter()
    print(f"[observe] from readbacks {t1 - t0:.3f} s, from the solve {t2 - t1:.3f} s", flush=True)
    assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
    assert np.isfinite(dev[1]).all() and dev[1].any()
    sm_only = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, None, cc, solve=s, **kw)
    assert np.array_equal(sm_only[0], host[0]) and not sm_only[1].any()
    # crater visibility from the observer's direction in each facet frame,
    # computed on the device: equals crater_sunlight's fractions fed to the
    # host path
    rng = np.random.default_rng(9)
    ob = rng.normal(size=(len(F), ntau, 3)).astype(np.float32)
    ob /= np.linalg.norm(ob, axis=2, keepdims=True)
    cvis2, _ = th.crater_sunlight(ctx, vis, ob, c["V"], c["F"])
    host2 = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, st, cc,
                       crater=(c["V"], c["F"], cvis2, ct), **kw)
    dev2 = th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, None, cc,
                      crater=(c["V"], c["F"], None, None), solve=s, observer_facet=ob, **kw)
    assert np.array_equal(host2[0], dev2[0]) and np.array_equal(host2[1], dev2[1])
    assert dev2[1].any() and not np.array_equal(dev2[1], dev[1])
    pending = th.Solve(ctx, shape, 300.0, with_craters=True)
    with pytest.raises(th.BtpmError):
        th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, None, cc, solve=pending, **kw)
    with pytest.raises(th.BtpmError):
        th.observe(ctx, V, F, ntau - 10, HEL, OBS, ROT, vis[:, :ntau - 10], None, cc,
                   solve=s, **kw)
    with pytest.raises(th.BtpmError):
        th.observe(ctx, V, F, ntau, HEL, OBS, ROT, vis, None, cc,
                   crater=(c["V"], c["F"][:50], cvis[:, :, :50], None), solve=s, **kw)
    pending.close()
    s.close()
    shape.close()
