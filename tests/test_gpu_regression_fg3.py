"""End-to-end regression against the reference's own shipped results: the
1996 FG3 light curves of observation set 4 (175706/Run-04/
1996FG3_binary_lowres_{smooth,rough}_flux(TI_4).txt, fixture
tests/golden/fg3_flux_set4.npz), regenerated with this library alone at the
reference's configuration (175706/Run-04/thermal.cpp:17-80, visibility.cpp:
18-49, observe_generator.cpp:18-45):

  illumination and eclipses  geometry producer (Sun at the heliocentric
                             position, 9:2 revolutions over 5850 steps)
  crater sunlight            the reference's rim-sample partial shadow from
                             the producer's Sun vectors
  thermal solve              smooth facets and 409,600 crater columns to
                             TACC = 0.001 K (A = 0.01, emissivity 0.95, no
                             shape self-heating, multiple crater scattering)
  visibility                 the same producers with the observer as source
  fluxes                     Observe synthesis, smooth temperatures rounded to
                             the 6 digits of the text file Observe reads

The shipped values carry 6 significant digits, so the median relative
difference sits at their rounding (~2e-6; ~1e-5 at TI 300, where the
0.001 K convergence tolerance weighs more).  One or two steps in 5850 differ
by up to ~1e-3, most in band 1 and least in band 3: the pattern of the colour
correction of a neighbouring colour temperature (a near tie in the fit), whose
band-1 corrections differ by up to ~1 % per kelvin.
"""
import math
import os
import time

import numpy as np
import pytest

from binarytpm_amd import formats
from binarytpm_amd import thermal as th

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI = 4.0 * math.atan(1.0)
HEL = (1.160088, (-1.9521 / 180.0) * PI, (198.3142 / 180.0) * PI)
OBS = (0.583905, (-3.882500 / 180.0) * PI, (257.492321 / 180.0) * PI)
ROT = ((-83.0 / 180.0) * PI, (266.0 / 180.0) * PI)
KW = dict(nv_primary=1026, rev_primary=9.0, rev_secondary=2.0)
NTAU = 5850


def _check(name, ours, ref):
    rel = np.abs(ours - ref) / ref
    med, p999, mx = np.median(rel, axis=0), np.percentile(rel, 99.9, axis=0), rel.max(axis=0)
    off = int((rel.max(axis=1) > 1e-4).sum())
    print(f"[regression] {name}: median {np.array2string(med, precision=1)}, p99.9 "
          f"{np.array2string(p999, precision=1)}, max {np.array2string(mx, precision=1)}, "
          f"steps beyond 1e-4: {off} of {len(ref)}", flush=True)
    assert (med < 3e-5).all() and (p999 < 1e-4).all() and (mx < 3e-3).all() and off <= 12


def _text6(T):
    """The temperatures as Observe reads them back from the 6-digit text file."""
    return np.array([float(f"{x:g}") for x in T.reshape(-1)], np.float64).reshape(T.shape)


@pytest.fixture(scope="module")
def fg3(ctx):
    d = np.load(os.path.join(GOLD, "fg3_observe.npz"))
    c = np.load(os.path.join(GOLD, "new_crater2.npz"))
    V, F, cc = d["V"], d["F"], d["colour"]
    t0 = time.time()
    sun = th.geometry_illumination(ctx, V, F, NTAU, HEL, ROT, sun_scale=10000.0, **KW)
    seen = th.geometry_illumination(ctx, V, F, NTAU, OBS, ROT, sun_scale=10000.0, **KW)
    print(f"[regression] geometry (Sun, observer): {time.time() - t0:.1f} s", flush=True)
    period = float(np.float32(2.0) * np.float32(16.15076) * np.float32(60.0) * np.float32(60.0))
    nf = len(F)
    numrevs = np.concatenate([np.full(2048, 9.0), np.full(nf - 2048, 2.0)]).astype(np.float32)
    return dict(V=V, F=F, cc=cc, c=c, sun=sun, seen=seen, numrevs=numrevs,
                p=th.params(NTAU, period, tacc=0.001, eps=0.95))


def test_fg3_smooth_light_curves(ctx, fg3):
    ref = np.load(os.path.join(GOLD, "fg3_flux_set4.npz"))
    shape = th.Shape(ctx, fg3["p"], fg3["sun"].illum, 0.01, numrevs=fg3["numrevs"], scatter=0)
    try:
        for ti in (10.0, 100.0, 300.0):
            s = th.Solve(ctx, shape, ti)
            th.batch_run(ctx, [s])
            inf = s.info()
            sm, _ = th.observe(ctx, fg3["V"], fg3["F"], NTAU, HEL, OBS, ROT, fg3["seen"].shadow,
                               _text6(s.surface_temps()), fg3["cc"], **KW)
            s.close()
            assert inf.converged
            _check(f"smooth TI {ti:g} ({inf.rotations} rotations)", sm,
                   ref[f"smooth_{int(ti)}"].astype(np.float64))
    finally:
        shape.close()


def test_fg3_rough_light_curve(ctx, fg3):
    """The crater path end to end: 4096 facets x 100 crater sub-facets."""
    ref = np.load(os.path.join(GOLD, "fg3_flux_set4.npz"))
    c = fg3["c"]
    cV, cF = c["V"].astype(np.float32), c["F"].astype(np.int32)
    sun, seen = fg3["sun"], fg3["seen"]
    csh, ccos = th.crater_sunlight(ctx, sun.shadow, sun.sun_facet, cV, cF)
    cvf = formats.crater_vf_matrix((c["rowptr"], c["colidx"], c["vf"]))
    cm = th.CraterModel(vf=cvf, shadow=csh, cosang=ccos, solar=sun.solar, multiple=True,
                        scatter_acc=1e-4)
    shape = th.Shape(ctx, fg3["p"], sun.illum, 0.01, numrevs=fg3["numrevs"], scatter=0, crater=cm)
    del csh, ccos, cm
    try:
        s = th.Solve(ctx, shape, 100.0, with_craters=True)
        t0 = time.time()
        th.batch_run(ctx, [s])
        inf = s.info()
        print(f"[regression] TI 100 solve: {inf.rotations} smooth and {inf.crater_rotations} "
              f"crater rotations in {time.time() - t0:.1f} s", flush=True)
        assert inf.converged and inf.crater_converged
        T6 = _text6(s.surface_temps())
        CT = s.crater_temps()
        s.close()
        cvis, _ = th.crater_sunlight(ctx, seen.shadow, seen.sun_facet, cV, cF)
        sm, ro = th.observe(ctx, fg3["V"], fg3["F"], NTAU, HEL, OBS, ROT, seen.shadow, T6,
                            fg3["cc"], crater=(cV, cF, cvis, CT), **KW)
        _check("rough TI 100", ro, ref["rough_100"].astype(np.float64))
        _check("smooth TI 100 (with the crater model)", sm, ref["smooth_100"].astype(np.float64))
    finally:
        shape.close()
