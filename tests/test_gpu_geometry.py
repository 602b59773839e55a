"""GPU parity of the illumination geometry and shadow producer
(btpm_geometry_illumination) against the FP64 oracle (oracle/geometry.c):
shadow flags bit for bit (integers), direct illumination within float
rounding of cos(angle), Sun directions in the facet frames within 2e-6."""
import math
import time

import numpy as np
import pytest

from binarytpm_amd import inputs
from binarytpm_amd import thermal as th
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _compare(name, gpu, ref):
    mism = int((gpu.shadow != ref.shadow).sum())
    di = float(np.abs(gpu.illum - ref.illum).max())
    print(f"[parity] {name}: shadow mismatches {mism} of {ref.shadow.size}, "
          f"max |d illum| {di:.2e} W/m^2, lit {ref.shadow.mean():.3f}", flush=True)
    assert mism == 0
    assert di <= 1e-3
    if gpu.sun_facet is not None:
        assert float(np.abs(gpu.sun_facet - ref.sun_facet).max()) <= 2e-6


def test_geometry_parity_dented_binary(ctx):
    """Self-shadowing in the troughs of a dented primary and mutual eclipses
    with a close secondary (Sun in the orbit plane), 600 steps of a 3:1 cycle."""
    equ, fv, npv = inputs.binary_shape(2, 1, sep=1.2, dent=0.35)
    kw = dict(nv_primary=npv, rev_primary=3.0, rev_secondary=1.0)
    pos, pole = (1.0, 0.0, 0.0), (math.pi / 2, 0.0)
    gpu = th.geometry_illumination(ctx, equ, fv, 600, pos, pole, **kw)
    ref = O.geometry(equ, fv, 600, pos, pole, **kw)
    assert gpu.solar == ref.solar
    assert int((ref.cosang >= 0).sum()) - int(ref.shadow.sum()) > 1000  # shadows beyond back faces
    _compare("dented binary, 600 steps", gpu, ref)


def test_geometry_parity_reference_pole_and_position(ctx):
    """The reference's own pole and heliocentric position
    (T175706/thermal.cpp:28-33), a dented single body over 130 steps."""
    equ, fv, npv = inputs.binary_shape(3, 0, sep=4.0, dent=0.3)
    pos = (1.160088, -1.9521 / 180.0 * math.pi, 198.3142 / 180.0 * math.pi)
    pole = (-83.0 / 180.0 * math.pi, 266.0 / 180.0 * math.pi)
    kw = dict(nv_primary=npv, rev_primary=9.0, rev_secondary=2.0)
    gpu = th.geometry_illumination(ctx, equ, fv, 130, pos, pole, **kw)
    ref = O.geometry(equ, fv, 130, pos, pole, **kw)
    _compare("reference pole/position, 130 steps", gpu, ref)


def test_geometry_config1_full_size(ctx):
    """Config 1 size: the 5120-facet primary and a 1280-facet secondary at
    3 km, 15:4 super-cycle over 9750 steps, on the GPU in full; the oracle
    recomputes three steps (6400^2 facet pairs each) for a bit-exact check."""
    equ, fv, npv = inputs.binary_shape(4, 3, axes_p=(0.6, 0.5, 0.45), radius_s=0.3, sep=3.0)
    kw = dict(nv_primary=npv, rev_primary=15.0, rev_secondary=4.0)
    pos, pole = (1.1, 0.0, 0.0), (math.pi / 2, 0.0)
    t0 = time.time()
    gpu = th.geometry_illumination(ctx, equ, fv, 9750, pos, pole, sun_facet=False, **kw)
    t_gpu = time.time() - t0
    eclipsed = int((gpu.illum == 0).sum())
    print(f"[geometry] config 1: {len(fv)} facets x 9750 steps in {t_gpu:.2f} s "
          f"(including transfers), {eclipsed} unlit facet-steps", flush=True)
    for s in (0, 2437, 7311):
        t0 = time.time()
        ref = O.geometry(equ, fv, 9750, pos, pole, steps=(s, 1), **kw)
        t_ref = time.time() - t0
        mism = int((gpu.shadow[:, s] != ref.shadow[:, 0]).sum())
        di = float(np.abs(gpu.illum[:, s] - ref.illum[:, 0]).max())
        print(f"[parity] config 1 step {s}: shadow mismatches {mism}, max |d illum| {di:.2e}, "
              f"oracle {t_ref:.2f} s for the step", flush=True)
        assert mism == 0 and di <= 1e-3


def test_crater_sunlight_parity(ctx):
    """Crater sunlight (btpm_crater_sunlight) bit for bit against the
    restatement: random Sun vectors (zenith and below-horizon ones included)
    and lit flags, then the chain from the geometry producer's
    sfcillumvector (sun_scale 10000) on a dented binary over 650 steps."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    c = np.load(os.path.join(gold, "new_crater2.npz"))
    cv, cf = c["V"].astype(np.float32), c["F"].astype(np.int32)
    rng = np.random.default_rng(4)
    nf, nt = 300, 40
    d = rng.standard_normal((nf, nt, 3))
    d[0, 0] = (0.0, 0.0, 1.0)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    sun = (10000.0 * d).astype(np.float32)
    lit = (rng.random((nf, nt)) > 0.3).astype(np.int32)
    g = th.crater_sunlight(ctx, lit, sun, cv, cf)
    o = O.crater_sun(lit, sun, cv, cf)
    for name, a, b in (("shadow", g[0], o[0]), ("cosang", g[1], o[1])):
        mism = int((a != b).sum())
        print(f"[parity] crater sunlight random {name}: {mism} mismatches of {a.size}", flush=True)
        assert mism == 0
    assert 0.2 < float((o[0] > 0).mean()) < 0.6 and float(((o[0] > 0) & (o[0] < 1)).mean()) > 0.05
    equ, fv, npv = inputs.binary_shape(2, 1, sep=1.2, dent=0.35)
    kw = dict(nv_primary=npv, rev_primary=3.0, rev_secondary=1.0)
    pos, pole = (1.0, 0.3, 0.0), (math.pi / 3, 0.4)
    gg = th.geometry_illumination(ctx, equ, fv, 650, pos, pole, sun_scale=10000.0, **kw)
    og = O.geometry(equ, fv, 650, pos, pole, sun_scale=10000.0, **kw)
    assert np.array_equal(gg.shadow, og.shadow)
    assert float(np.abs(gg.sun_facet - og.sun_facet).max()) <= 0.02   # 10000 x (2e-6)
    g = th.crater_sunlight(ctx, gg.shadow, gg.sun_facet, cv, cf)
    o = O.crater_sun(gg.shadow, gg.sun_facet, cv, cf)
    mism = int((g[0] != o[0]).sum()) + int((g[1] != o[1]).sum())
    print(f"[parity] crater sunlight chain: {mism} mismatches of {2 * g[0].size}", flush=True)
    assert mism == 0
