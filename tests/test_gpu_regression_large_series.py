"""Crater series of more than 2^32 elements.

A dispatch's grid size is a 32-bit count of work-items, so an element-wise
pass over such a series must not start one work-item per element: the grid
would be cut short and most of the crater source left as whatever the
memory held before (earlier solves' data once device memory is reused).
Here 20480 facets x 100 sub-facets x 2400 steps (4.9e9 elements per series)
are solved three times on one context, the middle solve at another thermal
inertia so that the reused memory holds different values: the first and
third solve must agree bit for bit.
"""
import numpy as np
import pytest

from binarytpm_amd import inputs

pytestmark = pytest.mark.gpu


def test_crater_series_over_2_32_elements_repeat_bitwise():
    from binarytpm_amd import thermal as th

    cg = inputs.crater_cap(45.0)
    table = inputs.CraterTable(cg)
    case, sun_local = inputs.serving_shape(5, seed=5000, ntau=2400)
    nf, ntau = case.illum.shape
    assert nf * 100 * ((ntau + 7) // 8) * 8 > 2**32
    p = th.params(ntau, case.rot_period, tacc=0.01, max_rev=2)
    cm = th.CraterModel(vf=cg.vf, sun_local=sun_local, vis_table=table.vis, normals=cg.normals,
                        solar=inputs.solar_at(1.0), multiple=True, scatter_acc=1e-4)
    with th.Context(0) as ctx:
        shape = th.Shape(ctx, p, case.illum, 0.1, crater=cm)
        out = []
        for ti in (200.0, 1603.0, 200.0):
            s = th.Solve(ctx, shape, ti, with_craters=True)
            th.batch_run(ctx, [s], th.batch_opts(autotune=False))
            out.append(s.crater_temps() if ti == 200.0 else None)
            s.close()
        shape.close()
    first, third = out[0], out[2]
    assert np.isfinite(first).all() and (first > 0).all()
    # the last facets' columns lie far past element 2^32 of the blocked series
    assert float(first[-64:].min()) > 0
    assert np.array_equal(first, third)
