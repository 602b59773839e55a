"""Facet self-heating kernel time per kernel on two view-factor patterns of a
5120-facet shape (4 smooth solves batched, TI 50 / 100 / 200 / 400, 650 steps):

  uniform  154 random emitters per receiver (the benchmark's pattern)
  skewed   20 random emitters per receiver, and every 97th receiver sees
           every other facet (the pattern of concave shapes, where a few
           facets see most of the body)

for each kernel forced through BTPM_SH_LAYOUT: csr (lane per receiver over
the per-chunk CSR), ell / ell32 (lane per receiver over the lane-interleaved
ELL, 64 / 32 steps per workgroup), grp (8 receivers per wave over the
receiver-grouped layout) and rows (lane per step over the per-chunk CSR).  Prints the average kernel time per launch
from the library's own event timing.

Usage (one GPU): python -m binarytpm_amd.bench_tools.selfheat_layouts
[--pattern uniform|skewed] [--layouts csr,ell,...] (default: both patterns,
every layout)
"""
import argparse
import os

import numpy as np

from binarytpm_amd import inputs
from binarytpm_amd import thermal as th


def skewed_viewfactors(nf, seed=11, every=97, k=20):
    rng = np.random.default_rng(seed)
    rowptr, cols, vals = [0], [], []
    for f in range(nf):
        if f % every == 0:
            c = np.delete(np.arange(nf), f)
            v = np.full(c.size, 0.5 / c.size)
        else:
            c = np.sort(rng.choice(nf - 1, size=k, replace=False))
            c = np.where(c >= f, c + 1, c)
            v = rng.uniform(0.001, 0.02, size=k)
        cols.append(c)
        vals.append(v)
        rowptr.append(rowptr[-1] + c.size)
    return (np.asarray(rowptr, np.int32), np.concatenate(cols).astype(np.int32),
            np.concatenate(vals).astype(np.float32))


LAYOUTS = ("csr", "ell", "ell32", "grp", "rows")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pattern", choices=("uniform", "skewed"), default=None)
    ap.add_argument("--layouts", default=",".join(LAYOUTS))
    args = ap.parse_args()
    layouts = [x for x in args.layouts.split(",") if x]
    for x in layouts:
        if x not in LAYOUTS:
            ap.error(f"unknown layout {x}")
    case, _ = inputs.serving_shape(4, 7)
    nf = case.illum.shape[0]
    p = th.params(case.ntau, case.rot_period, tacc=0.01)
    patterns = (("uniform", lambda: inputs.sparse_viewfactors(nf, 0.03, 0.2, 1)[0]),
                ("skewed", lambda: skewed_viewfactors(nf)))
    for name, make in patterns:
        if args.pattern and name != args.pattern:
            continue
        csr = make()
        for lay in layouts:
            os.environ["BTPM_SH_LAYOUT"] = lay
            with th.Context(0) as ctx:
                shape = th.Shape(ctx, p, case.illum, 0.1, csr=csr, scatter=0)
                solves = [th.Solve(ctx, shape, ti) for ti in (50.0, 100.0, 200.0, 400.0)]
                st = th.batch_run(ctx, solves, th.batch_opts(autotune=False, profile=True))
                ms = 1e3 * st.heat_kernel_s / max(st.heat_launches, 1)
                print(f"{name:8s} nnz {len(csr[1]):9d} {lay:5s}: {ms:.3f} ms per launch "
                      f"({st.heat_launches} launches)", flush=True)
                for s in solves:
                    s.close()
                shape.close()
    os.environ.pop("BTPM_SH_LAYOUT", None)


if __name__ == "__main__":
    main()
