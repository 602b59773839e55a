"""btpm_thermal: the Thermal stage at file level.

Reads the stage's inputs in the reference's file formats, runs a thermal-
inertia sweep on the GPU (all TIs batched) and writes the reference's
temperature files, which the Observe stage reads unchanged
(O3905/wise_main.cpp:222-244, 376-401).

    python -m binarytpm_amd.cli --shape DIR/name --nv NV --nf NF --ntau NTAU \
        --period SECONDS --ti 10 20 ... [--selfheating] [--crater DIR/new_crater2]

Inputs:  name.txt, name_illumination.dat (double[NV][3], int[NF][3],
float[NF][NTAU], the reference's binary_writeglobalillumination), optional
name_selfheating_list.txt; with --crater: new_crater2.txt,
new_crater2_selfheating_list.txt and name_roughness_illumination.dat
(float[NVc][3], int[NFc][3], float[NF][NFc][NTAU]); with both, also
name_selfheating_vector_list.txt, from which the shape->crater view factors
of the crater self-heating are computed on the GPU.
Outputs: name_temperature(TI).txt and .dat, name_roughness_temperature(TI).dat.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import formats
from . import thermal as th


def _ti_label(ti: float) -> str:
    """The reference's `stringstream << float` (6 significant digits)."""
    return f"{ti:g}"


def run(args) -> int:
    base = args.shape
    V, F = formats.read_shape(base + ".txt", args.nv, args.nf)
    _, _, illum = formats.read_temperature_dat(base + "_illumination.dat", args.nv, args.nf,
                                               args.ntau)
    csr = None
    if args.selfheating:
        csr = formats.read_selfheating_list(base + "_selfheating_list.txt", args.nf)
        if csr[0][-1] == 0:
            csr = None
    crater = None
    if args.crater:
        Vc, Fc = formats.read_shape(args.crater + ".txt", 61, 100)
        ccsr = formats.read_selfheating_list(args.crater + "_selfheating_list.txt", 100)
        cvf = formats.crater_vf_matrix(ccsr)
        _, _, cill = formats.read_roughness_dat(base + "_roughness_illumination.dat", 61, 100,
                                                args.nf, args.ntau)
        crater = th.CraterModel(vf=cvf, illum=np.ascontiguousarray(cill))
    p = th.params(args.ntau, args.period, tacc=args.tacc, eps=args.eps, ir_abs=args.ir_abs)
    numrevs = None
    if args.numrevs is not None:
        numrevs = np.loadtxt(args.numrevs, dtype=np.float32).reshape(-1)
    with th.Context(args.device) as ctx:
        if crater is not None and csr is not None:
            # shape->crater view factors from the facet lists and their
            # directions (calculate_surface_global_viewfactors,
            # T3905/main.cpp:1054-1181), on the GPU
            dirs = formats.read_selfheating_vectors(base + "_selfheating_vector_list.txt", csr[0])
            crater.drf = th.crater_viewfactors(ctx, V, F, csr[0], csr[2], dirs, Vc, Fc)
        shape = th.Shape(ctx, p, illum, args.albedo, numrevs=numrevs, csr=csr, scatter=0,
                         crater=crater)
        solves = [th.Solve(ctx, shape, ti, with_craters=crater is not None) for ti in args.ti]
        stats = th.batch_run(ctx, solves)
        for ti, s in zip(args.ti, solves):
            T = s.surface_temps()
            lab = _ti_label(ti)
            out = os.path.join(args.out, os.path.basename(base))
            th.write_temperature_txt(f"{out}_temperature({lab}).txt", V, F, T)
            formats.write_temperature_dat(f"{out}_temperature({lab}).dat", V, F, T)
            if crater is not None:
                s.write_crater_dat(f"{out}_roughness_temperature({lab}).dat", Vc, Fc)
            inf = s.info()
            print(f"TI {lab}: {inf.rotations} rotations (converged {inf.converged})"
                  + (f", crater {inf.crater_rotations} (converged {inf.crater_converged})"
                     if crater is not None else ""), flush=True)
            s.close()
        shape.close()
    print(f"{stats.solves} solves in {stats.wall_s:.2f} s", flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="btpm_thermal", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--shape", required=True, help="path prefix of the shape files")
    ap.add_argument("--nv", type=int, required=True)
    ap.add_argument("--nf", type=int, required=True)
    ap.add_argument("--ntau", type=int, required=True)
    ap.add_argument("--period", type=float, required=True, help="rotation period (s)")
    ap.add_argument("--ti", type=float, nargs="+", required=True)
    ap.add_argument("--albedo", type=float, default=0.1)
    ap.add_argument("--eps", type=float, default=0.95)
    ap.add_argument("--ir-abs", dest="ir_abs", type=float, default=1.0)
    ap.add_argument("--tacc", type=float, default=0.001)
    ap.add_argument("--numrevs", help="text file with one revolution count per facet")
    ap.add_argument("--selfheating", action="store_true")
    ap.add_argument("--crater", help="path prefix of the crater model files")
    ap.add_argument("--out", default=".")
    ap.add_argument("--device", type=int, default=0)
    return run(ap.parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
