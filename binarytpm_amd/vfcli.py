"""btpm_selfheating: the Shape_Selfheating producer at file level.

Reads a shape file and writes the facet view-factor files the Thermal stage
reads, computed on the GPU (3905/Shape_Selfheating/main.cpp: input
`<name>.txt` with ASTNUMVTX vertex and ASTNUMFACE facet lines, :173-183;
outputs :837-888):

    python -m binarytpm_amd.vfcli --shape DIR/name --nv NV --nf NF [--nsub 10]

Outputs (next to the shape, or in --out): name_selfheating_list.txt,
name_selfheating_vector_list.txt, name_photovectors.txt,
name_selfheating_map.txt, name_total_viewfactor_map.txt.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import formats
from . import thermal as th


def run(args) -> int:
    V, F = formats.read_shape(args.shape + ".txt", args.nv, args.nf)
    if F.min() < 0 or F.max() >= args.nv:
        raise SystemExit(f"{args.shape}.txt: facet vertex index outside 0..{args.nv - 1}")
    t0 = time.time()
    with th.Context(args.device) as ctx:
        r = th.viewfactors(ctx, V, F, nsub=args.nsub, cull=not args.no_cull,
                           refine_min=args.refine_min)
    dt = time.time() - t0
    out = os.path.join(args.out, os.path.basename(args.shape)) if args.out else args.shape
    Vf = V.astype(np.float32)
    formats.write_facet_map(out + "_selfheating_map.txt", Vf, F, np.diff(r.rowptr))
    formats.write_facet_map(out + "_total_viewfactor_map.txt", Vf, F, r.total)
    formats.write_selfheating_list(out + "_selfheating_list.txt", r.csr)
    formats.write_selfheating_vectors(out + "_selfheating_vector_list.txt", r.rowptr, r.dir)
    formats.write_photovectors(out + "_photovectors.txt", r.photon)
    print(f"{args.nf} facets: {r.n_facing} facing pairs, {len(r.col)} after the occlusion cull, "
          f"{r.n_refined} refined, in {dt:.2f} s", flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="btpm_selfheating", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--shape", required=True, help="path prefix of the shape file")
    ap.add_argument("--nv", type=int, required=True)
    ap.add_argument("--nf", type=int, required=True)
    ap.add_argument("--nsub", type=int, default=10,
                    help="sub-facet split per edge of close pairs (reference n = 10)")
    ap.add_argument("--refine-min", dest="refine_min", type=float, default=0.05,
                    help="refine pairs whose centroid view factor is at least this")
    ap.add_argument("--no-cull", dest="no_cull", action="store_true",
                    help="keep pairs hidden by a third facet")
    ap.add_argument("--out", default="", help="output directory (default: next to the shape)")
    ap.add_argument("--device", type=int, default=0)
    return run(ap.parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
