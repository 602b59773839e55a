"""btpm_stitch: join the crater temperature files of a run split into parts.

The reference runs large shapes in NUMPARTS parts of NUMROUGHFACE facets and
joins `<shape>_roughness_temperature_part_<k>(<TI>).dat` into
`<shape>_roughness_temperature(<TI>).dat` (stitch_temperatures,
175706/Run-04/thermal.cpp:104-244); this does the same, streamed, for one or
several thermal inertias:

    python -m binarytpm_amd.stitch --shape DIR/name --parts 8 --ntau 5850 --ti 10 20 ...
"""
from __future__ import annotations

import argparse
import sys

from . import thermal as th


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="btpm_stitch", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--shape", required=True, help="path prefix of the part files")
    ap.add_argument("--parts", type=int, required=True)
    ap.add_argument("--ntau", type=int, required=True)
    ap.add_argument("--ti", type=float, nargs="+", required=True)
    ap.add_argument("--ncv", type=int, default=61, help="crater model vertices")
    ap.add_argument("--ncf", type=int, default=100, help="crater model facets")
    a = ap.parse_args(argv)
    for ti in a.ti:
        lab = f"{ti:g}"
        parts = [f"{a.shape}_roughness_temperature_part_{k}({lab}).dat"
                 for k in range(1, a.parts + 1)]
        th.stitch_crater_parts(f"{a.shape}_roughness_temperature({lab}).dat", parts, a.ncv,
                               a.ncf, a.ntau)
        print(f"TI {lab}: {a.parts} parts joined", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
