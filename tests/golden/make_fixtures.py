"""Writes the small data fixtures the CPU tests use, from data files that ship
in the reference tree (data only: meshes, view-factor lists, photon momentum
vectors, the mesh statistics file, colour corrections, light curves).  Run in a checkout that has the reference
at REF:

    python tests/golden/make_fixtures.py [REF]

Outputs (committed): tests/golden/new_crater2.npz,
tests/golden/doppler_binary.npz, tests/golden/doppler_binary_properties.json,
tests/golden/fg3_observe.npz (WISE colour corrections and the 1996 FG3 mesh),
tests/golden/fg3_flux_set4.npz (shipped 1996 FG3 light curves of set 4 at 10
smooth and 6 rough thermal inertias), tests/golden/fg3_flux_sets013.npz (the
shipped curves of observation sets 0, 1 and 3, smooth and rough at TI 0 and
100)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from binarytpm_amd import formats  # noqa: E402

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
D = os.path.join(REF, "3905", "Thermal", "NominalPole")


def main():
    V, F = formats.read_shape(os.path.join(D, "new_crater2.txt"), 61, 100)
    csr = formats.read_selfheating_list(os.path.join(D, "new_crater2_selfheating_list.txt"), 100)
    photon = np.loadtxt(os.path.join(D, "new_crater2_photovectors.txt"), dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "new_crater2.npz"), V=V, F=F, rowptr=csr[0],
                        colidx=csr[1], vf=csr[2], photon=photon)
    V, F = formats.read_shape(os.path.join(D, "doppler_binary.txt"), 3204, 6400)
    np.savez_compressed(os.path.join(HERE, "doppler_binary.npz"), V=V, F=F.astype(np.int32))
    R = os.path.join(REF, "175706", "Run-04")
    # 466 lines of three bands; the reference reads 467 (observe_generator.cpp:
    # 865-869), so its last row is the 0 a failed stream extraction stores
    cc = np.loadtxt(os.path.join(R, "wise_colour_corrections.txt"), dtype=np.float64)
    cc = np.vstack([cc, np.zeros((467 - len(cc), 3))]).T.copy()
    V, F = formats.read_shape(os.path.join(R, "1996FG3_binary_lowres.txt"), 2052, 4096)
    np.savez_compressed(os.path.join(HERE, "fg3_observe.npz"), colour=cc, V=V,
                        F=F.astype(np.int32))
    # the shipped light curves of observation set 4 (Observe's output files):
    # smooth at TI 0 and every 50 from 50 to 400 plus 10, rough at TI 0, 10,
    # 100, 200, 300 and 400
    flux = {}
    for kind, tis in (("smooth", (0, 10, 50, 100, 150, 200, 250, 300, 350, 400)),
                      ("rough", (0, 10, 100, 200, 300, 400))):
        for ti in tis:
            flux[f"{kind}_{ti}"] = np.loadtxt(
                os.path.join(R, f"1996FG3_binary_lowres_{kind}_flux({ti}_4).txt")).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "fg3_flux_set4.npz"), **flux)
    # observation sets 0, 1 and 3 (their Sun and observer positions are in
    # Run-00, Run-01 and Run-03; the curves ship in Run-04 with the others)
    flux = {}
    for obs_set in (0, 1, 3):
        for kind in ("smooth", "rough"):
            for ti in (0, 100):
                flux[f"{kind}_{ti}_set{obs_set}"] = np.loadtxt(os.path.join(
                    R, f"1996FG3_binary_lowres_{kind}_flux({ti}_{obs_set}).txt")).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "fg3_flux_sets013.npz"), **flux)
    props = {}
    with open(os.path.join(D, "doppler_binary_properties.txt")) as fh:
        for line in fh:
            # "Key:\tv [Key2:\tv2]" -- split every "name:" token pair
            parts = line.replace(":", ":\t").split("\t")
            key = None
            for tok in parts:
                tok = tok.strip()
                if not tok:
                    continue
                if tok.endswith(":"):
                    key = tok[:-1]
                    props[key] = []
                else:
                    props[key].append(float(tok))
    with open(os.path.join(HERE, "doppler_binary_properties.json"), "w") as fh:
        json.dump(props, fh, indent=1)
    print("fixtures written")


if __name__ == "__main__":
    main()
