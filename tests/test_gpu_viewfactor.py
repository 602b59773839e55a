"""GPU parity of the view-factor producers (binarytpm_amd/csrc/viewfactor.hip)
against the float restatement (oracle/viewfactor.c).

Facet view factors (btpm_vf_build): the lists (row pointers, partners,
refined flags) bit for bit; refined view factors, their vectors and the
vector-list directions bit for bit (IEEE + - * / sqrt in the reference's
order on both sides); centroid view factors within a few 1e-7 of the
pair's geometric factor A_J / (pi d^2) (they go through cosf(acosf(c)),
evaluated by the device's libm), totals and photon vectors within the sum of
those bounds.  Shape->crater view factors (btpm_crater_viewfactors) bit for
bit.
"""
import os
import time

import numpy as np
import pytest

from binarytpm_amd import inputs
from binarytpm_amd import thermal as th
from binarytpm_amd._lib import BtpmError
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Centroid view factors go through cosf(acosf(c)) on two libms: each cosine
# may differ by a few 1e-8 absolute (the rounding of acos near pi/2), so the
# bound is relative to the geometric factor A_J / (pi d^2) of the pair plus a
# few ulps of the value.
COS_ATOL = 6e-7
RTOL = 4e-6


def _geom_factor(V, F, rowptr, col):
    """A_J / (pi d^2) per list entry (FP64 from the float mesh)."""
    V = np.asarray(V, np.float32).astype(np.float64)
    F = np.asarray(F)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    mid = (p0 + p1 + p2) / 3.0
    area = np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1) / 2.0
    rows = np.repeat(np.arange(len(F)), np.diff(rowptr))
    d2 = ((mid[col] - mid[rows]) ** 2).sum(axis=1)
    return area[col] / (np.pi * d2), rows


def _compare(name, g, o, V, F):
    assert np.array_equal(g.rowptr, o.rowptr), name
    assert np.array_equal(g.col, o.col), name
    assert g.n_facing == o.n_facing and g.n_refined == int(o.refined.sum()), name
    ref = o.refined
    exact_vf = bool(np.array_equal(g.vf[ref], o.vf[ref]))
    exact_vec = bool(np.array_equal(g.vec[ref], o.vec[ref]))
    gf, rows = _geom_factor(V, F, o.rowptr, o.col)
    tol = COS_ATOL * gf + RTOL * np.abs(o.vf)
    err = np.abs(g.vf.astype(np.float64) - o.vf)
    worst = float((err / tol).max()) if len(err) else 0.0
    # row sums: the per-entry bounds add up
    row_tol = np.bincount(rows, weights=tol, minlength=len(F)) + 1e-12
    tot_worst = float((np.abs(g.total.astype(np.float64) - o.total) / row_tol).max())
    ph_worst = float((np.abs(g.photon.astype(np.float64) - o.photon).max(axis=1) / row_tol).max())
    print(f"[parity] {name}: {len(o.col)} entries ({o.n_facing} facing, {int(ref.sum())} "
          f"refined): refined bitwise {exact_vf}/{exact_vec}, centroid error / bound "
          f"{worst:.2f}, total {tot_worst:.2f}, photon {ph_worst:.2f}", flush=True)
    assert exact_vf and exact_vec
    assert worst <= 1.0 and tot_worst <= 1.0 and ph_worst <= 1.0
    # vector-list directions: unit vectors (up to rounding)
    big = gf > 1e-12
    assert np.allclose(g.dir[big], o.dir[big], atol=2e-5)
    if ref.any():
        assert np.array_equal(g.dir[ref], o.dir[ref])


def test_vf_crater_fixture(ctx):
    """new_crater2 with every pair refined (as its shipped list): the GPU
    equals the restatement bit for bit and the shipped file to its printed
    precision; with the reference threshold (no pair reaches 0.05) the
    centroid values agree within RTOL."""
    d = np.load(os.path.join(GOLD, "new_crater2.npz"))
    g = th.viewfactors(ctx, d["V"], d["F"], nsub=10, refine_min=0.0)
    o = O.viewfactors(d["V"], d["F"], nsub=10, refine_min=0.0)
    _compare("crater, all refined", g, o, d["V"], d["F"])
    assert np.array_equal(g.vf, o.vf) and np.array_equal(g.total, o.total)
    assert np.array_equal(g.photon, o.photon)
    rel = np.abs(g.vf - d["vf"]) / d["vf"]
    assert rel.max() < 1e-4 and np.abs(g.photon - d["photon"]).max() < 2e-6
    g0 = th.viewfactors(ctx, d["V"], d["F"])
    o0 = O.viewfactors(d["V"], d["F"])
    _compare("crater, reference threshold", g0, o0, d["V"], d["F"])


@pytest.mark.parametrize("seed,cull,nsub", [(0, True, 10), (1, True, 10), (2, False, 10),
                                            (3, True, 0), (4, True, 3)])
def test_vf_occlusion_scene(ctx, seed, cull, nsub):
    """Floor, ceiling and blockers: occlusion cull, coplanar (jittered)
    neighbours at the facing threshold, refinement of the close pairs."""
    V, F = inputs.occlusion_scene(seed=seed)
    g = th.viewfactors(ctx, V, F, nsub=nsub, cull=cull)
    o = O.viewfactors(V, F, nsub=nsub, cull=cull)
    _compare(f"scene seed {seed} cull {cull} nsub {nsub}", g, o, V, F)
    if cull:
        assert o.n_facing - len(o.col) > 1000


def test_vf_scene_all_refined_bitwise(ctx):
    """Every pair refined (nsub 10): the whole list is bit-exact."""
    V, F = inputs.occlusion_scene(seed=5)
    g = th.viewfactors(ctx, V, F, nsub=10, refine_min=0.0)
    o = O.viewfactors(V, F, nsub=10, refine_min=0.0)
    _compare("scene all refined", g, o, V, F)
    for k in ("vf", "vec", "dir", "total", "photon"):
        assert np.array_equal(getattr(g, k), getattr(o, k)), k


def test_vf_doppler_binary_full_size(ctx):
    """The reference's own producer input: the 6400-facet Doppler binary
    (3905/Thermal/NominalPole/doppler_binary.txt) with n = 10."""
    d = np.load(os.path.join(GOLD, "doppler_binary.npz"))
    t0 = time.time()
    g = th.viewfactors(ctx, d["V"], d["F"])
    tg = time.time() - t0
    t0 = time.time()
    o = O.viewfactors(d["V"], d["F"])
    to = time.time() - t0
    print(f"[viewfactors] doppler_binary: GPU {tg:.2f} s (incl. transfers), oracle {to:.2f} s",
          flush=True)
    _compare("doppler_binary", g, o, d["V"], d["F"])
    assert len(o.col) > 4_000_000


def test_vf_edge_cases(ctx):
    """One facet; a degenerate facet (repeated vertex: no normal, faces
    nothing); two coaxial parallel triangles (cosines at exactly 1 or just
    above it, where the reference's acos gives NaN); the largest split."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    g = th.viewfactors(ctx, tri, np.array([[0, 1, 2]]))
    assert g.rowptr.tolist() == [0, 0] and len(g.col) == 0 and g.total[0] == 0.0
    V = np.vstack([tri, tri[:, [0, 1, 2]] + [0, 0, 1], [[0.3, 0.3, 0.5]]]).astype(np.float32)
    F = np.array([[0, 1, 2], [3, 5, 4], [0, 1, 1]], np.int32)
    for nsub in (0, 10, 32):
        g = th.viewfactors(ctx, V, F, nsub=nsub)
        o = O.viewfactors(V, F, nsub=nsub)
        _compare(f"edge cases nsub {nsub}", g, o, V, F)
    assert not np.isin(2, o.col) and o.rowptr[3] == o.rowptr[2]
    for bad in (dict(nsub=33), dict(nsub=-1)):
        with pytest.raises(BtpmError):
            th.viewfactors(ctx, V, F, **bad)
    with pytest.raises(BtpmError):
        th.viewfactors(ctx, V, np.array([[0, 1, 7]]))


def _crater_model():
    d = np.load(os.path.join(GOLD, "new_crater2.npz"))
    return d["V"].astype(np.float32), d["F"].astype(np.int32)


def test_drf_scene_bitwise(ctx):
    """Shape->crater view factors of the occlusion scene's lists against the
    shipped crater model: every entry, bit for bit."""
    V, F = inputs.occlusion_scene(seed=1)
    vf = th.viewfactors(ctx, V, F)
    cv, cf = _crater_model()
    g = th.crater_viewfactors(ctx, V, F, vf.rowptr, vf.vf, vf.dir, cv, cf)
    o = O.crater_drf(V, F, vf.rowptr, vf.vf, vf.dir, cv, cf)
    nz = int((o > 0).sum())
    print(f"[parity] drf scene: {o.shape[0]} entries x {o.shape[1]}, nonzero {nz}, "
          f"mismatches {int((g != o).sum())}", flush=True)
    assert nz > 1000 and np.array_equal(g, o)


def test_drf_doppler_binary(ctx):
    """Full size: the Doppler binary's lists (4.1 M entries) x 100 crater
    sub-facets on the GPU; the restatement recomputes 64 rows spread over
    the shape and must agree bit for bit."""
    d = np.load(os.path.join(GOLD, "doppler_binary.npz"))
    V, F = d["V"], d["F"]
    vf = th.viewfactors(ctx, V, F)
    cv, cf = _crater_model()
    t0 = time.time()
    g = th.crater_viewfactors(ctx, V, F, vf.rowptr, vf.vf, vf.dir, cv, cf)
    tg = time.time() - t0
    rows = np.flatnonzero(np.diff(vf.rowptr) > 0)
    pick = rows[np.linspace(0, len(rows) - 1, 64).astype(int)]
    mism = 0
    for f in pick:
        a, b = vf.rowptr[f], vf.rowptr[f + 1]
        o = O.crater_drf(V, F[f:f + 1], np.array([0, b - a], np.int32), vf.vf[a:b], vf.dir[a:b],
                         cv, cf)
        mism += int((g[a:b] != o).sum())
    print(f"[parity] drf doppler_binary: {g.shape[0]} x {g.shape[1]} in {tg:.2f} s "
          f"(incl. transfers), 64 rows checked, mismatches {mism}", flush=True)
    assert mism == 0


def _write_shape(path, V, F):
    with open(path, "w") as fh:
        for v in V:
            fh.write(f"{v[0]:.9g}\t{v[1]:.9g}\t{v[2]:.9g}\n")
        for f in F:
            fh.write(f"{f[0]}\t{f[1]}\t{f[2]}\n")


def test_selfheating_cli_reproduces_shipped_crater_files(tmp_path):
    """btpm_selfheating on the shipped crater mesh (every pair refined, as
    its shipped list) writes the reference's list and photon vectors."""
    from binarytpm_amd import formats, vfcli
    d = np.load(os.path.join(GOLD, "new_crater2.npz"))
    base = str(tmp_path / "new_crater2")
    _write_shape(base + ".txt", d["V"], d["F"])
    assert vfcli.main(["--shape", base, "--nv", "61", "--nf", "100", "--refine-min", "0"]) == 0
    rp, ci, vf = formats.read_selfheating_list(base + "_selfheating_list.txt", 100)
    assert np.array_equal(rp, d["rowptr"]) and np.array_equal(ci, d["colidx"])
    assert (np.abs(vf - d["vf"]) / d["vf"]).max() < 1e-4
    ph = np.loadtxt(base + "_photovectors.txt")
    assert np.abs(ph - d["photon"]).max() < 2e-6
    dirs = formats.read_selfheating_vectors(base + "_selfheating_vector_list.txt", rp)
    # refined pairs: the VF-weighted mean of the sub-pair unit directions
    nrm = np.linalg.norm(dirs, axis=1)
    assert (nrm <= 1.0 + 1e-5).all() and (nrm > 0.85).all()
    tot = np.loadtxt(base + "_total_viewfactor_map.txt", skiprows=61)
    assert tot.shape == (100, 4) and np.allclose(tot[:, 3], np.bincount(
        np.repeat(np.arange(100), np.diff(rp)), weights=vf, minlength=100), rtol=1e-4)


def test_thermal_cli_uses_producer_files(ctx, tmp_path):
    """File-level chain: btpm_selfheating writes the lists of a dented
    binary, btpm_thermal reads them with the crater model, computes the
    shape->crater view factors from the vector list on the GPU and solves;
    its temperature files equal the same chain through the library API."""
    import math

    from binarytpm_amd import cli, formats, vfcli
    equ, fv, npv = inputs.binary_shape(2, 0, sep=1.4, dent=0.35)
    nf, ntau, period = len(fv), 650, 4.3 * 3600.0
    base = str(tmp_path / "body")
    _write_shape(base + ".txt", equ, fv)
    assert vfcli.main(["--shape", base, "--nv", str(len(equ)), "--nf", str(nf)]) == 0
    ill = th.geometry_illumination(ctx, equ, fv, ntau, (1.1, 0.0, 0.0),
                                   (math.pi / 2, 0.0), nv_primary=npv, sun_facet=False).illum
    formats.write_temperature_dat(base + "_illumination.dat", equ, fv, ill)
    d = np.load(os.path.join(GOLD, "new_crater2.npz"))
    crater = str(tmp_path / "new_crater2")
    _write_shape(crater + ".txt", d["V"], d["F"])
    formats.write_selfheating_list(crater + "_selfheating_list.txt",
                                   (d["rowptr"], d["colidx"], d["vf"]))
    cil = np.repeat(ill[:, None, :], 100, axis=1) * np.linspace(0.5, 1.0, 100)[None, :, None]
    formats.write_roughness_dat(base + "_roughness_illumination.dat", d["V"], d["F"],
                                np.ascontiguousarray(np.transpose(cil, (0, 2, 1))))
    rc = cli.main(["--shape", base, "--nv", str(len(equ)), "--nf", str(nf), "--ntau", str(ntau),
                   "--period", str(period), "--ti", "100", "--tacc", "0.01", "--selfheating",
                   "--crater", crater, "--out", str(tmp_path)])
    assert rc == 0
    _, _, T_cli = formats.read_temperature_dat(base + "_temperature(100).dat", len(equ), nf, ntau)
    Tc_cli = formats.read_roughness_dat(base + "_roughness_temperature(100).dat", 61, 100, nf,
                                        ntau)[2]
    # the same chain through the API
    csr = formats.read_selfheating_list(base + "_selfheating_list.txt", nf)
    dirs = formats.read_selfheating_vectors(base + "_selfheating_vector_list.txt", csr[0])
    assert csr[0][-1] > 100
    Vc, Fc = formats.read_shape(crater + ".txt", 61, 100)
    Vb, Fb = formats.read_shape(base + ".txt", len(equ), nf)   # the vertices as the CLI read them
    drf = th.crater_viewfactors(ctx, Vb, Fb, csr[0], csr[2], dirs, Vc, Fc)
    assert (drf > 0).sum() > 1000
    _, _, cill = formats.read_roughness_dat(base + "_roughness_illumination.dat", 61, 100, nf, ntau)
    cm = th.CraterModel(vf=formats.crater_vf_matrix(
        formats.read_selfheating_list(crater + "_selfheating_list.txt", 100)),
        illum=np.ascontiguousarray(cill), drf=drf)
    p = th.params(ntau, period, tacc=0.01)
    _, _, ill_back = formats.read_temperature_dat(base + "_illumination.dat", len(equ), nf, ntau)
    shape = th.Shape(ctx, p, ill_back, 0.1, csr=csr, scatter=0, crater=cm)
    s = th.Solve(ctx, shape, 100.0, with_craters=True)
    th.batch_run(ctx, [s])
    try:
        assert np.array_equal(s.surface_temps(), T_cli)
        assert np.array_equal(s.crater_temps(), Tc_cli)
    finally:
        s.close()
        shape.close()
