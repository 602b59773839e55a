"""File contract of the stage (CPU): readers/writers round-trip, the
self-heating list reader matches its writer, the crater list transposes to
the kernel storage VF(j->i), and text output keeps 6 significant digits."""
import os

import numpy as np

from binarytpm_amd import formats, inputs


def test_shape_and_temperature_roundtrip(tmp_path):
    V, F = inputs.ellipsoid(1, (0.6, 0.5, 0.45))
    rng = np.random.default_rng(0)
    T = rng.uniform(100, 400, (len(F), 650)).astype(np.float32)
    p = str(tmp_path / "body_temperature(200).dat")
    formats.write_temperature_dat(p, V, F, T)
    assert os.path.getsize(p) == len(V) * 24 + len(F) * 12 + T.size * 4
    V2, F2, T2 = formats.read_temperature_dat(p, len(V), len(F), 650)
    np.testing.assert_array_equal(T2, T)
    np.testing.assert_array_equal(F2, F)
    txt = str(tmp_path / "body_temperature(200).txt")
    formats.write_temperature_txt(txt, V, F, T)
    lines = open(txt).read().split("\n")
    assert len(lines) - 1 == len(V) + len(F) + T.size
    vals = np.array(lines[len(V) + len(F):-1], dtype=np.float64)
    np.testing.assert_allclose(vals, T.reshape(-1), rtol=5e-6)
    shp = str(tmp_path / "body.txt")
    with open(shp, "w") as fh:
        for v in V:
            fh.write(f"{v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for f in F:
            fh.write(f"{f[0]} {f[1]} {f[2]}\n")
    V3, F3 = formats.read_shape(shp, len(V), len(F))
    np.testing.assert_allclose(V3, V, rtol=1e-8)
    np.testing.assert_array_equal(F3, F)


def test_roughness_file_layout(tmp_path):
    cg = inputs.crater_cap(45.0)
    rng = np.random.default_rng(1)
    Tc = rng.uniform(100, 400, (3, 650, 100)).astype(np.float32)
    p = str(tmp_path / "r.dat")
    formats.write_roughness_dat(p, cg.V, cg.F, Tc)
    raw = np.fromfile(p, dtype=np.float32, offset=61 * 12 + 100 * 12).reshape(3, 100, 650)
    np.testing.assert_array_equal(raw, np.transpose(Tc, (0, 2, 1)))   # [f][j][s] on disk
    _, _, back = formats.read_roughness_dat(p, 61, 100, 3, 650)
    np.testing.assert_array_equal(back, Tc)


def test_selfheating_list_roundtrip(tmp_path):
    csr, _ = inputs.sparse_viewfactors(50, 0.1, 0.2, seed=2)
    p = str(tmp_path / "x_selfheating_list.txt")
    formats.write_selfheating_list(p, csr)
    rp, ci, vf = formats.read_selfheating_list(p, 50)
    np.testing.assert_array_equal(rp, csr[0])
    np.testing.assert_array_equal(ci, csr[1])
    np.testing.assert_allclose(vf, csr[2], rtol=1e-5)
    # missing trailing facets read as zero dependents, like a failed stream read
    rp2, _, _ = formats.read_selfheating_list(p, 60)
    assert rp2[-1] == rp[-1] and np.all(np.diff(rp2[50:]) == 0)


def test_crater_list_transposition():
    rows = (np.array([0, 2, 3, 3], np.int32), np.array([1, 2, 0], np.int32),
            np.array([0.25, 0.5, 0.125], np.float32))
    M = formats.crater_vf_matrix(rows, nsf=3).reshape(3, 3)
    # row j of the list holds (i, VF(j->i)); storage M[i][j] = VF(j->i)
    assert M[1, 0] == 0.25 and M[2, 0] == 0.5 and M[0, 1] == 0.125
    assert M.trace() == 0


def test_producer_files_layout(tmp_path):
    """Shape_Selfheating's writers put a tab after every value
    (3905/Shape_Selfheating/main.cpp:837-888); the readers take them back."""
    rp = np.array([0, 2, 2, 3], np.int32)
    ci = np.array([1, 2, 0], np.int32)
    vf = np.array([0.125, 1e-7, 0.5], np.float32)
    dirs = np.array([[1, 0, 0], [0, 0.6, 0.8], [0, 0, -1]], np.float32)
    p = str(tmp_path / "s")
    formats.write_selfheating_list(p + "_selfheating_list.txt", (rp, ci, vf))
    assert open(p + "_selfheating_list.txt").read() == "2\t1\t2\t0.125\t1e-07\t\n0\t\n1\t0\t0.5\t\n"
    formats.write_selfheating_vectors(p + "_v.txt", rp, dirs)
    assert open(p + "_v.txt").read().splitlines()[1] == ""
    np.testing.assert_array_equal(formats.read_selfheating_vectors(p + "_v.txt", rp), dirs)
    formats.write_photovectors(p + "_photovectors.txt", dirs)
    assert open(p + "_photovectors.txt").read().splitlines()[1] == "0\t0.6\t0.8"
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1.5, 0]], np.float32)
    F = np.array([[0, 1, 2], [0, 2, 1], [1, 2, 0]])
    formats.write_facet_map(p + "_map.txt", V, F, np.diff(rp))
    assert open(p + "_map.txt").read().splitlines()[2:4] == ["0\t1.5\t0", "0\t1\t2\t2"]
    formats.write_facet_map(p + "_tot.txt", V, F, np.array([0.5, 0, 0.25], np.float32))
    assert open(p + "_tot.txt").read().splitlines()[-1] == "1\t2\t0\t0.25"


def test_text6_matches_printf():
    """The vectorised 6-digit read-back of the light-curve regression equals
    formatting each float with %g and parsing it, on random temperatures and
    on the values next to every decade boundary."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "fg3reg", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                               "test_gpu_regression_fg3.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.default_rng(0)
    edges = np.float32([0.01, 0.1, 1.0, 10.0, 100.0, 1000.0])
    a = np.concatenate([rng.uniform(0, 600, 200_000), 10.0 ** rng.uniform(-3, 3, 50_000),
                        np.nextafter(edges, np.float32(0)), np.nextafter(edges, np.float32(1e4)),
                        edges, [0.0, 99.9999995, 999.9995, 123.4565]]).astype(np.float32)
    slow = np.array([float(f"{x:g}") for x in a], np.float64)
    assert np.array_equal(mod._text6(a), slow)
