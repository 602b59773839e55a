"""Readers and writers of the Thermal stage's on-disk contract.

Formats (whitespace-separated text unless noted):
  - shape `<name>.txt`: NV lines "x y z", then NF lines "v0 v1 v2"
    (Asteroid::readmodel, T3905/main.cpp:400-410);
  - view-factor list `<name>_selfheating_list.txt`: per facet
    "n idx_1 .. idx_n vf_1 .. vf_n" (readselfheating, T3905/main.cpp:418-434);
  - `<name>_selfheating_vector_list.txt`: per facet 3 floats per dependent;
  - `<name>_photovectors.txt`: per facet "x y z" (the photon momentum
    vector); `<name>_selfheating_map.txt` / `_total_viewfactor_map.txt`: vertex
    lines, then "v0 v1 v2 value" per facet (Shape_Selfheating's writers,
    3905/Shape_Selfheating/main.cpp:837-888: every value followed by a tab);
  - temperature text `<shape>_temperature(<TI>).txt`: vertex lines, facet
    lines, then NF*NTAU lines of surface T, facet-major
    (writeglobaltemperature, T3905/main.cpp:2290-2308);
  - temperature binary `<shape>_temperature(<TI>).dat`: double[NV][3],
    int[NF][3], float[NF][NTAU] (binary_writeglobaltemperature,
    175706/Run-04/thermal.cpp:3291-3316);
  - crater temperature binary `<shape>_roughness_temperature(<TI>).dat`:
    float[NVc][3], int[NFc][3], float[NF][NFc][NTAU]
    (binary_writesurfacetemperature, T3905/main.cpp:2314-2341).
"""
from __future__ import annotations

import numpy as np


def _tokens(path):
    with open(path) as fh:
        return fh.read().split()


def read_shape(path, nv: int, nf: int):
    """Vertices [nv,3] (float64) and facets [nf,3] (int) of a shape file."""
    t = _tokens(path)
    need = 3 * nv + 3 * nf
    if len(t) < need:
        raise ValueError(f"{path}: expected {need} values, found {len(t)}")
    V = np.array(t[:3 * nv], dtype=np.float64).reshape(nv, 3)
    F = np.array(t[3 * nv:need], dtype=np.int64).reshape(nf, 3)
    return V, F


def read_shape_auto(path):
    """Shape file with unknown counts: vertex lines have 3 floats with a
    decimal point or exponent, facet lines 3 integers."""
    V, F = [], []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if len(p) != 3:
                continue
            if F or all(_is_int(x) for x in p) and V and _looks_facet(p):
                F.append([int(x) for x in p])
            else:
                V.append([float(x) for x in p])
    return np.array(V, dtype=np.float64), np.array(F, dtype=np.int64)


def _is_int(s):
    try:
        int(s)
        return True
    except ValueError:
        return False


def _looks_facet(p):
    return all("." not in x and "e" not in x.lower() for x in p)


def read_selfheating_list(path, nf: int):
    """CSR (rowptr, colidx, vf) of a `_selfheating_list.txt` file.  Rows
    are receiving facets; missing lines read as zero dependents, as the
    reference's stream reads do."""
    t = _tokens(path)
    pos = 0
    rowptr = np.zeros(nf + 1, dtype=np.int32)
    cols, vals = [], []
    for f in range(nf):
        if pos >= len(t):
            n = 0
        else:
            n = int(t[pos])
            pos += 1
        c = np.array(t[pos:pos + n], dtype=np.int64)
        v = np.array(t[pos + n:pos + 2 * n], dtype=np.float64)
        pos += 2 * n
        cols.append(c)
        vals.append(v)
        rowptr[f + 1] = rowptr[f] + n
    colidx = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
    vf = np.concatenate(vals).astype(np.float32) if vals else np.zeros(0, np.float32)
    return rowptr, colidx, vf


def _g(x) -> str:
    """A float as an ostream prints it by default (6 significant digits)."""
    return f"{float(x):g}"


def write_selfheating_list(path, csr):
    """`_selfheating_list.txt` as Shape_Selfheating writes it
    (3905/Shape_Selfheating/main.cpp:857-869): per facet the count, the
    partner indices, then the view factors, each followed by a tab."""
    rowptr, colidx, vf = csr
    with open(path, "w") as fh:
        for f in range(len(rowptr) - 1):
            a, b = rowptr[f], rowptr[f + 1]
            fh.write("".join([f"{b - a}\t"] + [f"{int(x)}\t" for x in colidx[a:b]]
                             + [f"{_g(x)}\t" for x in vf[a:b]]) + "\n")


def write_selfheating_vectors(path, rowptr, dirs):
    """`_selfheating_vector_list.txt` (main.cpp:870-880): per facet the unit
    direction of each partner, 3 values each followed by a tab."""
    with open(path, "w") as fh:
        for f in range(len(rowptr) - 1):
            a, b = rowptr[f], rowptr[f + 1]
            fh.write("".join(f"{_g(x)}\t" for x in np.asarray(dirs[a:b]).reshape(-1)) + "\n")


def read_selfheating_vectors(path, rowptr):
    """Directions [nnz][3] of a `_selfheating_vector_list.txt` for the rows
    of a view-factor list (readselfheating_vectors, T3905/main.cpp:436-446)."""
    t = np.array(_tokens(path), dtype=np.float64)
    n = int(rowptr[-1])
    if len(t) < 3 * n:
        raise ValueError(f"{path}: expected {3 * n} values, found {len(t)}")
    return t[:3 * n].reshape(n, 3).astype(np.float32)


def write_photovectors(path, photon):
    """`_photovectors.txt` (main.cpp:882-886): "x\ty\tz" per facet."""
    with open(path, "w") as fh:
        for p in np.asarray(photon):
            fh.write(f"{_g(p[0])}\t{_g(p[1])}\t{_g(p[2])}\n")


def write_facet_map(path, V, F, values):
    """`_selfheating_map.txt` / `_total_viewfactor_map.txt` (main.cpp:
    837-856): vertex lines "x\ty\tz", then "v0\tv1\tv2\tvalue" per facet."""
    with open(path, "w") as fh:
        for v in np.asarray(V):
            fh.write(f"{_g(v[0])}\t{_g(v[1])}\t{_g(v[2])}\n")
        for f, x in zip(np.asarray(F), np.asarray(values)):
            val = str(int(x)) if np.issubdtype(np.asarray(values).dtype, np.integer) else _g(x)
            fh.write(f"{int(f[0])}\t{int(f[1])}\t{int(f[2])}\t{val}\n")


def crater_vf_matrix(csr, nsf: int = 100):
    """Dense crater view factors in the kernel storage vf[i*nsf + j] =
    VF(j -> i) from a crater `_selfheating_list.txt` whose row j lists
    (i, VF(j -> i)) (the transposition of T3905/main.cpp:1994-2000),
    zero diagonal."""
    rowptr, colidx, vals = csr
    M = np.zeros((nsf, nsf), dtype=np.float32)
    for j in range(nsf):
        for k in range(rowptr[j], rowptr[j + 1]):
            M[colidx[k], j] = vals[k]
    np.fill_diagonal(M, 0.0)
    return M.reshape(-1)


def shape_properties(V, F):
    """Total area, enclosed volume and volume-equivalent diameter of a
    closed triangle mesh (Shape_Properties, 3905/Shape_Properties)."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    vol = abs(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0
    deq = (6.0 * vol / np.pi) ** (1.0 / 3.0)
    return area, vol, deq


def write_temperature_txt(path, V, F, stemp):
    """writeglobaltemperature layout: vertices, facets, then NF*NTAU values
    (facet-major) with 6 significant digits (default ostream precision)."""
    with open(path, "w") as fh:
        for v in V:
            fh.write(f"{v[0]:g}\t{v[1]:g}\t{v[2]:g}\n")
        for f in F:
            fh.write(f"{int(f[0])}\t{int(f[1])}\t{int(f[2])}\n")
        np.savetxt(fh, np.asarray(stemp, dtype=np.float32).reshape(-1, 1), fmt="%.6g")


def write_temperature_dat(path, V, F, stemp):
    with open(path, "wb") as fh:
        fh.write(np.ascontiguousarray(V, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(F, dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(stemp, dtype=np.float32).tobytes())


def read_temperature_dat(path, nv, nf, ntau):
    raw = open(path, "rb").read()
    o = 0
    V = np.frombuffer(raw, np.float64, nv * 3, o).reshape(nv, 3)
    o += nv * 24
    F = np.frombuffer(raw, np.int32, nf * 3, o).reshape(nf, 3)
    o += nf * 12
    T = np.frombuffer(raw, np.float32, nf * ntau, o).reshape(nf, ntau)
    return V, F, T


def write_roughness_dat(path, Vc, Fc, crater_temps):
    """crater_temps [nf][ntau][nsf] -> file order float[nf][nsf][ntau]."""
    with open(path, "wb") as fh:
        fh.write(np.ascontiguousarray(Vc, dtype=np.float32).tobytes())
        fh.write(np.ascontiguousarray(Fc, dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(np.transpose(crater_temps, (0, 2, 1)),
                                      dtype=np.float32).tobytes())


def read_roughness_dat(path, nvc, nfc, nf, ntau):
    raw = open(path, "rb").read()
    o = 0
    Vc = np.frombuffer(raw, np.float32, nvc * 3, o).reshape(nvc, 3)
    o += nvc * 12
    Fc = np.frombuffer(raw, np.int32, nfc * 3, o).reshape(nfc, 3)
    o += nfc * 12
    T = np.frombuffer(raw, np.float32, nf * nfc * ntau, o).reshape(nf, nfc, ntau)
    return Vc, Fc, np.transpose(T, (0, 2, 1))
