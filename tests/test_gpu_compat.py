"""The reference's extern "C" crater API (include/btpm_compat.h) driven in
single_model's call order (175706/Run-04/thermal.cpp:3465-3818) through
ctypes, against the FP64 oracle and the exact cap visibility."""
import ctypes as C
import os

import numpy as np
import pytest

from binarytpm_amd import _lib, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ES = 0.95 * 5.67e-8
EPS, SIGMA, IR_ABS = 0.95, 5.67e-8, 1.0


class SfcVertex(C.Structure):
    _fields_ = [("pos", C.c_float * 3)]


class SfcFacet(C.Structure):
    _fields_ = [("sfcvertices", C.c_int * 3), ("midpos", C.c_float * 3),
                ("trivector", (C.c_float * 3) * 2), ("normcalc", (C.c_float * 3) * 3),
                ("normal", C.c_float * 3)]


class Surface(C.Structure):
    _fields_ = [("NV", C.c_int), ("NF", C.c_int), ("SV", C.POINTER(SfcVertex)),
                ("SF", C.POINTER(SfcFacet))]


def _lib_compat():
    L = _lib.lib()
    fp, ip, bp = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_bool)
    sig = {
        "btpm_compat_configure": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "btpm_compat_status": (C.c_int, []),
        "set_cuda_device": (None, []),
        "initialise_GPU_constants": (None, [C.c_float, fp, Surface]),
        "set_GPU_facet_properties": (None, [bp, fp, fp, fp]),
        "obtain_GPU_facet_properties": (None, [bp, fp, fp, fp]),
        "load_GPU_globalillumination": (None, [fp]),
        "load_GPU_globalheating": (None, [fp]),
        "GPU_sfcpartialshadow": (None, [ip, fp, fp, fp]),
        "obtain_GPU_roughness_shadows": (None, [ip, fp, fp, fp]),
        "GPU_single_scattering": (None, [C.c_float]),
        "GPU_multiple_scattering": (None, [C.c_float, C.c_float]),
        "GPU_initialise_temperatures": (None, [C.c_float, C.c_float, C.c_float]),
        "GPU_inertia_difference": (None, [C.c_float, C.c_float, C.c_float, C.c_float]),
        "GPU_inst_equilibrium": (None, [C.c_float, C.c_float, C.c_float]),
        "GPU_converged": (None, [C.c_float]),
        "obtain_GPU_roughness_temperatures": (None, [fp, fp, fp]),
        "obtain_GPU_roughness_illumination": (None, [fp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    for name in ("initialise_GPU_facet_properties_arrays", "initialise_GPU_shadow_arrays",
                 "initialise_GPU_illumination_arrays", "initialise_GPU_globalillumination_arrays",
                 "initialise_GPU_globalheating_arrays", "initialise_GPU_temperature_arrays",
                 "freeup_GPU_shadow_arrays", "freeup_GPU_globalillumination_arrays",
                 "freeup_GPU_globalheating_arrays", "freeup_GPU_temperature_arrays",
                 "GPU_zero_temperatures"):
        f = getattr(L, name)
        f.restype, f.argtypes = None, []
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _surface(cg, m=10):
    """GPU_Surface of a cap built the way surfacegeometry fills it
    (trivector[0] = (v0 - v2)/m, trivector[1] = (v1 - v0)/m)."""
    sv = (SfcVertex * len(cg.V))()
    for i, v in enumerate(cg.V):
        sv[i].pos[:] = [float(x) for x in v]
    sf = (SfcFacet * len(cg.F))()
    for j, (a, b, c) in enumerate(cg.F):
        sf[j].sfcvertices[:] = [int(a), int(b), int(c)]
        t0 = (cg.V[a] - cg.V[c]) / m
        t1 = (cg.V[b] - cg.V[a]) / m
        sf[j].trivector[0][:] = [float(x) for x in t0]
        sf[j].trivector[1][:] = [float(x) for x in t1]
    s = Surface(len(cg.V), len(cg.F), sv, sf)
    s._keep = (sv, sf)
    return s


def test_partial_shadow_matches_exact_hemisphere():
    L = _lib_compat()
    nf, nt = 2, 24
    assert L.btpm_compat_configure(nf, nt, 56) == 0
    cg = inputs.crater_cap(90.0)
    V10 = cg.V * 10.0                                      # radius-10 hemisphere
    cg10 = inputs.CraterGeom(V10, cg.F, cg.normals, cg.area * 100, cg.samples * 10,
                             cg.z_rim * 10, cg.gamma, cg.vf)
    L.set_cuda_device()
    vf = np.ascontiguousarray(cg.vf.reshape(100, 100), np.float32)
    L.initialise_GPU_constants(15480.0, _fp(vf), _surface(cg10))
    L.initialise_GPU_facet_properties_arrays()
    L.initialise_GPU_shadow_arrays()
    rng = np.random.default_rng(3)
    el = rng.uniform(0.05, 1.5, nf * nt)
    az = rng.uniform(0, 2 * np.pi, nf * nt)
    d = np.stack([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)], 1)
    shadow = np.ones((nf, nt), np.int32)
    sfc = np.ones((nf, nt, 100), np.float32)
    vec = np.ascontiguousarray(d.reshape(nf, nt, 3), np.float32)
    ang = np.zeros((nf, nt, 100), np.float32)
    L.GPU_sfcpartialshadow(shadow.ctypes.data_as(C.POINTER(C.c_int)), _fp(sfc), _fp(vec), _fp(ang))
    L.obtain_GPU_roughness_shadows(shadow.ctypes.data_as(C.POINTER(C.c_int)), _fp(sfc), _fp(vec),
                                   _fp(ang))
    assert L.btpm_compat_status() == 0
    exact = inputs.crater_visibility(cg, d).reshape(nf, nt, 100)
    assert np.abs(sfc - exact).max() <= 0.021


def test_compat_crater_loop_matches_oracle():
    """initialise -> scattering -> RGH -> initial temperatures -> rotations
    with GPU_converged after rev > 5 -> obtain temperatures."""
    L = _lib_compat()
    case, crater = inputs.rough_primary(subdiv=1)          # 80 facets
    nf, nt = case.illum.shape
    ti = 400.0
    p = O.params(nt, case.rot_period, tacc=0.01)
    F = O.facet_scatter(case.illum, 0.1, case.csr, multiple=True, acc=1e-4)
    rgi = O.rgi(case.csr, crater["drf"], F, np.full(nf, 0.1), nf).astype(np.float32)
    sm = O.smooth_solve(p, F, 0.1, ti, csr=case.csr)
    rgh = O.rgh(case.csr, crater["drf"], sm.stemp, ES, nf).astype(np.float32)
    il = O.crater_scatter(crater["vf"], 0.1, crater["shadow"], crater["cosang"], rgi,
                          crater["solar"], multiple=True, acc=1e-4)
    ref = O.crater_solve(p, crater["vf"], 0.1, ti, il, rgh)

    assert L.btpm_compat_configure(nf, nt, 56) == 0
    L.set_cuda_device()
    vf = np.ascontiguousarray(crater["vf"].reshape(100, 100), np.float32)
    L.initialise_GPU_constants(np.float32(case.rot_period), _fp(vf), _surface(crater["geom"]))
    L.initialise_GPU_facet_properties_arrays()
    conv = np.zeros(nf, np.bool_)
    alb = np.full(nf, 0.1, np.float32)
    tis = np.full(nf, ti, np.float32)
    nrv = np.ones(nf, np.float32)
    bp = conv.ctypes.data_as(C.POINTER(C.c_bool))
    L.set_GPU_facet_properties(bp, _fp(alb), _fp(tis), _fp(nrv))
    L.initialise_GPU_shadow_arrays()
    L.initialise_GPU_globalillumination_arrays()
    shadow = np.ones((nf, nt), np.int32)
    sh = np.ascontiguousarray(crater["shadow"], np.float32)
    cs = np.ascontiguousarray(crater["cosang"], np.float32)
    vec = np.zeros((nf, nt, 3), np.float32)
    # the inputs already carry the partial shadows: skip GPU_sfcpartialshadow
    L.load_GPU_globalillumination(_fp(rgi))
    L.initialise_GPU_illumination_arrays()
    # scattering reads the shadow / angle arrays uploaded by the shadow call
    L.GPU_sfcpartialshadow(np.zeros((nf, nt), np.int32).ctypes.data_as(C.POINTER(C.c_int)),
                           _fp(sh), _fp(vec), _fp(cs))
    L.GPU_single_scattering(np.float32(crater["solar"]))
    L.GPU_multiple_scattering(np.float32(crater["solar"]), np.float32(1e-4))
    L.freeup_GPU_shadow_arrays()
    L.freeup_GPU_globalillumination_arrays()
    L.initialise_GPU_temperature_arrays()
    L.initialise_GPU_globalheating_arrays()
    L.load_GPU_globalheating(_fp(rgh))
    L.GPU_initialise_temperatures(EPS, SIGMA, IR_ABS)
    rev = 0
    done = False
    while not done and rev < 301:
        L.GPU_inertia_difference(EPS, SIGMA, IR_ABS, np.float32(0.01))
        if rev > 5:
            L.GPU_converged(np.float32(0.01))
            L.obtain_GPU_facet_properties(bp, _fp(alb), _fp(tis), _fp(nrv))
            done = bool(conv.all())
        rev += 1
    s0 = np.zeros((nf, nt, 100), np.float32)
    s1 = np.zeros((nf, nt, 100), np.float32)
    tp = np.zeros((nf, 57, 100), np.float32)
    L.obtain_GPU_roughness_temperatures(_fp(s0), _fp(s1), _fp(tp))
    assert L.btpm_compat_status() == 0, _lib.lib().btpm_last_error()
    d = np.abs(s1.astype(np.float64) - ref.stemp)
    print(f"[parity] compat crater loop: rev {rev} (oracle {ref.rotations}), "
          f"max|dT| {d.max():.5f} K rms {np.sqrt((d ** 2).mean()):.5f} K", flush=True)
    assert done and abs(rev - ref.rotations) <= 1
    assert d.max() <= 0.05
    assert np.abs(tp.astype(np.float64) - ref.temp).max() <= 0.05
    assert np.abs(s1 - s0).max() < 0.01          # frozen: last two rotations agree
    L.freeup_GPU_temperature_arrays()
    L.freeup_GPU_globalheating_arrays()


def test_compat_crater_inst_equilibrium_matches_oracle():
    """TI = 0 through the reference's calls, in its order for
    instant_equib_model = 1 (175706/Run-04/thermal.cpp:3567-3574, 3736-3745):
    load_GPU_globalheating, GPU_zero_temperatures, then GPU_inst_equilibrium
    rotations with GPU_converged after rev > 5; against the oracle's
    instantaneous crater equilibrium from zero temperatures."""
    L = _lib_compat()
    case, crater = inputs.rough_primary(subdiv=1)          # 80 facets
    nf, nt = case.illum.shape
    p = O.params(nt, case.rot_period, tacc=0.01)
    F = O.facet_scatter(case.illum, 0.1, case.csr, multiple=True, acc=1e-4)
    rgi = O.rgi(case.csr, crater["drf"], F, np.full(nf, 0.1), nf).astype(np.float32)
    sm = O.smooth_solve(p, F, 0.1, 0.0, csr=case.csr)
    rgh = O.rgh(case.csr, crater["drf"], sm.stemp, ES, nf).astype(np.float32)
    il = O.crater_scatter(crater["vf"], 0.1, crater["shadow"], crater["cosang"], rgi,
                          crater["solar"], multiple=True, acc=1e-4)
    ref = O.crater_solve(p, crater["vf"], 0.1, 0.0, il, rgh, temp=np.zeros((nf, 57, 100)),
                         stemp=np.zeros(il.shape))

    assert L.btpm_compat_configure(nf, nt, 56) == 0
    L.set_cuda_device()
    vf = np.ascontiguousarray(crater["vf"].reshape(100, 100), np.float32)
    L.initialise_GPU_constants(np.float32(case.rot_period), _fp(vf), _surface(crater["geom"]))
    L.initialise_GPU_facet_properties_arrays()
    conv = np.zeros(nf, np.bool_)
    alb = np.full(nf, 0.1, np.float32)
    tis = np.zeros(nf, np.float32)
    nrv = np.ones(nf, np.float32)
    bp = conv.ctypes.data_as(C.POINTER(C.c_bool))
    L.set_GPU_facet_properties(bp, _fp(alb), _fp(tis), _fp(nrv))
    L.initialise_GPU_shadow_arrays()
    L.initialise_GPU_globalillumination_arrays()
    sh = np.ascontiguousarray(crater["shadow"], np.float32)
    cs = np.ascontiguousarray(crater["cosang"], np.float32)
    vec = np.zeros((nf, nt, 3), np.float32)
    L.load_GPU_globalillumination(_fp(rgi))
    L.initialise_GPU_illumination_arrays()
    L.GPU_sfcpartialshadow(np.zeros((nf, nt), np.int32).ctypes.data_as(C.POINTER(C.c_int)),
                           _fp(sh), _fp(vec), _fp(cs))
    L.GPU_single_scattering(np.float32(crater["solar"]))
    L.GPU_multiple_scattering(np.float32(crater["solar"]), np.float32(1e-4))
    L.freeup_GPU_shadow_arrays()
    L.freeup_GPU_globalillumination_arrays()
    L.initialise_GPU_temperature_arrays()
    L.initialise_GPU_globalheating_arrays()
    L.load_GPU_globalheating(_fp(rgh))
    L.GPU_zero_temperatures()
    s0 = np.ones((nf, nt, 100), np.float32)
    s1 = np.ones((nf, nt, 100), np.float32)
    tp = np.ones((nf, 57, 100), np.float32)
    L.obtain_GPU_roughness_temperatures(_fp(s0), _fp(s1), _fp(tp))
    assert not s0.any() and not s1.any() and not tp.any()
    rev = 0
    done = False
    while not done and rev < 301:
        L.GPU_inst_equilibrium(EPS, SIGMA, IR_ABS)
        if rev > 5:
            L.GPU_converged(np.float32(0.01))
            L.obtain_GPU_facet_properties(bp, _fp(alb), _fp(tis), _fp(nrv))
            done = bool(conv.all())
        rev += 1
    L.obtain_GPU_roughness_temperatures(_fp(s0), _fp(s1), _fp(tp))
    assert L.btpm_compat_status() == 0, _lib.lib().btpm_last_error()
    d = np.abs(s1.astype(np.float64) - ref.stemp)
    print(f"[parity] compat crater TI 0: rev {rev} (oracle {ref.rotations}), "
          f"max|dT| {d.max():.5f} K rms {np.sqrt((d ** 2).mean()):.5f} K", flush=True)
    assert done and abs(rev - ref.rotations) <= 1
    assert d.max() <= 0.05
    # radiative equilibrium with the previous rotation's crater self-heating,
    # es T^4 = (1 - A) illum + RGH + sum_i VF(j->i) es T_prev,i^4, up to the
    # GPU illumination's difference from the oracle's (< 0.05 W/m^2)
    e_prev = ES * s0.astype(np.float64) ** 4
    q = rgh + np.einsum("fsi,ij->fsj", e_prev, vf.astype(np.float64))
    lhs = ES * s1.astype(np.float64) ** 4
    rhs = 0.9 * il + q
    assert (np.abs(lhs - rhs) - 1e-5 * rhs).max() < 0.06
    L.freeup_GPU_temperature_arrays()
    L.freeup_GPU_globalheating_arrays()


def test_cli_file_contract(tmp_path):
    """btpm_thermal reads the reference's input files and writes its
    temperature files; their contents match the FP64 oracle."""
    from binarytpm_amd import cli, formats
    case = inputs.config0(subdiv=1)
    V, F = inputs.ellipsoid(1, (0.6, 0.5, 0.45))
    base = str(tmp_path / "body")
    with open(base + ".txt", "w") as fh:
        for v in V:
            fh.write(f"{v[0]:.9g}\t{v[1]:.9g}\t{v[2]:.9g}\n")
        for f in F:
            fh.write(f"{f[0]}\t{f[1]}\t{f[2]}\n")
    formats.write_temperature_dat(base + "_illumination.dat", V, F, case.illum)
    rc = cli.main(["--shape", base, "--nv", str(len(V)), "--nf", str(len(F)), "--ntau", "650",
                   "--period", str(case.rot_period), "--ti", "50", "250", "--tacc", "0.01",
                   "--out", str(tmp_path)])
    assert rc == 0
    for ti in (50.0, 250.0):
        _, _, T = formats.read_temperature_dat(str(tmp_path / f"body_temperature({ti:g}).dat"),
                                               len(V), len(F), 650)
        ref = O.smooth_solve(O.params(650, case.rot_period, tacc=0.01), case.illum, 0.1, ti)
        assert np.abs(T.astype(np.float64) - ref.stemp).max() <= 0.05
        assert os.path.exists(str(tmp_path / f"body_temperature({ti:g}).txt"))
