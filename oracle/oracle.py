"""ctypes front end of the CPU checker libraries (test infrastructure only).

Imported by tests/, by ``__graft_entry__.smoke()`` and by the cpu_baseline leg
of bench.py.  The product path (``binarytpm_amd``) never imports this module.

``liboracle.so`` is the FP64 restatement (btpm_oracle.c); ``librefcpu.so`` is
the float OpenMP restatement of the reference CPU loop shape (refcpu.c).
Both are built by ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC = None
_RC = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class OrcParams(C.Structure):
    _fields_ = [
        ("ntau", C.c_int), ("nx", C.c_int), ("rot_period", C.c_double),
        ("eps", C.c_double), ("sigma", C.c_double), ("ir_abs", C.c_double),
        ("tacc", C.c_double), ("newton_tol", C.c_double), ("max_rev", C.c_int),
        ("check_after", C.c_int), ("newton_cap", C.c_int),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("rotations", C.c_int), ("converged", C.c_int), ("newton_iters", C.c_long),
        ("newton_capped", C.c_long), ("column_steps", C.c_long),
    ]


class OrcGeometryIn(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("ntau", C.c_int), ("nv_primary", C.c_int),
        ("equpos", C.POINTER(C.c_double)), ("fv", C.POINTER(C.c_int)),
        ("rev_primary", C.c_double), ("rev_secondary", C.c_double),
        ("pol", C.c_double * 3), ("rot", C.c_double * 2), ("au_km", C.c_double),
        ("s_begin", C.c_int), ("s_count", C.c_int), ("sun_scale", C.c_double),
    ]


AU_KM = 149.6e6  # T175706/thermal.cpp:13


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `make -C {_HERE}`")
    return C.CDLL(path)


def lib():
    global _ORC
    if _ORC is None:
        _ORC = _load("liboracle.so")
    return _ORC


def rclib():
    global _RC
    if _RC is None:
        _RC = _load("librefcpu.so")
        _RC.rc_max_threads.restype = C.c_int
    return _RC


def _d(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.int32)


def _f(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def params(ntau, rot_period, tacc=0.01, eps=0.95, sigma=5.67e-8, ir_abs=1.0,
           nx=56, max_rev=301, check_after=5, newton_tol=None, newton_cap=200):
    p = OrcParams()
    p.ntau, p.nx, p.rot_period = ntau, nx, rot_period
    p.eps, p.sigma, p.ir_abs = eps, sigma, ir_abs
    p.tacc = tacc
    p.newton_tol = tacc / 10.0 if newton_tol is None else newton_tol
    p.max_rev, p.check_after, p.newton_cap = max_rev, check_after, newton_cap
    return p


@dataclass
class SolveResult:
    stemp: np.ndarray
    temp: np.ndarray
    rotations: int
    converged: bool
    frozen_rev: np.ndarray
    newton_iters: int
    newton_capped: int
    column_steps: int


def _csr(csr):
    if csr is None:
        return None, None, None
    rowptr, colidx, vals = csr
    return _i(rowptr), _i(colidx), _d(vals)


def smooth_init(p, F, albedo):
    F = _d(F)
    nf = F.shape[0]
    temp = np.zeros((nf, p.nx + 1))
    stemp = np.zeros((nf, p.ntau))
    lib().orc_smooth_init(C.byref(p), nf, _ptr(F, _dp), _ptr(_d(albedo), _dp),
                          _ptr(temp, _dp), _ptr(stemp, _dp))
    return temp, stemp


def smooth_solve(p, F, albedo, ti, numrevs=None, csr=None, temp=None, stemp=None):
    """Full smooth solve; cold start from smooth_init unless temp/stemp given."""
    F = _d(F)
    nf = F.shape[0]
    albedo = _d(np.broadcast_to(albedo, (nf,)))
    ti = _d(np.broadcast_to(ti, (nf,)))
    numrevs = _d(np.ones(nf) if numrevs is None else np.broadcast_to(numrevs, (nf,)))
    if temp is None:
        temp, stemp = smooth_init(p, F, albedo)
    temp, stemp = _d(temp).copy(), _d(stemp).copy()
    rp, ci, vv = _csr(csr)
    frozen_rev = np.zeros(nf, dtype=np.int32)
    st = OrcStats()
    lib().orc_smooth_solve(C.byref(p), nf, _ptr(F, _dp), _ptr(albedo, _dp), _ptr(ti, _dp),
                           _ptr(numrevs, _dp), _ptr(rp, _ip), _ptr(ci, _ip), _ptr(vv, _dp),
                           _ptr(temp, _dp), _ptr(stemp, _dp), _ptr(frozen_rev, _ip),
                           C.byref(st))
    return SolveResult(stemp, temp, st.rotations, bool(st.converged), frozen_rev,
                       st.newton_iters, st.newton_capped, st.column_steps)


def smooth_rotation(p, F, albedo, ti, numrevs, temp, stemp, csr=None, frozen=None):
    F = _d(F)
    nf = F.shape[0]
    temp, stemp = _d(temp).copy(), _d(stemp).copy()
    dmax = np.zeros(nf)
    rp, ci, vv = _csr(csr)
    fz = None if frozen is None else np.ascontiguousarray(frozen, dtype=np.uint8)
    lib().orc_smooth_rotation(C.byref(p), nf, _ptr(F, _dp), _ptr(_d(albedo), _dp),
                              _ptr(_d(ti), _dp), _ptr(_d(numrevs), _dp), _ptr(rp, _ip),
                              _ptr(ci, _ip), _ptr(vv, _dp), _ptr(fz, _up), _ptr(temp, _dp),
                              _ptr(stemp, _dp), _ptr(dmax, _dp), None)
    return temp, stemp, dmax


def newton_sbc(src, m1, t1, es, t0, tol, cap=200):
    it = C.c_int(0)
    f = lib().orc_newton_sbc
    f.restype = C.c_double
    f.argtypes = [C.c_double] * 6 + [C.c_int, C.POINTER(C.c_int)]
    x = f(src, m1, t1, es, t0, tol, cap, C.byref(it))
    return x, it.value


def facet_scatter(direct, albedo, csr, multiple=True, acc=1e-4):
    direct = _d(direct)
    nf, nt = direct.shape
    albedo = _d(np.broadcast_to(albedo, (nf,)))
    rp, ci, vv = _csr(csr)
    F = np.zeros_like(direct)
    lib().orc_facet_scatter_single(nf, nt, _ptr(direct, _dp), _ptr(albedo, _dp), _ptr(rp, _ip),
                                   _ptr(ci, _ip), _ptr(vv, _dp), _ptr(F, _dp))
    if multiple:
        fn = lib().orc_facet_scatter_multiple
        fn.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, _ip, _dp, C.c_double, _dp]
        fn(nf, nt, _ptr(direct, _dp), _ptr(albedo, _dp), _ptr(rp, _ip), _ptr(ci, _ip),
           _ptr(vv, _dp), acc, _ptr(F, _dp))
    return F


def rgi(csr, drf, F, albedo, nfac):
    rp, ci, _ = _csr(csr)
    drf, F = _d(drf), _d(F)
    ns = drf.shape[1]
    nt = F.shape[1]
    out = np.zeros((nfac, nt, ns))
    lib().orc_rgi(nfac, nt, ns, _ptr(rp, _ip), _ptr(ci, _ip), _ptr(drf, _dp), _ptr(F, _dp),
                  _ptr(_d(np.broadcast_to(albedo, (F.shape[0],))), _dp), _ptr(out, _dp))
    return out


def rgh(csr, drf, stemp_smooth, eps_sigma, nfac):
    rp, ci, _ = _csr(csr)
    drf, st = _d(drf), _d(stemp_smooth)
    ns = drf.shape[1]
    nt = st.shape[1]
    out = np.zeros((nfac, nt, ns))
    fn = lib().orc_rgh
    fn.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_double, _dp]
    fn(nfac, nt, ns, _ptr(rp, _ip), _ptr(ci, _ip), _ptr(drf, _dp), _ptr(st, _dp), eps_sigma,
       _ptr(out, _dp))
    return out


def crater_scatter(vf, albedo, shadow, cosang, rgi_, solar, multiple=True, acc=1e-4):
    vf, shadow, cosang, rgi_ = _d(vf), _d(shadow), _d(cosang), _d(rgi_)
    nfac, nt, ns = shadow.shape
    albedo = _d(np.broadcast_to(albedo, (nfac,)))
    out = np.zeros((nfac, nt, ns))
    fn = lib().orc_crater_scatter_single
    fn.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
    fn(nfac, nt, ns, _ptr(vf, _dp), _ptr(albedo, _dp), _ptr(shadow, _dp), _ptr(cosang, _dp),
       _ptr(rgi_, _dp), solar, _ptr(out, _dp))
    if multiple:
        fn = lib().orc_crater_scatter_multiple
        fn.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double,
                       C.c_double, _dp]
        fn(nfac, nt, ns, _ptr(vf, _dp), _ptr(albedo, _dp), _ptr(shadow, _dp),
           _ptr(cosang, _dp), _ptr(rgi_, _dp), solar, acc, _ptr(out, _dp))
    return out


def crater_init(p, vf, albedo, illum, rgh_):
    vf, illum, rgh_ = _d(vf), _d(illum), _d(rgh_)
    nfac, nt, ns = illum.shape
    albedo = _d(np.broadcast_to(albedo, (nfac,)))
    temp = np.zeros((nfac, p.nx + 1, ns))
    stemp = np.zeros((nfac, nt, ns))
    lib().orc_crater_init(C.byref(p), nfac, ns, _ptr(vf, _dp), _ptr(albedo, _dp),
                          _ptr(illum, _dp), _ptr(rgh_, _dp), _ptr(temp, _dp), _ptr(stemp, _dp))
    return temp, stemp


def crater_solve(p, vf, albedo, ti, illum, rgh_, numrevs=None, temp=None, stemp=None):
    vf, illum, rgh_ = _d(vf), _d(illum), _d(rgh_)
    nfac, nt, ns = illum.shape
    albedo = _d(np.broadcast_to(albedo, (nfac,)))
    ti = _d(np.broadcast_to(ti, (nfac,)))
    numrevs = _d(np.ones(nfac) if numrevs is None else np.broadcast_to(numrevs, (nfac,)))
    if temp is None:
        temp, stemp = crater_init(p, vf, albedo, illum, rgh_)
    temp, stemp = _d(temp).copy(), _d(stemp).copy()
    frozen_rev = np.zeros(nfac, dtype=np.int32)
    st = OrcStats()
    lib().orc_crater_solve(C.byref(p), nfac, ns, _ptr(vf, _dp), _ptr(albedo, _dp),
                           _ptr(ti, _dp), _ptr(numrevs, _dp), _ptr(illum, _dp),
                           _ptr(rgh_, _dp), _ptr(temp, _dp), _ptr(stemp, _dp),
                           _ptr(frozen_rev, _ip), C.byref(st))
    return SolveResult(stemp, temp, st.rotations, bool(st.converged), frozen_rev,
                       st.newton_iters, st.newton_capped, st.column_steps)


def ti_sweep(p, F, albedo, tis, numrevs=None, csr=None, crater=None, es=None):
    """The reference's thermal-inertia sweep with its carried state
    (T175706/thermal.cpp:3554-3576 initialise once, :3676-3760 every TI from
    the previous TI's final state; T3905/main.cpp:2478-2517 the same with the
    initialisation at the first TI > 0): per TI the smooth facets to
    convergence from the previous TI's final profiles and surface series,
    then RGH from the converged smooth temperatures and the crater columns
    from the previous TI's final crater state.  Convergence status restarts
    at every TI (reset_asteroid_status), so the rev > 5 gate applies afresh.

    The first TI starts from the initial state of a single solve: smooth
    facets from smooth_init; craters from crater_init with the RGH of the
    converged smooth temperatures (the reference builds its first crater
    state from the RGH of the initial smooth temperatures; the converged
    periodic state is the same), or from zero temperatures at TI 0
    (roughness_zero_temperatures, T175706/thermal_GPU.cu:348-367).  A TI-0
    smooth solve carries no depth profile, so the next TI > 0 starts from the
    initial state again, as the reference initialises at its first TI > 0.

    crater = dict(vf=[100*100], illum=[nf][ntau][100], drf=[nnz][100] or
    None) with the crater csr equal to `csr`; es = eps*sigma for RGH.
    Returns [(smooth SolveResult, crater SolveResult or None)] per TI."""
    F = _d(F)
    nf = F.shape[0]
    es = p.eps * p.sigma if es is None else es
    out = []
    st_s = None            # carried smooth (temp, stemp), None = initial state
    st_c = None            # carried crater (temp, stemp)
    for ti in tis:
        ti = float(ti)
        if st_s is None:
            sm = smooth_solve(p, F, albedo, ti, numrevs=numrevs, csr=csr)
        else:
            sm = smooth_solve(p, F, albedo, ti, numrevs=numrevs, csr=csr, temp=st_s[0],
                              stemp=st_s[1])
        st_s = None if ti == 0.0 else (sm.temp, sm.stemp)
        cr = None
        if crater is not None:
            ns = crater["illum"].shape[2]
            if crater.get("drf") is not None and csr is not None:
                rgh_ = rgh(csr, crater["drf"], sm.stemp, es, nf)
            else:
                rgh_ = np.zeros_like(_d(crater["illum"]))
            if st_c is not None:
                tc, sc = st_c
            elif ti == 0.0:
                tc = np.zeros((nf, p.nx + 1, ns))
                sc = np.zeros(crater["illum"].shape)
            else:
                tc = sc = None
            cr = crater_solve(p, crater["vf"], albedo, ti, crater["illum"], rgh_,
                              numrevs=numrevs, temp=tc, stemp=sc)
            st_c = None if ti == 0.0 else (cr.temp, cr.stemp)
        out.append((sm, cr))
    return out


# ---- refcpu (float OpenMP restatement, the timed CPU baseline) -------------

def refcpu_threads():
    return rclib().rc_max_threads()


def refcpu_smooth_solve(F, albedo, ti, ntau, rot_period, tacc=0.01, numrevs=None, csr=None,
                        eps=0.95, sigma=5.67e-8, ir_abs=1.0, nx=56, max_rev=301,
                        check_after=5, nthreads=0, state=None):
    """refcpu smooth solve.  state: None (initial state), or a dict with
    float32 'temp' [nf][nx+1] and 'stemp' [nf][ntau] to start from (the TI
    sweep's carried state); the dict is updated with the final state."""
    F = _f(F)
    nf = F.shape[0]
    albedo = _f(np.broadcast_to(albedo, (nf,)))
    ti = _f(np.broadcast_to(ti, (nf,)))
    numrevs = _f(np.ones(nf) if numrevs is None else np.broadcast_to(numrevs, (nf,)))
    if csr is not None:
        rp, ci, vv = _i(csr[0]), _i(csr[1]), _f(csr[2])
    else:
        rp = ci = vv = None
    warm = state is not None and "temp" in state
    out = np.ascontiguousarray(state["stemp"], np.float32).copy() if warm else \
        np.zeros((nf, ntau), dtype=np.float32)
    temp = np.ascontiguousarray(state["temp"], np.float32).copy() if warm else \
        np.zeros((nf, nx + 1), dtype=np.float32)
    steps = C.c_long(0)
    fn = rclib().rc_smooth_solve
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                   C.c_float, C.c_int, C.c_int, _fp, _fp, _fp, _fp, _ip, _ip, _fp, _fp,
                   C.c_int, C.POINTER(C.c_long), _fp, C.c_int]
    fn.restype = C.c_int
    rev = fn(nf, ntau, nx, rot_period, eps, sigma, ir_abs, tacc, max_rev, check_after,
             _ptr(F, _fp), _ptr(albedo, _fp), _ptr(ti, _fp), _ptr(numrevs, _fp),
             _ptr(rp, _ip), _ptr(ci, _ip), _ptr(vv, _fp), _ptr(out, _fp), nthreads,
             C.byref(steps), _ptr(temp, _fp), int(warm))
    if state is not None:
        state["temp"], state["stemp"] = temp, out.copy()
    return out, rev, steps.value


def refcpu_crater_solve(vf, albedo, ti, illum, rgh_, rot_period, tacc=0.01, numrevs=None,
                        eps=0.95, sigma=5.67e-8, ir_abs=1.0, nx=56, max_rev=301,
                        check_after=5, nthreads=0, state=None):
    """refcpu crater solve; state as for refcpu_smooth_solve ('temp'
    [nfac][nx+1][ns], 'stemp' [nfac][ntau][ns])."""
    illum, rgh_, vf = _f(illum), _f(rgh_), _f(vf)
    nfac, nt, ns = illum.shape
    albedo = _f(np.broadcast_to(albedo, (nfac,)))
    ti = _f(np.broadcast_to(ti, (nfac,)))
    numrevs = _f(np.ones(nfac) if numrevs is None else np.broadcast_to(numrevs, (nfac,)))
    warm = state is not None and "temp" in state
    out = np.ascontiguousarray(state["stemp"], np.float32).copy() if warm else \
        np.zeros((nfac, nt, ns), dtype=np.float32)
    temp = np.ascontiguousarray(state["temp"], np.float32).copy() if warm else \
        np.zeros((nfac, nx + 1, ns), dtype=np.float32)
    steps = C.c_long(0)
    fn = rclib().rc_crater_solve
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                   C.c_float, C.c_float, C.c_int, C.c_int, _fp, _fp, _fp, _fp, _fp, _fp, _fp,
                   C.c_int, C.POINTER(C.c_long), _fp, C.c_int]
    fn.restype = C.c_int
    rev = fn(nfac, ns, nt, nx, rot_period, eps, sigma, ir_abs, tacc, max_rev, check_after,
             _ptr(vf, _fp), _ptr(albedo, _fp), _ptr(ti, _fp), _ptr(numrevs, _fp),
             _ptr(illum, _fp), _ptr(rgh_, _fp), _ptr(out, _fp), nthreads, C.byref(steps),
             _ptr(temp, _fp), int(warm))
    if state is not None:
        state["temp"], state["stemp"] = temp, out.copy()
    return out, rev, steps.value


@dataclass
class Geometry:
    shadow: np.ndarray     # [nf][ntau] int32: 1 lit, 0 back-facing or shadowed
    cosang: np.ndarray     # [nf][ntau] float64
    illum: np.ndarray      # [nf][ntau] float32: solar cos(angle) shadow
    sun_facet: np.ndarray  # [nf][ntau][3] float32: unit Sun direction, facet frame
    solar: float


def geometry(equpos, fv, ntau, pol, rot, nv_primary=None, rev_primary=1.0, rev_secondary=1.0,
             au_km=AU_KM, steps=None, sun_scale=0.0):
    """Illumination geometry and shadows (geometry.c): the shape's vertices
    [nv][3] (km) and facets [nf][3], the first nv_primary vertices on the
    primary; pol = (r AU, ecliptic lat, lon), rot = (pole lat, lon), radians.
    steps = (first, count) computes that range of the cycle only."""
    equpos = np.ascontiguousarray(equpos, dtype=np.float64)
    fv = np.ascontiguousarray(fv, dtype=np.int32)
    nv, nf = equpos.shape[0], fv.shape[0]
    g = OrcGeometryIn()
    g.nv, g.nf, g.ntau = nv, nf, int(ntau)
    g.nv_primary = nv if nv_primary is None else int(nv_primary)
    g.equpos = _ptr(equpos, _dp)
    g.fv = _ptr(fv, _ip)
    g.rev_primary, g.rev_secondary = float(rev_primary), float(rev_secondary)
    g.pol[:] = [float(x) for x in pol]
    g.rot[:] = [float(x) for x in rot]
    g.au_km = float(au_km)
    g.s_begin, g.s_count = (0, int(ntau)) if steps is None else (int(steps[0]), int(steps[1]))
    g.sun_scale = float(sun_scale)
    ns = g.s_count
    shadow = np.zeros((nf, ns), np.int32)
    cosang = np.zeros((nf, ns), np.float64)
    illum = np.zeros((nf, ns), np.float32)
    sun = np.zeros((nf, ns, 3), np.float32)
    solar = C.c_float()
    lib().orc_geometry(C.byref(g), _ptr(shadow, _ip), _ptr(cosang, _dp), _ptr(illum, _fp),
                       _ptr(sun, _fp), C.byref(solar))
    return Geometry(shadow, cosang, illum, sun, float(solar.value))


class OrcVfIn(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("nsub", C.c_int), ("cull", C.c_int),
        ("refine_min", C.c_float), ("vtx", _fp), ("fv", _ip),
    ]


class OrcVfOut(C.Structure):
    _fields_ = [
        ("n_facing", C.c_longlong), ("nnz", C.c_longlong), ("n_refined", C.c_longlong),
        ("rowptr", _ip), ("col", _ip), ("vf", _fp), ("vec", _fp), ("dir", _fp),
        ("refined", _up), ("total", _fp), ("photon", _fp),
    ]


class OrcDrfIn(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("vtx", _dp), ("fv", _ip), ("rowptr", _ip),
        ("vf", _fp), ("dir", _fp), ("ncv", C.c_int), ("ncf", C.c_int), ("cvtx", _fp),
        ("cfv", _ip), ("m", C.c_int), ("crater_radius", C.c_float),
    ]


@dataclass
class ViewFactors:
    rowptr: np.ndarray    # [nf+1] int32
    col: np.ndarray       # [nnz] int32: facet seen by the row's facet
    vf: np.ndarray        # [nnz] float32
    vec: np.ndarray       # [nnz][3] float32: A_I VF d/|d| (refined: sub-facet sum)
    dir: np.ndarray       # [nnz][3] float32: vec / (VF A_I), the vector-list value
    refined: np.ndarray   # [nnz] bool
    total: np.ndarray     # [nf] float32: total view factor
    photon: np.ndarray    # [nf][3] float32: photon momentum vector
    n_facing: int         # facing pairs before the occlusion cull

    @property
    def csr(self):
        return self.rowptr, self.col, self.vf


def viewfactors(vtx, fv, nsub=10, cull=True, refine_min=0.05):
    """Facet view factors of a mesh (viewfactor.c): vertices [nv][3] rounded
    to float as the producer reads them, facets [nf][3]."""
    vtx = np.ascontiguousarray(vtx, dtype=np.float32)
    fv = np.ascontiguousarray(fv, dtype=np.int32)
    nf = fv.shape[0]
    q = OrcVfIn()
    q.nv, q.nf, q.nsub, q.cull = vtx.shape[0], nf, int(nsub), int(bool(cull))
    q.refine_min = float(refine_min)
    q.vtx, q.fv = _ptr(vtx, _fp), _ptr(fv, _ip)
    o = OrcVfOut()
    if lib().orc_viewfactors(C.byref(q), C.byref(o)) != 0:
        raise RuntimeError("orc_viewfactors failed")
    try:
        n = int(o.nnz)

        def arr(p, count, dt):
            return np.ctypeslib.as_array(p, shape=(count,)).astype(dt, copy=True) if count else \
                np.zeros(0, dt)
        res = ViewFactors(arr(o.rowptr, nf + 1, np.int32), arr(o.col, n, np.int32),
                          arr(o.vf, n, np.float32), arr(o.vec, 3 * n, np.float32).reshape(n, 3),
                          arr(o.dir, 3 * n, np.float32).reshape(n, 3),
                          arr(o.refined, n, np.uint8).astype(bool), arr(o.total, nf, np.float32),
                          arr(o.photon, 3 * nf, np.float32).reshape(nf, 3), int(o.n_facing))
    finally:
        lib().orc_vf_free(C.byref(o))
    return res


def vf_subdivide(v9, n):
    """(vertex table, facet table, sub-facet centroids) of a triangle [3][3]
    split n x n, in the producer's order (viewfactor.c)."""
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(9)
    nv = (n + 1) * (n + 2) // 2
    vpos = np.zeros((nv, 3), np.float32)
    tri = np.zeros((n * n, 3), np.int32)
    mid = np.zeros((n * n, 3), np.float32)
    lib().orc_vf_subdivide(_ptr(v9, _fp), C.c_int(n), _ptr(vpos, _fp), _ptr(tri, _ip),
                           _ptr(mid, _fp))
    return vpos, tri, mid


def crater_drf(vtx, fv, rowptr, vf, dirs, cvtx, cfv, m=10, crater_radius=10.0):
    """Shape->crater view factors drf [nnz][ncf] (viewfactor.c): the mesh in
    double, the facet view-factor CSR rows and values with their vector-list
    directions, the crater mesh in float."""
    vtx = np.ascontiguousarray(vtx, dtype=np.float64)
    fv = np.ascontiguousarray(fv, dtype=np.int32)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    vf = np.ascontiguousarray(vf, dtype=np.float32)
    dirs = np.ascontiguousarray(dirs, dtype=np.float32)
    cvtx = np.ascontiguousarray(cvtx, dtype=np.float32)
    cfv = np.ascontiguousarray(cfv, dtype=np.int32)
    q = OrcDrfIn()
    q.nv, q.nf = vtx.shape[0], fv.shape[0]
    q.vtx, q.fv, q.rowptr = _ptr(vtx, _dp), _ptr(fv, _ip), _ptr(rowptr, _ip)
    q.vf, q.dir = _ptr(vf, _fp), _ptr(dirs, _fp)
    q.ncv, q.ncf = cvtx.shape[0], cfv.shape[0]
    q.cvtx, q.cfv = _ptr(cvtx, _fp), _ptr(cfv, _ip)
    q.m, q.crater_radius = int(m), float(crater_radius)
    out = np.zeros((int(rowptr[-1]), q.ncf), np.float32)
    lib().orc_crater_drf(C.byref(q), _ptr(out, _fp))
    return out


class OrcObserveIn(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("ntau", C.c_int), ("nv_primary", C.c_int),
        ("equpos", _dp), ("fv", _ip), ("rev_primary", C.c_double), ("rev_secondary", C.c_double),
        ("hel", C.c_double * 3), ("obs", C.c_double * 3), ("rot", C.c_double * 2),
        ("au_m", C.c_double), ("visible", _ip), ("stemp", _fp), ("rough", C.c_int),
        ("ncv", C.c_int), ("ncf", C.c_int), ("cvtx", _fp), ("cfv", _ip), ("cvis", _fp),
        ("ctemp", _fp), ("nwl", C.c_int), ("wavelength_um", _dp), ("colour", _dp),
        ("tmin", C.c_int), ("tcount", C.c_int), ("eps", C.c_double), ("s_begin", C.c_int),
        ("s_count", C.c_int),
    ]


AU_M = 149.6e9                                # observe_generator.cpp:9
WISE_UM = (4.6028, 11.0984, 22.6405)          # observe_generator.cpp:43


def observe(equpos, fv, ntau, hel, obs, rot, visible, stemp, colour, nv_primary=None,
            rev_primary=1.0, rev_secondary=1.0, crater=None, eps=0.95, wavelengths=WISE_UM,
            au_m=AU_M, steps=None):
    """Observed fluxes [s_count][3] (smooth, rough) of observe.c.  crater =
    (cvtx [ncv][3], cfv [ncf][3], cvis [nf][ntau][ncf], ctemp [nf][ntau][ncf])
    or None (no roughness); colour [3][tcount] for 100 K upwards."""
    eq = _d(equpos)
    f = _i(fv)
    vis = _i(visible)
    st = _f(stemp)
    col = _d(colour)
    wl = _d(wavelengths)
    q = OrcObserveIn()
    q.nv, q.nf, q.ntau = eq.shape[0], f.shape[0], int(ntau)
    q.nv_primary = eq.shape[0] if nv_primary is None else int(nv_primary)
    q.equpos, q.fv = _ptr(eq, _dp), _ptr(f, _ip)
    q.rev_primary, q.rev_secondary = float(rev_primary), float(rev_secondary)
    q.hel[:] = [float(x) for x in hel]
    q.obs[:] = [float(x) for x in obs]
    q.rot[:] = [float(x) for x in rot]
    q.au_m = float(au_m)
    q.visible, q.stemp = _ptr(vis, _ip), _ptr(st, _fp)
    keep = []
    if crater is not None:
        cv, cf, cvis, ct = (_f(crater[0]), _i(crater[1]), _f(crater[2]), _f(crater[3]))
        keep += [cv, cf, cvis, ct]
        q.rough, q.ncv, q.ncf = 1, cv.shape[0], cf.shape[0]
        q.cvtx, q.cfv, q.cvis, q.ctemp = (_ptr(cv, _fp), _ptr(cf, _ip), _ptr(cvis, _fp),
                                          _ptr(ct, _fp))
    q.nwl, q.wavelength_um, q.colour = len(wl), _ptr(wl, _dp), _ptr(col, _dp)
    q.tmin, q.tcount, q.eps = 100, col.shape[1], float(eps)
    q.s_begin, q.s_count = (0, int(ntau)) if steps is None else (int(steps[0]), int(steps[1]))
    sm = np.zeros((q.s_count, len(wl)), np.float64)
    ro = np.zeros((q.s_count, len(wl)), np.float64)
    lib().orc_observe(C.byref(q), _ptr(sm, _dp), _ptr(ro, _dp))
    return sm, ro


class OrcCraterSunIn(C.Structure):
    _fields_ = [
        ("nf", C.c_int), ("ntau", C.c_int), ("shadow_in", _ip), ("sun", _fp), ("ncv", C.c_int),
        ("ncf", C.c_int), ("cvtx", _fp), ("cfv", _ip), ("m", C.c_int),
        ("crater_radius", C.c_float),
    ]


def crater_sun(shadow, sun, cvtx, cfv, m=10, crater_radius=10.0):
    """(sunlit fraction, cosine) [nf][ntau][ncf] of the crater sub-facets
    (viewfactor.c orc_crater_sun)."""
    lit = _i(shadow)
    sv = _f(sun)
    cv, cf = _f(cvtx), _i(cfv)
    nf, ntau = lit.shape
    q = OrcCraterSunIn()
    q.nf, q.ntau, q.shadow_in, q.sun = nf, ntau, _ptr(lit, _ip), _ptr(sv, _fp)
    q.ncv, q.ncf, q.cvtx, q.cfv = cv.shape[0], cf.shape[0], _ptr(cv, _fp), _ptr(cf, _ip)
    q.m, q.crater_radius = int(m), float(crater_radius)
    sh = np.zeros((nf, ntau, cf.shape[0]), np.float32)
    co = np.zeros((nf, ntau, cf.shape[0]), np.float32)
    lib().orc_crater_sun(C.byref(q), _ptr(sh, _fp), _ptr(co, _fp))
    return sh, co
