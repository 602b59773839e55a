"""Python front end of libbtpm: the Thermal stage's solve objects.

Mirrors the reference's thermal-stage operations (single_model's per-TI
solve: initial temperatures, rotations to convergence with per-facet
freezing, then the crater columns driven by the shape->crater heating;
175706/Run-04/thermal.cpp:3465-3818) over the C ABI of include/btpm.h.
Arrays use the reference layouts: facet series [nf][ntau], crater series
[nf][ntau][100], crater view factors vf[i*100+j] = VF(j->i).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import (NS, NX, BatchOpts, BatchStats, Body, BtpmError, Crater, CraterSunIn, DrfIn,
                   Geometry, ObserveIn, Params, SolveInfo, VfMesh, check, lib)

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _f32(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


def _ptr(a, t=_fp):
    return None if a is None else a.ctypes.data_as(t)


def _out(out, shape):
    """A C-contiguous float32 output array of `shape`: `out` itself (checked)
    or a new one."""
    if out is None:
        return np.empty(shape, dtype=np.float32)
    if out.dtype != np.float32 or out.shape != tuple(shape) or not out.flags.c_contiguous:
        raise ValueError(f"out must be a C-contiguous float32 array of shape {tuple(shape)}")
    return out


def params(ntau: int, rot_period: float, tacc: float = 0.01, **kw) -> Params:
    """Model constants (reference globals, thermal.cpp:14-83); keyword
    arguments override eps, sigma, ir_abs, newton_tol, max_rev, check_after,
    newton_cap."""
    p = Params()
    lib().btpm_params_default(C.byref(p), int(ntau), float(rot_period), float(tacc))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Context:
    """One GPU and its stream (set_cuda_device, thermal_GPU.cu:698-702)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().btpm_ctx_create(int(device), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def mem_info(self):
        fr, tot = C.c_size_t(), C.c_size_t()
        check(lib().btpm_ctx_mem_info(self._h, C.byref(fr), C.byref(tot)))
        return fr.value, tot.value

    def set_mem_limit(self, nbytes: int):
        """Cap the device memory the library holds on this device (0: none)."""
        check(lib().btpm_ctx_set_mem_limit(self._h, int(nbytes)))

    def close(self):
        if self._h:
            lib().btpm_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


@dataclass
class CraterModel:
    """Crater (roughness) inputs; see btpm_crater in include/btpm.h."""
    vf: np.ndarray                        # [100][100], vf[i*100+j] = VF(j->i)
    illum: Optional[np.ndarray] = None    # [nf][ntau][100] total illumination
    shadow: Optional[np.ndarray] = None   # [nf][ntau][100] sunlit fraction
    cosang: Optional[np.ndarray] = None   # [nf][ntau][100] cos incidence
    solar: float = 0.0
    multiple: bool = True
    scatter_acc: float = 1e-4
    drf: Optional[np.ndarray] = None      # [nnz][100] shape->crater view factors
    # direct sunlight evaluated on the GPU instead of shadow/cosang:
    sun_local: Optional[np.ndarray] = None  # [nf][ntau][3] Sun in the crater frame x sunlit fraction
    vis_table: Optional[np.ndarray] = None  # [n_el][n_az][100] sub-facet sunlit fraction
    normals: Optional[np.ndarray] = None    # [100][3] sub-facet normals


class Shape:
    """Device-resident inputs shared by solves (facet illumination, view
    factors, crater model and crater illumination after scattering)."""

    def __init__(self, ctx: Context, p: Params, illum, albedo, numrevs=None, csr=None,
                 scatter: int = 0, scatter_acc: float = 1e-4,
                 crater: Optional[CraterModel] = None):
        illum = _f32(illum)
        nf, nt = illum.shape
        if nt != p.ntau:
            raise ValueError(f"illum has {nt} steps, params.ntau = {p.ntau}")
        self.nf, self.ntau = nf, nt
        self.params = p
        keep = [illum]
        albedo = _f32(np.broadcast_to(np.asarray(albedo, dtype=np.float32), (nf,)))
        keep.append(albedo)
        nr = None if numrevs is None else _f32(np.broadcast_to(np.asarray(numrevs, np.float32), (nf,)))
        keep.append(nr)
        b = Body()
        b.nf = nf
        b.illum, b.albedo, b.numrevs = _ptr(illum), _ptr(albedo), _ptr(nr)
        b.scatter, b.scatter_acc = int(scatter), float(scatter_acc)
        self.nnz = 0
        if csr is not None:
            rp = np.ascontiguousarray(csr[0], dtype=np.int32)
            ci = np.ascontiguousarray(csr[1], dtype=np.int32)
            vv = _f32(csr[2])
            keep += [rp, ci, vv]
            b.nnz = int(rp[-1])
            b.rowptr, b.colidx, b.vf = _ptr(rp, _ip), _ptr(ci, _ip), _ptr(vv)
            self.nnz = b.nnz
        cr_p = None
        self.has_crater = crater is not None
        if crater is not None:
            cr = Crater()
            vf = _f32(crater.vf, (NS * NS,))
            keep.append(vf)
            cr.vf = _ptr(vf)
            for name in ("illum", "shadow", "cosang"):
                arr = getattr(crater, name)
                if arr is not None:
                    arr = _f32(arr, (nf, nt, NS))
                    keep.append(arr)
                    setattr(cr, name, _ptr(arr))
            if crater.sun_local is not None:
                sun = _f32(crater.sun_local, (nf, nt, 3))
                vis = _f32(crater.vis_table)
                if vis.ndim != 3 or vis.shape[2] != NS:
                    raise ValueError("vis_table must be [n_el][n_az][100]")
                nrm = _f32(crater.normals, (NS, 3))
                keep += [sun, vis, nrm]
                cr.sun_local, cr.vis_table, cr.normals = _ptr(sun), _ptr(vis), _ptr(nrm)
                cr.n_el, cr.n_az = int(vis.shape[0]), int(vis.shape[1])
            if crater.drf is not None:
                drf = _f32(crater.drf, (self.nnz, NS))
                keep.append(drf)
                cr.drf = _ptr(drf)
            cr.solar = float(crater.solar)
            cr.multiple = int(bool(crater.multiple))
            cr.scatter_acc = float(crater.scatter_acc)
            cr_p = C.byref(cr)
        h = C.c_void_p()
        check(lib().btpm_shape_create(ctx.handle, C.byref(p), C.byref(b), cr_p, C.byref(h)))
        del keep
        self._h = h
        self.ctx = ctx

    @property
    def handle(self):
        return self._h

    def facet_illum(self) -> np.ndarray:
        out = np.empty((self.nf, self.ntau), dtype=np.float32)
        check(lib().btpm_shape_facet_illum(self._h, _ptr(out)))
        return out

    def crater_illum(self) -> np.ndarray:
        out = np.empty((self.nf, self.ntau, NS), dtype=np.float32)
        check(lib().btpm_shape_crater_illum(self._h, _ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().btpm_shape_destroy(self._h)
            self._h = None


class Solve:
    """One thermal-inertia solve of a shape (smooth facets, then craters)."""

    def __init__(self, ctx: Context, shape: Shape, ti: float, ti_per_facet=None,
                 with_craters: bool = False, flux_scale: float = 1.0,
                 warm_from: Optional["Solve"] = None):
        tpf = None if ti_per_facet is None else _f32(ti_per_facet, (shape.nf,))
        h = C.c_void_p()
        check(lib().btpm_solve_create(ctx.handle, shape.handle, float(ti), _ptr(tpf),
                                      int(bool(with_craters)), C.byref(h)))
        self._h = h
        rc = 0
        if flux_scale != 1.0:
            rc = lib().btpm_solve_set_flux_scale(h, float(flux_scale))
        if rc == 0 and warm_from is not None:
            rc = lib().btpm_solve_warm_start(h, warm_from.handle)
        if rc != 0:
            msg = lib().btpm_last_error()
            lib().btpm_solve_destroy(h)
            self._h = None
            raise BtpmError(rc, msg.decode() if msg else "")
        self.shape = shape
        self.ti = float(ti)
        self.with_craters = bool(with_craters)

    @property
    def handle(self):
        return self._h

    def run(self):
        """Run alone with a host convergence check every rotation."""
        check(lib().btpm_solve_run(self._h))
        return self

    def run_stepwise(self, rotation_cap: int = 0):
        """Run alone in the reference's GPU execution structure (two launches
        and a host sync per time step); rotation_cap > 0 limits each phase."""
        check(lib().btpm_solve_run_stepwise(self._h, int(rotation_cap)))
        return self

    def info(self) -> SolveInfo:
        inf = SolveInfo()
        check(lib().btpm_solve_get_info(self._h, C.byref(inf)))
        return inf

    def surface_temps(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """[nf][ntau] final surface temperatures (into `out` when given)."""
        shape = (self.shape.nf, self.shape.ntau)
        out = _out(out, shape)
        check(lib().btpm_solve_surface_temps(self._h, _ptr(out)))
        return out

    def crater_temps(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """[nf][ntau][100] final crater temperatures (into `out` when given)."""
        shape = (self.shape.nf, self.shape.ntau, NS)
        out = _out(out, shape)
        check(lib().btpm_solve_crater_temps(self._h, _ptr(out)))
        return out

    def write_crater_dat(self, path, Vc, Fc):
        """The crater temperature file (binary_writesurfacetemperature) straight
        from device memory, streamed in facet chunks."""
        cv = _f32(Vc)
        cf = np.ascontiguousarray(Fc, dtype=np.int32)
        check(lib().btpm_solve_write_crater_dat(self._h, os.fsencode(path), _ptr(cv), cv.shape[0],
                                                _ptr(cf, _ip), cf.shape[0]))

    def profile(self) -> np.ndarray:
        out = np.empty((self.shape.nf, NX + 1), dtype=np.float32)
        check(lib().btpm_solve_profile(self._h, _ptr(out)))
        return out

    def crater_profile(self) -> np.ndarray:
        out = np.empty((self.shape.nf, NX + 1, NS), dtype=np.float32)
        check(lib().btpm_solve_crater_profile(self._h, _ptr(out)))
        return out

    def close(self):
        """Destroy the solve (an error while another solve still has to take
        its state as a warm start; the handle stays valid then)."""
        if self._h:
            check(lib().btpm_solve_destroy(self._h))
            self._h = None


def batch_opts(rotations_per_sync=4, use_graph=True, max_active_craters=0,
               mem_fraction=0.8, profile=False, autotune=True, fast_extra=3) -> BatchOpts:
    o = BatchOpts()
    lib().btpm_batch_opts_default(C.byref(o))
    o.profile = int(bool(profile))
    o.autotune = int(bool(autotune))
    o.fast_extra = int(fast_extra)
    o.rotations_per_sync = int(rotations_per_sync)
    o.use_graph = int(bool(use_graph))
    o.max_active_craters = int(max_active_craters)
    o.mem_fraction = float(mem_fraction)
    return o


def batch_run(ctx: Context, solves: Sequence[Solve], opts: Optional[BatchOpts] = None
              ) -> BatchStats:
    arr = (C.c_void_p * len(solves))(*[s.handle for s in solves])
    st = BatchStats()
    o = opts if opts is not None else batch_opts()
    check(lib().btpm_batch_run(ctx.handle, arr, len(solves), C.byref(o), C.byref(st)))
    return st


class Server:
    """Continuous-batching request loop (btpm_server_*)."""

    def __init__(self, ctx: Context, opts: Optional[BatchOpts] = None):
        h = C.c_void_p()
        o = opts if opts is not None else batch_opts()
        check(lib().btpm_server_start(ctx.handle, C.byref(o), C.byref(h)))
        self._h = h

    def submit(self, s: Solve):
        check(lib().btpm_server_submit(self._h, s.handle))

    def wait(self, s: Solve, timeout_s: float = 600.0):
        check(lib().btpm_server_wait(self._h, s.handle, float(timeout_s)))

    def wait_any(self, solves: Sequence["Solve"], timeout_s: float = 600.0) -> int:
        """Index of a completed solve among `solves` (blocks until one is)."""
        arr = (C.c_void_p * len(solves))(*[s.handle for s in solves])
        idx = C.c_int(-1)
        check(lib().btpm_server_wait_any(self._h, arr, len(solves), float(timeout_s),
                                         C.byref(idx)))
        return idx.value

    def stop(self) -> BatchStats:
        st = BatchStats()
        h, self._h = self._h, None
        if h:
            check(lib().btpm_server_stop(h, C.byref(st)))
        return st


AU_KM = 149.6e6  # astronomical unit of the reference, km


@dataclass
class Illumination:
    shadow: np.ndarray     # [nf][ntau] int32: 1 lit, 0 back-facing, shadowed or eclipsed
    illum: np.ndarray      # [nf][ntau] float32: direct illumination, W/m^2
    sun_facet: Optional[np.ndarray]  # [nf][ntau][3] unit Sun direction, facet frame
    solar: float


def geometry_illumination(ctx: Context, equpos, facets, ntau: int, pol, rot, nv_primary=None,
                          rev_primary=1.0, rev_secondary=1.0, au_km=AU_KM,
                          sun_facet=True, sun_scale=0.0) -> Illumination:
    """Direct illumination and shadows of a (binary) shape over one cycle on
    the GPU: vertices equpos [nv][3] (km, each body's equatorial frame, the
    first nv_primary on the primary), facets [nf][3]; pol = (r AU, ecliptic
    latitude, longitude) and rot = (pole latitude, longitude) in radians."""
    eq = np.ascontiguousarray(equpos, dtype=np.float64)
    fv = np.ascontiguousarray(facets, dtype=np.int32)
    nv, nf = eq.shape[0], fv.shape[0]
    g = Geometry()
    g.nv, g.nf, g.ntau = nv, nf, int(ntau)
    g.nv_primary = nv if nv_primary is None else int(nv_primary)
    g.equpos = eq.ctypes.data_as(C.POINTER(C.c_double))
    g.facets = fv.ctypes.data_as(C.POINTER(C.c_int))
    g.rev_primary, g.rev_secondary = float(rev_primary), float(rev_secondary)
    g.pol[:] = [float(x) for x in pol]
    g.rot[:] = [float(x) for x in rot]
    g.au_km = float(au_km)
    g.sun_scale = float(sun_scale)
    shadow = np.empty((nf, ntau), np.int32)
    illum = np.empty((nf, ntau), np.float32)
    sun = np.empty((nf, ntau, 3), np.float32) if sun_facet else None
    solar = C.c_float()
    check(lib().btpm_geometry_illumination(ctx.handle, C.byref(g),
                                           shadow.ctypes.data_as(C.POINTER(C.c_int)),
                                           _ptr(illum), _ptr(sun) if sun is not None else None,
                                           C.byref(solar)))
    return Illumination(shadow, illum, sun, float(solar.value))


@dataclass
class ViewFactors:
    """Facet view-factor lists of a shape (the Shape_Selfheating outputs)."""
    rowptr: np.ndarray    # [nf+1] int32
    col: np.ndarray       # [nnz] int32: facet seen by the row's facet, increasing
    vf: np.ndarray        # [nnz] float32
    dir: np.ndarray       # [nnz][3] float32: _selfheating_vector_list.txt value
    vec: np.ndarray       # [nnz][3] float32: area-weighted view-factor vector
    total: np.ndarray     # [nf] float32: total view factor
    photon: np.ndarray    # [nf][3] float32: photon momentum vector (_photovectors.txt)
    n_facing: int         # facing pairs before the occlusion cull
    n_refined: int        # entries refined by sub-facet quadrature

    @property
    def csr(self):
        return self.rowptr, self.col, self.vf


def viewfactors(ctx: Context, vtx, facets, nsub: int = 10, cull: bool = True,
                refine_min: float = 0.05) -> ViewFactors:
    """Facet view factors of a mesh on the GPU (Shape_Selfheating): vertices
    [nv][3] (rounded to float, as the producer reads them), facets [nf][3].
    nsub = sub-facet split per edge of pairs with VF >= refine_min (0: none);
    cull = drop pairs hidden by a third facet."""
    v = np.ascontiguousarray(vtx, dtype=np.float32)
    fv = np.ascontiguousarray(facets, dtype=np.int32)
    nf = fv.shape[0]
    m = VfMesh()
    m.nv, m.nf = v.shape[0], nf
    m.vtx, m.facets = _ptr(v), _ptr(fv, _ip)
    m.nsub, m.cull, m.refine_min = int(nsub), int(bool(cull)), float(refine_min)
    h = C.c_void_p()
    check(lib().btpm_vf_build(ctx.handle, C.byref(m), C.byref(h)))
    try:
        n1, n, nr = C.c_longlong(), C.c_longlong(), C.c_longlong()
        check(lib().btpm_vf_counts(h, C.byref(n1), C.byref(n), C.byref(nr)))
        nnz = int(n.value)
        rowptr = np.empty(nf + 1, np.int32)
        col = np.empty(nnz, np.int32)
        vf = np.empty(nnz, np.float32)
        dirs = np.empty((nnz, 3), np.float32)
        vec = np.empty((nnz, 3), np.float32)
        total = np.empty(nf, np.float32)
        photon = np.empty((nf, 3), np.float32)
        check(lib().btpm_vf_fetch(h, _ptr(rowptr, _ip), _ptr(col, _ip), _ptr(vf), _ptr(dirs),
                                  _ptr(vec), _ptr(total), _ptr(photon)))
    finally:
        lib().btpm_vf_free(h)
    return ViewFactors(rowptr, col, vf, dirs, vec, total, photon, int(n1.value), int(nr.value))


def crater_viewfactors(ctx: Context, vtx, facets, rowptr, vf, dirs, crater_vtx, crater_facets,
                       m: int = 10, crater_radius: float = 10.0) -> np.ndarray:
    """Shape->crater view factors drf [nnz][ncf] on the GPU
    (calculate_surface_global_viewfactors): the shape in double, its facet
    view-factor rows with their vector-list directions, the crater model."""
    v = np.ascontiguousarray(vtx, dtype=np.float64)
    fv = np.ascontiguousarray(facets, dtype=np.int32)
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    vfs = np.ascontiguousarray(vf, dtype=np.float32)
    d = np.ascontiguousarray(dirs, dtype=np.float32)
    cv = np.ascontiguousarray(crater_vtx, dtype=np.float32)
    cf = np.ascontiguousarray(crater_facets, dtype=np.int32)
    q = DrfIn()
    q.nv, q.nf = v.shape[0], fv.shape[0]
    q.vtx, q.facets = v.ctypes.data_as(C.POINTER(C.c_double)), _ptr(fv, _ip)
    q.rowptr, q.vf, q.dir = _ptr(rp, _ip), _ptr(vfs), _ptr(d)
    q.ncv, q.ncf = cv.shape[0], cf.shape[0]
    q.crater_vtx, q.crater_facets = _ptr(cv), _ptr(cf, _ip)
    q.m, q.crater_radius = int(m), float(crater_radius)
    out = np.empty((int(rp[-1]), q.ncf), np.float32)
    check(lib().btpm_crater_viewfactors(ctx.handle, C.byref(q), _ptr(out)))
    return out


def write_temperature_txt(path, V, F, stemp, threads: int = 0):
    """writeglobaltemperature's text file, formatted natively on `threads`
    threads (0: all cores); byte-identical to formats.write_temperature_txt."""
    v = np.ascontiguousarray(V, dtype=np.float64)
    f = np.ascontiguousarray(F, dtype=np.int32)
    t = _f32(stemp)
    nf, ntau = t.shape
    check(lib().btpm_write_temperature_txt(os.fsencode(path), v.ctypes.data_as(C.POINTER(C.c_double)),
                                           v.shape[0], _ptr(f, _ip), f.shape[0], _ptr(t), ntau,
                                           int(threads)))


def write_crater_dat(path, Vc, Fc, temps):
    """binary_writesurfacetemperature's file from host temperatures
    [nf][ntau][ncf] (transposed a facet at a time)."""
    cv = _f32(Vc)
    cf = np.ascontiguousarray(Fc, dtype=np.int32)
    t = _f32(temps)
    nf, ntau, ncf = t.shape
    check(lib().btpm_write_crater_dat(os.fsencode(path), _ptr(cv), cv.shape[0], _ptr(cf, _ip), ncf,
                                      _ptr(t), nf, ntau))


def stitch_crater_parts(out_path, parts, ncv: int, ncf: int, ntau: int):
    """stitch_temperatures: join part files of consecutive facet ranges."""
    arr = (C.c_char_p * len(parts))(*[os.fsencode(p) for p in parts])
    check(lib().btpm_stitch_crater_parts(os.fsencode(out_path), arr, len(parts), int(ncv), int(ncf),
                                         int(ntau)))


AU_M = 149.6e9                          # observe_generator.cpp:9 (metres)
WISE_UM = (4.6028, 11.0984, 22.6405)    # observe_generator.cpp:43


def observe(ctx: Context, equpos, facets, ntau: int, hel, obs, rot, visible, stemp, colour,
            nv_primary=None, rev_primary=1.0, rev_secondary=1.0, crater=None, eps=0.95,
            wavelengths=WISE_UM, au_m=AU_M, tmin=100, solve=None, observer_facet=None,
            crater_m: int = 10, crater_radius: float = 10.0):
    """Observed fluxes ([ntau][3] smooth, [ntau][3] rough) on the GPU (the
    Observe stage's flux synthesis): the shape and its motion as for
    geometry_illumination, hel / obs = (r AU, ecliptic latitude, longitude)
    of the system from the Sun / from the observer, visible [nf][ntau] facet
    flags, stemp [nf][ntau]; crater = (vertices [ncv][3], facets [ncf][3],
    visible fractions [nf][ntau][ncf], temperatures [nf][ntau][ncf]) or None;
    colour [3][tcount] corrections from tmin K up.  With solve= a finished
    Solve, the surface and crater temperatures are that solve's, read in
    device memory (btpm_observe_solve): stemp and the crater temperatures
    may then be None; with observer_facet ([nf][ntau][3], the observer's
    direction in each facet frame) as well, the crater visible fractions are
    computed on the device (crater_sunlight with the observer as the source,
    crater_m x crater_m rim samples) and may be None too."""
    eq = np.ascontiguousarray(equpos, dtype=np.float64)
    fv = np.ascontiguousarray(facets, dtype=np.int32)
    vis = np.ascontiguousarray(visible, dtype=np.int32)
    col = np.ascontiguousarray(colour, dtype=np.float64)
    wl = np.ascontiguousarray(wavelengths, dtype=np.float64)
    nv, nf = eq.shape[0], fv.shape[0]
    st = _f32(stemp) if stemp is not None or solve is None else None
    if (vis.shape != (nf, ntau) or (st is not None and st.shape != (nf, ntau))
            or wl.shape != (3,) or col.shape[0] != 3):
        raise ValueError("visible/stemp must be [nf][ntau], wavelengths [3], colour [3][n]")
    q = ObserveIn()
    q.nv, q.nf, q.ntau = nv, nf, int(ntau)
    q.nv_primary = nv if nv_primary is None else int(nv_primary)
    q.equpos = eq.ctypes.data_as(C.POINTER(C.c_double))
    q.facets = _ptr(fv, _ip)
    q.rev_primary, q.rev_secondary = float(rev_primary), float(rev_secondary)
    q.hel[:] = [float(x) for x in hel]
    q.obs[:] = [float(x) for x in obs]
    q.rot[:] = [float(x) for x in rot]
    q.au_m = float(au_m)
    q.visible = _ptr(vis, _ip)
    if st is not None:
        q.stemp = _ptr(st)
    keep = []
    if crater is not None:
        cv = _f32(crater[0])
        cf = np.ascontiguousarray(crater[1], dtype=np.int32)
        on_dev = solve is not None and observer_facet is not None
        cvis = _f32(crater[2]) if crater[2] is not None or not on_dev else None
        ct = _f32(crater[3]) if crater[3] is not None or solve is None else None
        want = (nf, ntau, cf.shape[0])
        if (cvis is not None and cvis.shape != want) or (ct is not None and ct.shape != want):
            raise ValueError("crater arrays must be [nf][ntau][ncf]")
        keep += [cv, cf, cvis, ct]
        q.rough, q.ncv, q.ncf = 1, cv.shape[0], cf.shape[0]
        q.crater_vtx, q.crater_facets = _ptr(cv), _ptr(cf, _ip)
        if cvis is not None:
            q.crater_visible = _ptr(cvis)
        if ct is not None:
            q.crater_temp = _ptr(ct)
    q.wavelength_um = wl.ctypes.data_as(C.POINTER(C.c_double))
    q.colour = col.ctypes.data_as(C.POINTER(C.c_double))
    q.tmin, q.tcount, q.eps = int(tmin), col.shape[1], float(eps)
    sm = np.empty((ntau, 3), np.float64)
    ro = np.empty((ntau, 3), np.float64)
    if solve is None:
        check(lib().btpm_observe(ctx.handle, C.byref(q), sm.ctypes.data_as(C.POINTER(C.c_double)),
                                 ro.ctypes.data_as(C.POINTER(C.c_double))))
    else:
        ob = None if observer_facet is None else _f32(observer_facet, (nf, int(ntau), 3))
        check(lib().btpm_observe_solve(ctx.handle, C.byref(q), solve.handle, _ptr(ob),
                                       int(crater_m), float(crater_radius),
                                       sm.ctypes.data_as(C.POINTER(C.c_double)),
                                       ro.ctypes.data_as(C.POINTER(C.c_double))))
    return sm, ro


def crater_sunlight(ctx: Context, shadow, sun, crater_vtx, crater_facets, m: int = 10,
                    crater_radius: float = 10.0):
    """Crater sunlight (the reference's surfacegeometry and sfcpartialshadow):
    facet lit flags [nf][ntau] and Sun vectors in the facet frames
    [nf][ntau][3] -> (sunlit fraction, cosine), each [nf][ntau][ncf]."""
    lit = np.ascontiguousarray(shadow, dtype=np.int32)
    sv = _f32(sun)
    cv = _f32(crater_vtx)
    cf = np.ascontiguousarray(crater_facets, dtype=np.int32)
    nf, ntau = lit.shape
    if sv.shape != (nf, ntau, 3):
        raise ValueError("sun must be [nf][ntau][3]")
    q = CraterSunIn()
    q.nf, q.ntau, q.shadow, q.sun = nf, ntau, _ptr(lit, _ip), _ptr(sv)
    q.ncv, q.ncf, q.crater_vtx, q.crater_facets = cv.shape[0], cf.shape[0], _ptr(cv), _ptr(cf, _ip)
    q.m, q.crater_radius = int(m), float(crater_radius)
    sh = np.empty((nf, ntau, cf.shape[0]), np.float32)
    co = np.empty((nf, ntau, cf.shape[0]), np.float32)
    check(lib().btpm_crater_sunlight(ctx.handle, C.byref(q), _ptr(sh), _ptr(co)))
    return sh, co


def crater_contract(ctx: Context, vf, e) -> np.ndarray:
    """q[n][j] = sum_i VF(j->i) e[n][i] through the MFMA split product."""
    vf = _f32(vf, (NS * NS,))
    e = _f32(e)
    n = e.shape[0]
    q = np.empty((n, NS), dtype=np.float32)
    check(lib().btpm_crater_contract(ctx.handle, _ptr(vf), _ptr(e), n, _ptr(q)))
    return q
