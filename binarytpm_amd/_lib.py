"""ctypes binding of libbtpm.so (the C ABI declared in include/btpm.h).

The library is loaded from this package directory; a missing library is an
error (there is no CPU fallback in the product path).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbtpm.so")

BTPM_OK = 0
BTPM_ERR_ARG = -1
BTPM_ERR_HIP = -2
BTPM_ERR_NOMEM = -3
BTPM_ERR_STATE = -4
NS = 100
NX = 56


class BtpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libbtpm error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("ntau", C.c_int), ("nx", C.c_int), ("rot_period", C.c_float), ("eps", C.c_float),
        ("sigma", C.c_float), ("ir_abs", C.c_float), ("tacc", C.c_float),
        ("newton_tol", C.c_float), ("max_rev", C.c_int), ("check_after", C.c_int),
        ("newton_cap", C.c_int),
    ]


class Body(C.Structure):
    _fields_ = [
        ("nf", C.c_int), ("illum", C.POINTER(C.c_float)), ("albedo", C.POINTER(C.c_float)),
        ("numrevs", C.POINTER(C.c_float)), ("nnz", C.c_int), ("rowptr", C.POINTER(C.c_int)),
        ("colidx", C.POINTER(C.c_int)), ("vf", C.POINTER(C.c_float)), ("scatter", C.c_int),
        ("scatter_acc", C.c_float),
    ]


class Crater(C.Structure):
    _fields_ = [
        ("vf", C.POINTER(C.c_float)), ("illum", C.POINTER(C.c_float)),
        ("shadow", C.POINTER(C.c_float)), ("cosang", C.POINTER(C.c_float)),
        ("solar", C.c_float), ("multiple", C.c_int), ("scatter_acc", C.c_float),
        ("drf", C.POINTER(C.c_float)),
        ("sun_local", C.POINTER(C.c_float)), ("vis_table", C.POINTER(C.c_float)),
        ("n_el", C.c_int), ("n_az", C.c_int), ("normals", C.POINTER(C.c_float)),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("rotations", C.c_int), ("converged", C.c_int), ("column_steps", C.c_longlong),
        ("crater_rotations", C.c_int), ("crater_converged", C.c_int),
        ("crater_column_steps", C.c_longlong), ("latency_s", C.c_double),
    ]


class BatchOpts(C.Structure):
    _fields_ = [
        ("rotations_per_sync", C.c_int), ("use_graph", C.c_int),
        ("max_active_craters", C.c_int), ("mem_fraction", C.c_double), ("profile", C.c_int),
        ("autotune", C.c_int), ("fast_extra", C.c_int),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("ntau", C.c_int), ("nv_primary", C.c_int),
        ("equpos", C.POINTER(C.c_double)), ("facets", C.POINTER(C.c_int)),
        ("rev_primary", C.c_double), ("rev_secondary", C.c_double),
        ("pol", C.c_double * 3), ("rot", C.c_double * 2), ("au_km", C.c_double),
        ("sun_scale", C.c_double),
    ]


class VfMesh(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("vtx", C.POINTER(C.c_float)),
        ("facets", C.POINTER(C.c_int)), ("nsub", C.c_int), ("cull", C.c_int),
        ("refine_min", C.c_float),
    ]


class DrfIn(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("vtx", C.POINTER(C.c_double)),
        ("facets", C.POINTER(C.c_int)), ("rowptr", C.POINTER(C.c_int)),
        ("vf", C.POINTER(C.c_float)), ("dir", C.POINTER(C.c_float)), ("ncv", C.c_int),
        ("ncf", C.c_int), ("crater_vtx", C.POINTER(C.c_float)),
        ("crater_facets", C.POINTER(C.c_int)), ("m", C.c_int), ("crater_radius", C.c_float),
    ]


class ObserveIn(C.Structure):
    _fields_ = [
        ("nv", C.c_int), ("nf", C.c_int), ("ntau", C.c_int), ("nv_primary", C.c_int),
        ("equpos", C.POINTER(C.c_double)), ("facets", C.POINTER(C.c_int)),
        ("rev_primary", C.c_double), ("rev_secondary", C.c_double),
        ("hel", C.c_double * 3), ("obs", C.c_double * 3), ("rot", C.c_double * 2),
        ("au_m", C.c_double), ("visible", C.POINTER(C.c_int)), ("stemp", C.POINTER(C.c_float)),
        ("rough", C.c_int), ("ncv", C.c_int), ("ncf", C.c_int),
        ("crater_vtx", C.POINTER(C.c_float)), ("crater_facets", C.POINTER(C.c_int)),
        ("crater_visible", C.POINTER(C.c_float)), ("crater_temp", C.POINTER(C.c_float)),
        ("wavelength_um", C.POINTER(C.c_double)), ("colour", C.POINTER(C.c_double)),
        ("tmin", C.c_int), ("tcount", C.c_int), ("eps", C.c_double),
    ]


class CraterSunIn(C.Structure):
    _fields_ = [
        ("nf", C.c_int), ("ntau", C.c_int), ("shadow", C.POINTER(C.c_int)),
        ("sun", C.POINTER(C.c_float)), ("ncv", C.c_int), ("ncf", C.c_int),
        ("crater_vtx", C.POINTER(C.c_float)), ("crater_facets", C.POINTER(C.c_int)),
        ("m", C.c_int), ("crater_radius", C.c_float),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double), ("solves", C.c_int), ("converged", C.c_int),
        ("column_steps", C.c_longlong), ("column_steps_all", C.c_longlong),
        ("graph_launches", C.c_int), ("syncs", C.c_int),
        ("smooth_kernel_s", C.c_double), ("crater_kernel_s", C.c_double),
        ("smooth_launches", C.c_int), ("crater_launches", C.c_int),
        ("smooth_column_steps", C.c_longlong), ("crater_column_steps", C.c_longlong),
        ("max_resident", C.c_int), ("deferrals", C.c_int),
        ("tune_rotations", C.c_int), ("tune_decisions", C.c_int),
        ("fast_rotations", C.c_longlong), ("mem_waits", C.c_int),
        ("heat_kernel_s", C.c_double), ("heat_launches", C.c_int),
        ("heat_set_rotations", C.c_longlong),
    ]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)

# (name, restype, argtypes)
PROTOTYPES = [
    ("btpm_version", C.c_int, []),
    ("btpm_last_error", C.c_char_p, []),
    ("btpm_params_default", None, [C.POINTER(Params), C.c_int, C.c_float, C.c_float]),
    ("btpm_batch_opts_default", None, [C.POINTER(BatchOpts)]),
    ("btpm_ctx_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("btpm_ctx_destroy", C.c_int, [_vp]),
    ("btpm_ctx_mem_info", C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("btpm_ctx_set_mem_limit", C.c_int, [_vp, C.c_size_t]),
    ("btpm_shape_create", C.c_int, [_vp, C.POINTER(Params), C.POINTER(Body), C.POINTER(Crater),
                                    C.POINTER(_vp)]),
    ("btpm_shape_destroy", C.c_int, [_vp]),
    ("btpm_shape_facet_illum", C.c_int, [_vp, _fp]),
    ("btpm_shape_crater_illum", C.c_int, [_vp, _fp]),
    ("btpm_solve_create", C.c_int, [_vp, _vp, C.c_float, _fp, C.c_int, C.POINTER(_vp)]),
    ("btpm_solve_destroy", C.c_int, [_vp]),
    ("btpm_solve_set_flux_scale", C.c_int, [_vp, C.c_float]),
    ("btpm_solve_warm_start", C.c_int, [_vp, _vp]),
    ("btpm_solve_run_stepwise", C.c_int, [_vp, C.c_int]),
    ("btpm_solve_run", C.c_int, [_vp]),
    ("btpm_solve_get_info", C.c_int, [_vp, C.POINTER(SolveInfo)]),
    ("btpm_solve_surface_temps", C.c_int, [_vp, _fp]),
    ("btpm_solve_crater_temps", C.c_int, [_vp, _fp]),
    ("btpm_solve_profile", C.c_int, [_vp, _fp]),
    ("btpm_solve_crater_profile", C.c_int, [_vp, _fp]),
    ("btpm_batch_run", C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(BatchOpts),
                                 C.POINTER(BatchStats)]),
    ("btpm_server_start", C.c_int, [_vp, C.POINTER(BatchOpts), C.POINTER(_vp)]),
    ("btpm_server_submit", C.c_int, [_vp, _vp]),
    ("btpm_server_wait", C.c_int, [_vp, _vp, C.c_double]),
    ("btpm_server_wait_any", C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_double,
                                       C.POINTER(C.c_int)]),
    ("btpm_server_stop", C.c_int, [_vp, C.POINTER(BatchStats)]),
    ("btpm_crater_contract", C.c_int, [_vp, _fp, _fp, C.c_int, _fp]),
    ("btpm_crater_contract_noalloc", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp]),
    ("btpm_geometry_illumination", C.c_int,
     [_vp, C.POINTER(Geometry), C.POINTER(C.c_int), _fp, _fp, _fp]),
    ("btpm_vf_build", C.c_int, [_vp, C.POINTER(VfMesh), C.POINTER(_vp)]),
    ("btpm_vf_counts", C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                 C.POINTER(C.c_longlong)]),
    ("btpm_vf_fetch", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), _fp, _fp, _fp, _fp,
                                _fp]),
    ("btpm_vf_free", None, [_vp]),
    ("btpm_crater_viewfactors", C.c_int, [_vp, C.POINTER(DrfIn), _fp]),
    ("btpm_crater_sunlight", C.c_int, [_vp, C.POINTER(CraterSunIn), _fp, _fp]),
    ("btpm_write_temperature_txt", C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.c_int,
                                             C.POINTER(C.c_int), C.c_int, _fp, C.c_int, C.c_int]),
    ("btpm_write_crater_dat", C.c_int, [C.c_char_p, _fp, C.c_int, C.POINTER(C.c_int), C.c_int,
                                        _fp, C.c_int, C.c_int]),
    ("btpm_solve_write_crater_dat", C.c_int, [_vp, C.c_char_p, _fp, C.c_int, C.POINTER(C.c_int),
                                              C.c_int]),
    ("btpm_stitch_crater_parts", C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.c_int,
                                           C.c_int, C.c_int]),
    ("btpm_observe", C.c_int, [_vp, C.POINTER(ObserveIn), C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]),
    ("btpm_observe_solve", C.c_int, [_vp, C.POINTER(ObserveIn), _vp, _fp, C.c_int, C.c_float,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]),
]

_LIB = None


def lib():
    """Load libbtpm.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m binarytpm_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in PROTOTYPES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(code):
    if code != BTPM_OK:
        msg = lib().btpm_last_error()
        raise BtpmError(code, msg.decode() if msg else "")
    return code
