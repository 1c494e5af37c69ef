"""ctypes mirror of include/monte.h (the C ABI of libmonte.so).

The shared library is built in-tree by monte_amd.build.  Importing this module
without it raises immediately: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmonte.so")

TABLE_ROWS = 201
MAX_LABELS = 16

OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_IO, ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
PRESET_STANDARD, PRESET_BP3D20 = 0, 1
FILTER_RAMLAK, FILTER_SHEPP_LOGAN = 0, 1
COH_FORWARD, COH_RAYLEIGH = 0, 1
SHAPE_ELLIPSOID, SHAPE_CYLINDER_Z, SHAPE_BOX, SHAPE_CYLINDER_X = 0, 1, 2, 3


class MonteError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"monte error {code}: {msg}")
        self.code = code


class Geom(C.Structure):
    _fields_ = [
        ("preset", C.c_int32), ("n_views", C.c_int32), ("nu", C.c_int32), ("nv", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("sod", C.c_double), ("sdd", C.c_double), ("du", C.c_double), ("dv", C.c_double),
        ("u_off", C.c_double), ("v_off", C.c_double),
        ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double),
        ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double),
        ("angle0", C.c_double), ("dangle", C.c_double),
        ("odd", C.c_double), ("scale", C.c_double), ("step_deg", C.c_int32),
        ("roi_x0", C.c_int32), ("roi_x1", C.c_int32), ("mask_r", C.c_double), ("filter", C.c_int32),
    ]


class Phantom(C.Structure):
    _fields_ = [
        ("labels", C.c_void_p), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_labels", C.c_int32),
        ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double),
        ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double),
        ("mat_table", C.c_int32 * MAX_LABELS), ("density", C.c_double * MAX_LABELS),
    ]


class McCfg(C.Structure):
    _fields_ = [
        ("photons_per_view", C.c_uint64), ("e_kev", C.c_double), ("spectrum", C.c_void_p),
        ("max_interactions", C.c_int32), ("coherent_mode", C.c_int32), ("seed", C.c_uint64),
        ("job", C.c_uint32), ("view_begin", C.c_int32), ("view_end", C.c_int32),
    ]


class McStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "histories", "steps", "interactions", "photo", "coherent", "compton",
        "tallied_primary", "tallied_scatter", "missed", "capped")] + [("kernel_ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class McJob(C.Structure):
    _fields_ = [
        ("geom", C.POINTER(Geom)), ("phantom", C.POINTER(Phantom)), ("cfg", C.POINTER(McCfg)),
        ("primary", C.c_void_p), ("scatter", C.c_void_p),
    ]


class E2EBuffers(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("primary", "scatter", "lineint", "vol", "vol_ref", "residual", "grad",
                                          "loss")]


class Tuning(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "mc_block", "mc_batch", "mc_blocks_per_cu", "bp_zchunk", "bp_tile_x", "bp_tile_y", "bp_view_chunk",
        "pj_zchunk", "mc_ray_brick")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Primitive(C.Structure):
    _fields_ = [
        ("shape", C.c_int32), ("label", C.c_int32),
        ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
        ("rx", C.c_double), ("ry", C.c_double), ("rz", C.c_double),
    ]


# every symbol include/monte.h declares, with its ctypes signature
_P = C.c_void_p
_CTX = C.c_void_p
SIGNATURES = {
    "monte_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "monte_destroy": (None, [_CTX]),
    "monte_set_stream": (C.c_int, [_CTX, _P]),
    "monte_set_pointer_mode": (C.c_int, [_CTX, C.c_int]),
    "monte_synchronize": (C.c_int, [_CTX]),
    "monte_status_str": (C.c_char_p, [C.c_int]),
    "monte_last_error": (C.c_char_p, [_CTX]),
    "monte_get_tuning": (C.c_int, [_CTX, C.POINTER(Tuning)]),
    "monte_set_tuning": (C.c_int, [_CTX, C.POINTER(Tuning)]),
    "monte_read_xcom_csv": (C.c_int, [C.c_char_p, _P]),
    "monte_load_tables": (C.c_int, [_CTX, _P, C.c_int]),
    "monte_fdk": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int]),
    "monte_fdk_filter": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int]),
    "monte_fdk_backproject": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int, C.c_int]),
    "monte_fdk_project": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int]),
    "monte_fdk_filter_adjoint": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int]),
    "monte_fdk_adjoint": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int]),
    "monte_mc_project": (C.c_int, [_CTX, C.POINTER(Geom), C.POINTER(Phantom), C.POINTER(McCfg), _P, _P,
                                   C.POINTER(McStats)]),
    "monte_mc_project_noalloc": (C.c_int, [_CTX, C.POINTER(Geom), C.POINTER(Phantom), C.POINTER(McCfg), _P, _P,
                                           C.POINTER(McStats)]),
    "monte_mc_first_flight_sets": (C.c_int, [_CTX, C.POINTER(Geom), C.POINTER(Phantom), C.c_int, _P]),
    "monte_mc_batch": (C.c_int, [_CTX, C.POINTER(McJob), C.c_int, C.POINTER(McStats)]),
    "monte_primary_exact": (C.c_int, [_CTX, C.POINTER(Geom), C.POINTER(Phantom), C.c_double, _P, _P]),
    "monte_counts_to_lineint": (C.c_int, [_CTX, C.POINTER(Geom), _P, C.c_uint64, _P]),
    "monte_counts_to_lineint_flat": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int, _P]),
    "monte_phantom_paint": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(Primitive), C.c_int]),
    "monte_autotune": (C.c_int, [_CTX, C.c_int, C.POINTER(Geom), C.POINTER(Phantom), C.POINTER(McCfg),
                                 C.POINTER(Tuning)]),
    "monte_fdk_graph": (C.c_int, [_CTX, C.POINTER(Geom), _P, _P, C.c_int]),
    "monte_e2e_step": (C.c_int, [_CTX, C.POINTER(Geom), C.POINTER(Phantom), C.POINTER(McCfg), C.POINTER(E2EBuffers),
                                 C.c_int, C.POINTER(McStats)]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m monte_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, ctx=None, what: str = "") -> int:
    if rc < 0:
        msg = lib().monte_status_str(rc).decode()
        if ctx is not None:
            detail = lib().monte_last_error(ctx).decode()
            if detail:
                msg = f"{msg}: {detail}"
        raise MonteError(rc, f"{what}: {msg}" if what else msg)
    return rc
