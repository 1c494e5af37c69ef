"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the correctness reference or the timed CPU baseline.  The
product package (monte_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PRESET_STANDARD, PRESET_BP3D20 = 0, 1
FILTER_RAMLAK, FILTER_SHEPP_LOGAN = 0, 1
COH_FORWARD, COH_RAYLEIGH = 0, 1
NROWS = 201
MAX_MAT = 16


class Geom(C.Structure):
    _fields_ = [
        ("preset", C.c_int), ("n_views", C.c_int), ("nu", C.c_int), ("nv", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("sod", C.c_double), ("sdd", C.c_double), ("du", C.c_double), ("dv", C.c_double),
        ("u_off", C.c_double), ("v_off", C.c_double),
        ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double),
        ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double),
        ("angle0", C.c_double), ("dangle", C.c_double),
        ("odd", C.c_double), ("scale", C.c_double), ("step_deg", C.c_int),
        ("roi_x0", C.c_int), ("roi_x1", C.c_int), ("mask_r", C.c_double), ("filter", C.c_int),
    ]


class Phantom(C.Structure):
    _fields_ = [
        ("labels", C.c_void_p), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("vx", C.c_double), ("vy", C.c_double), ("vz", C.c_double),
        ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double),
        ("n_labels", C.c_int), ("mat_table", C.c_int * MAX_MAT), ("density", C.c_double * MAX_MAT),
    ]


class Tables(C.Structure):
    _fields_ = [("n_tables", C.c_int), ("tab", C.c_void_p)]


class McCfg(C.Structure):
    _fields_ = [
        ("photons_per_view", C.c_uint64), ("e_kev", C.c_double), ("spectrum_cdf", C.c_void_p),
        ("max_interactions", C.c_int), ("coherent_mode", C.c_int),
        ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32), ("job", C.c_uint32),
        ("view_begin", C.c_int), ("view_end", C.c_int), ("ray_brick", C.c_int),
    ]


class McStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "histories", "steps", "interactions", "photo", "coherent", "compton",
        "tallied_primary", "tallied_scatter", "missed", "capped")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_u01.restype = C.c_double
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_read_xcom_csv.restype = C.c_int
        L.orc_read_xcom_csv.argtypes = [C.c_char_p, P]
        for n in ("orc_bp3d20_weight", "orc_bp3d20_filter", "orc_fdk_std_weight", "orc_fdk_std_filter",
                  "orc_fdk_std_filter_T", "orc_fdk_std_weight_T"):
            getattr(L, n).argtypes = [C.POINTER(Geom), P, P]
        for n in ("orc_bp3d20_backproject", "orc_fdk_std_backproject", "orc_fdk_std_backproject_T",
                  "orc_fdk_std", "orc_fdk_std_adjoint"):
            getattr(L, n).argtypes = [C.POINTER(Geom), P, P, C.c_int]
        L.orc_primary_lineint.argtypes = [C.POINTER(Geom), C.POINTER(Phantom), C.POINTER(Tables),
                                          C.c_double, P, P, C.c_int]
        L.orc_siddon_lengths.argtypes = [C.POINTER(Phantom), P, P, P]
        L.orc_mc_run.argtypes = [C.POINTER(Geom), C.POINTER(Phantom), C.POINTER(Tables),
                                 C.POINTER(McCfg), P, P, C.POINTER(McStats), C.c_int]
        L.orc_kahn_lambda_ratio.restype = C.c_double
        L.orc_kahn_lambda_ratio.argtypes = [C.c_double, P, C.POINTER(C.c_int)]
        L.orc_thomson_mu.restype = C.c_double
        L.orc_thomson_mu.argtypes = [C.c_double]
        L.orc_rotate_dir.argtypes = [P, C.c_double, C.c_double]
        L.orc_brick_masks.argtypes = [C.POINTER(Phantom), C.c_int, P]
        L.orc_ray_masks.argtypes = [C.POINTER(Geom), C.POINTER(Phantom), P, P, P, C.c_int, C.c_int, C.c_int, P]
        L.orc_flight_box.argtypes = [C.POINTER(Phantom), C.POINTER(Tables), P, P]
        L.orc_spectrum_bin.restype = C.c_int
        L.orc_spectrum_bin.argtypes = [P, C.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


# ------------------------------------------------------------------ geometry
def geom(**kw) -> Geom:
    """Build an oracle Geom from keyword fields (same names as the C struct)."""
    g = Geom()
    defaults = dict(preset=PRESET_STANDARD, n_views=1, nu=1, nv=1, nx=1, ny=1, nz=1, sod=1.0, sdd=2.0,
                    du=1.0, dv=1.0, u_off=0.0, v_off=0.0, vx=1.0, vy=1.0, vz=1.0, ox=0.0, oy=0.0, oz=0.0,
                    angle0=0.0, dangle=None, odd=60.0, scale=1.0, step_deg=1, roi_x0=0, roi_x1=None,
                    mask_r=-1.0, filter=FILTER_RAMLAK)
    defaults.update(kw)
    if defaults["dangle"] is None:
        defaults["dangle"] = 2 * np.pi / defaults["n_views"]
    if defaults["roi_x1"] is None:
        defaults["roi_x1"] = defaults["nx"]
    for k, v in defaults.items():
        setattr(g, k, v)
    return g


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def u01(x):
    return lib().orc_u01(int(x))


def read_xcom_csv(path):
    tab = np.zeros((4, NROWS), dtype=np.float64)
    n = lib().orc_read_xcom_csv(path.encode(), _p(tab))
    if n < 0:
        raise FileNotFoundError(path)
    return tab, n


# ---------------------------------------------------------------- bp3d20 FDK
def bp3d20_weight(g, proj):
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    out = np.zeros(proj.shape, dtype=np.float64)
    lib().orc_bp3d20_weight(C.byref(g), _p(proj), _p(out))
    return out


def bp3d20_filter(g, pw):
    pw = np.ascontiguousarray(pw, dtype=np.float64)
    out = np.zeros((g.n_views, g.nv, g.nu), dtype=np.float64)
    lib().orc_bp3d20_filter(C.byref(g), _p(pw), _p(out))
    return out


def bp3d20_backproject(g, filt, nthreads=None):
    filt = np.ascontiguousarray(filt, dtype=np.float64)
    vol = np.zeros((g.nz, g.ny, g.nx), dtype=np.float64)
    lib().orc_bp3d20_backproject(C.byref(g), _p(filt), _p(vol), nthreads or nthreads_default())
    return vol


def bp3d20_fdk(g, proj, nthreads=None):
    f = bp3d20_filter(g, bp3d20_weight(g, proj))
    return bp3d20_backproject(g, f, nthreads), f


# -------------------------------------------------------------- standard FDK
def fdk_std_weight(g, proj):
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    out = np.zeros_like(proj)
    lib().orc_fdk_std_weight(C.byref(g), _p(proj), _p(out))
    return out


def fdk_std_filter(g, pw):
    pw = np.ascontiguousarray(pw, dtype=np.float64)
    out = np.zeros((g.n_views, g.nv, g.nu), dtype=np.float64)
    lib().orc_fdk_std_filter(C.byref(g), _p(pw), _p(out))
    return out


def fdk_std_filter_T(g, filt):
    filt = np.ascontiguousarray(filt, dtype=np.float64)
    out = np.zeros((g.n_views, g.nu, g.nv), dtype=np.float64)
    lib().orc_fdk_std_filter_T(C.byref(g), _p(filt), _p(out))
    return out


def fdk_std_backproject(g, filt, nthreads=None):
    filt = np.ascontiguousarray(filt, dtype=np.float64)
    vol = np.zeros((g.nz, g.ny, g.nx), dtype=np.float64)
    lib().orc_fdk_std_backproject(C.byref(g), _p(filt), _p(vol), nthreads or nthreads_default())
    return vol


def fdk_std_backproject_T(g, vol, nthreads=None):
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    out = np.zeros((g.n_views, g.nv, g.nu), dtype=np.float64)
    lib().orc_fdk_std_backproject_T(C.byref(g), _p(vol), _p(out), nthreads or nthreads_default())
    return out


def fdk_std(g, proj, nthreads=None):
    proj = np.ascontiguousarray(proj, dtype=np.float64)
    vol = np.zeros((g.nz, g.ny, g.nx), dtype=np.float64)
    lib().orc_fdk_std(C.byref(g), _p(proj), _p(vol), nthreads or nthreads_default())
    return vol


def fdk_std_adjoint(g, dvol, nthreads=None):
    dvol = np.ascontiguousarray(dvol, dtype=np.float64)
    out = np.zeros((g.n_views, g.nu, g.nv), dtype=np.float64)
    lib().orc_fdk_std_adjoint(C.byref(g), _p(dvol), _p(out), nthreads or nthreads_default())
    return out


def fdk(g, proj, nthreads=None):
    """Whole FDK in the preset named by g.preset (FP64)."""
    if g.preset == PRESET_BP3D20:
        return bp3d20_fdk(g, proj, nthreads)[0]
    return fdk_std(g, proj, nthreads)


# ------------------------------------------------------------ phantom / MC
def phantom(labels, vx, vy=None, vz=None, materials=((-1, 0.0), (0, 1.0)), origin=(0.0, 0.0, 0.0)):
    """materials: per label (table index or -1 for vacuum, density)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    ph = Phantom()
    ph.labels = labels.ctypes.data
    ph.nz, ph.ny, ph.nx = labels.shape
    ph.vx = vx
    ph.vy = vy if vy is not None else vx
    ph.vz = vz if vz is not None else vx
    ph.ox, ph.oy, ph.oz = origin
    ph.n_labels = len(materials)
    for i, (t, d) in enumerate(materials):
        ph.mat_table[i] = t
        ph.density[i] = d
    ph._keep = labels
    return ph


def tables(tabs):
    """tabs: array [n_tables][4][201] (coh, incoh, photo, total)."""
    arr = np.ascontiguousarray(tabs, dtype=np.float64)
    t = Tables()
    t.n_tables = arr.shape[0]
    t.tab = arr.ctypes.data
    t._keep = arr
    return t


def primary_lineint(g, ph, tb, e_kev=60.0, spectrum=None, nthreads=None):
    out = np.zeros((g.n_views, g.nu, g.nv), dtype=np.float64)
    sp = None
    if spectrum is not None:
        sp = np.ascontiguousarray(spectrum, dtype=np.float64)
    lib().orc_primary_lineint(C.byref(g), C.byref(ph), C.byref(tb), float(e_kev),
                              _p(sp) if sp is not None else None, _p(out), nthreads or nthreads_default())
    return out


def siddon_lengths(ph, p0, p1, n_labels=256):
    a = np.ascontiguousarray(p0, dtype=np.float64)
    b = np.ascontiguousarray(p1, dtype=np.float64)
    out = np.zeros(n_labels, dtype=np.float64)
    lib().orc_siddon_lengths(C.byref(ph), _p(a), _p(b), _p(out))
    return out


def mc_run(g, ph, tb, photons_per_view, e_kev=60.0, spectrum_cdf=None, max_interactions=5,
           coherent_mode=COH_FORWARD, seed=20261202, job=0, views=None, nthreads=None, ray_brick=0):
    cfg = McCfg()
    cfg.photons_per_view = int(photons_per_view)
    cfg.e_kev = float(e_kev)
    cdf = None
    if spectrum_cdf is not None:
        cdf = np.ascontiguousarray(spectrum_cdf, dtype=np.float64)
        cfg.spectrum_cdf = cdf.ctypes.data
    cfg.max_interactions = max_interactions
    cfg.coherent_mode = coherent_mode
    cfg.seed_lo = seed & 0xFFFFFFFF
    cfg.seed_hi = (seed >> 32) & 0xFFFFFFFF
    cfg.job = job
    v0, v1 = (0, g.n_views) if views is None else views
    cfg.view_begin, cfg.view_end = v0, v1
    cfg.ray_brick = int(ray_brick)
    prim = np.zeros((g.n_views, g.nu, g.nv), dtype=np.float64)
    scat = np.zeros_like(prim)
    st = McStats()
    lib().orc_mc_run(C.byref(g), C.byref(ph), C.byref(tb), C.byref(cfg), _p(prim), _p(scat), C.byref(st),
                     nthreads or nthreads_default())
    return prim, scat, st.as_dict()


def kahn_ratio(e_kev, uniforms):
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    n = C.c_int(0)
    r = lib().orc_kahn_lambda_ratio(float(e_kev), _p(u), C.byref(n))
    return r, n.value


def thomson_mu(u):
    return lib().orc_thomson_mu(float(u))


def rotate_dir(d, cos_theta, phi):
    """The oracle's direction update: d turned by polar cosine cos_theta and azimuth phi."""
    v = np.ascontiguousarray(d, dtype=np.float64).copy()
    lib().orc_rotate_dir(_p(v), float(cos_theta), float(phi))
    return v


def spectrum_bin(cdf, u):
    c = np.ascontiguousarray(cdf, dtype=np.float64)
    return lib().orc_spectrum_bin(_p(c), float(u))


def brick_masks(ph, b):
    """Label set (bit l: label l < 4, bit 4: label >= 4) of every brick of b^3
    voxels widened by one voxel: [ceil(nz/b)][ceil(ny/b)][ceil(nx/b)] uint8."""
    out = np.zeros((-(-ph.nz // b), -(-ph.ny // b), -(-ph.nx // b)), dtype=np.uint8)
    lib().orc_brick_masks(C.byref(ph), int(b), _p(out))
    return out


def flight_box(ph, tb):
    c0 = np.zeros(3)
    c1 = np.zeros(3)
    lib().orc_flight_box(C.byref(ph), C.byref(tb), _p(c0), _p(c1))
    return c0, c1


def ray_masks(g, ph, tb, b, views=None):
    """First-flight label sets per view and pixel (aim order iv * nu + iu):
    [n_views][nv * nu] uint8."""
    v0, v1 = (0, g.n_views) if views is None else views
    c0, c1 = flight_box(ph, tb)
    bricks = brick_masks(ph, b)
    out = np.zeros((v1 - v0, g.nu * g.nv), dtype=np.uint8)
    lib().orc_ray_masks(C.byref(g), C.byref(ph), _p(c0), _p(c1), _p(bricks), int(b), v0, v1 - v0, _p(out))
    return out
