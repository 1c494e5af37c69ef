"""Scan geometries: the reference's own and the workloads this engine is measured on.

Units are cm.  A Geometry is the Python face of monte_geom (include/monte.h);
`.c()` returns the ctypes struct.

Reference geometries (preset BP3D20, reproducing recon/bp3d20*.cpp):
  bp3d20_325()  SOD 160, SDD 220, 325x325 detector at 0.1 (bp3d20_325.cpp:17-20,
                CBCT_real325.cu:447), 360 views at 1 degree, 256^3 volume at 0.1
                with only x columns 125..129 reconstructed (:104), scale 2.7*5 (:170)
  bp3d20_65()   the 65x65 detector at 0.5 (bp3d20.cpp:17, :40), scale 2.7 and the
                sphere mask of radius 118 voxels (:145)
Workloads (preset STANDARD):
  small()       65^3 at 1 mm, SOD 500 mm / SDD 1000 mm, 128^2 detector at 1 mm, 36 views
  full()        325^3 at 0.5 mm, 512^2 detector at 1.2 mm, SOD 600 / SDD 1200 mm, 360 views,
                Shepp-Logan; the pitch makes the magnified field cover the whole cube
  e2e()         512^3 at 0.4 mm, 1024 x 768 detector at 0.8 mm, 720 views
  fd_check()    33^3 / 64^2 / 30 views for finite-difference checks of the adjoint
"""
from __future__ import annotations

import dataclasses
import math

from . import _lib


@dataclasses.dataclass
class Geometry:
    preset: int = _lib.PRESET_STANDARD
    n_views: int = 1
    nu: int = 1
    nv: int = 1
    nx: int = 1
    ny: int = 1
    nz: int = 1
    sod: float = 1.0
    sdd: float = 2.0
    du: float = 1.0
    dv: float = 1.0
    u_off: float = 0.0
    v_off: float = 0.0
    vx: float = 1.0
    vy: float = 1.0
    vz: float = 1.0
    ox: float = 0.0
    oy: float = 0.0
    oz: float = 0.0
    angle0: float = 0.0
    dangle: float | None = None
    odd: float = 60.0
    scale: float = 1.0
    step_deg: int = 1
    roi_x0: int = 0
    roi_x1: int | None = None
    mask_r: float = -1.0
    filter: int = _lib.FILTER_RAMLAK

    def c(self) -> _lib.Geom:
        g = _lib.Geom()
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name == "dangle" and v is None:
                v = 2.0 * math.pi / self.n_views
            if f.name == "roi_x1" and v is None:
                v = self.nx
            setattr(g, f.name, v)
        return g

    def as_dict(self) -> dict:
        d = dataclasses.asdict(self)
        if d["dangle"] is None:
            d["dangle"] = 2.0 * math.pi / self.n_views
        if d["roi_x1"] is None:
            d["roi_x1"] = self.nx
        return d

    @property
    def proj_shape(self):
        return (self.n_views, self.nu, self.nv)

    @property
    def filt_shape(self):
        return (self.n_views, self.nv, self.nu)

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    def replace(self, **kw) -> "Geometry":
        return dataclasses.replace(self, **kw)


def bp3d20_325(n_views: int = 360) -> Geometry:
    return Geometry(preset=_lib.PRESET_BP3D20, n_views=n_views, nu=325, nv=325, nx=256, ny=256, nz=256,
                    sod=160.0, sdd=220.0, du=0.1, dv=0.1, vx=0.1, vy=0.1, vz=0.1, odd=60.0, scale=2.7 * 5,
                    step_deg=1, roi_x0=125, roi_x1=130, mask_r=-1.0)


def bp3d20_65(n_views: int = 360) -> Geometry:
    return Geometry(preset=_lib.PRESET_BP3D20, n_views=n_views, nu=65, nv=65, nx=256, ny=256, nz=256,
                    sod=160.0, sdd=220.0, du=0.5, dv=0.5, vx=0.1, vy=0.1, vz=0.1, odd=60.0, scale=2.7,
                    step_deg=1, roi_x0=125, roi_x1=130, mask_r=118.0)


def small(n_views: int = 36) -> Geometry:
    return Geometry(n_views=n_views, nu=128, nv=128, nx=65, ny=65, nz=65, sod=50.0, sdd=100.0, du=0.1, dv=0.1,
                    vx=0.1, vy=0.1, vz=0.1, filter=_lib.FILTER_RAMLAK)


def full(n_views: int = 360) -> Geometry:
    return Geometry(n_views=n_views, nu=512, nv=512, nx=325, ny=325, nz=325, sod=60.0, sdd=120.0, du=0.12,
                    dv=0.12, vx=0.05, vy=0.05, vz=0.05, filter=_lib.FILTER_SHEPP_LOGAN)


def e2e(n_views: int = 720) -> Geometry:
    return Geometry(n_views=n_views, nu=1024, nv=768, nx=512, ny=512, nz=512, sod=60.0, sdd=120.0, du=0.08,
                    dv=0.08, vx=0.04, vy=0.04, vz=0.04, filter=_lib.FILTER_SHEPP_LOGAN)


def fd_check() -> Geometry:
    return Geometry(n_views=30, nu=64, nv=64, nx=33, ny=33, nz=33, sod=30.0, sdd=60.0, du=0.1, dv=0.1,
                    vx=0.1, vy=0.1, vz=0.1, filter=_lib.FILTER_RAMLAK)
