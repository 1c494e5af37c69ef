"""Voxel label phantoms ("phantom in"), painted by the library's monte_phantom_paint.

The reference generates its single water sphere with a triple loop
(monte_cpp/make_image01.cpp:8-33) and uses analytic rods inside the transport
kernel (monte_cu/CBCT_real325.cu:916-928) or a fixed label map
(CBCT_real325im.cu:921).  Here every phantom is an explicit uint8 label map
[z][y][x] plus a material list (table index, density) per label.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .tables import CALCIUM, WATER

CA_DENSITY = 1.55  # CBCT_real325.cu:123
WATER_DENSITY = 1.0


@dataclasses.dataclass
class LabelPhantom:
    labels: np.ndarray          # uint8 [nz][ny][nx]
    voxel: tuple                # (vx, vy, vz) cm
    materials: list             # per label: (table index or -1, density)
    origin: tuple = (0.0, 0.0, 0.0)

    def c(self, labels_ptr: int) -> _lib.Phantom:
        p = _lib.Phantom()
        p.labels = labels_ptr
        p.nz, p.ny, p.nx = self.labels.shape
        p.vx, p.vy, p.vz = self.voxel
        p.ox, p.oy, p.oz = self.origin
        p.n_labels = len(self.materials)
        for i, (t, d) in enumerate(self.materials):
            p.mat_table[i] = t
            p.density[i] = d
        return p


def paint(shape, voxel, prims) -> np.ndarray:
    """prims: iterable of (shape, label, (cx,cy,cz), (rx,ry,rz)) in cm; painted in order."""
    nz, ny, nx = shape
    labels = np.zeros(shape, dtype=np.uint8)
    arr = (_lib.Primitive * max(1, len(prims)))()
    for i, (s, lab, c, r) in enumerate(prims):
        arr[i].shape, arr[i].label = s, lab
        arr[i].cx, arr[i].cy, arr[i].cz = c
        arr[i].rx, arr[i].ry, arr[i].rz = r
    rc = _lib.lib().monte_phantom_paint(labels.ctypes.data_as(C.c_void_p), nx, ny, nz, voxel[0], voxel[1], voxel[2],
                                        arr, len(prims))
    _lib.check(rc, what="monte_phantom_paint")
    return labels


MATERIALS = [(-1, 0.0), (WATER, WATER_DENSITY), (CALCIUM, CA_DENSITY)]


def small_phantom() -> LabelPhantom:
    """65^3 at 1 mm: water cylinder r = 25 voxels along z, two Ca cylinders r = 4 at x = +-12 voxels."""
    v = 0.1
    n = 65
    h = 0.5 * n * v + 1.0
    prims = [
        (_lib.SHAPE_CYLINDER_Z, 1, (0.0, 0.0, 0.0), (25 * v, 25 * v, h)),
        (_lib.SHAPE_CYLINDER_Z, 2, (12 * v, 0.0, 0.0), (4 * v, 4 * v, h)),
        (_lib.SHAPE_CYLINDER_Z, 2, (-12 * v, 0.0, 0.0), (4 * v, 4 * v, h)),
    ]
    return LabelPhantom(paint((n, n, n), (v, v, v), prims), (v, v, v), list(MATERIALS))


def full_phantom(seed: int = 20261202, n: int = 325, v: float = 0.05) -> LabelPhantom:
    """325^3 at 0.5 mm: water ellipsoid with semi-axes 150 x 125 x 150 voxels and 12
    calcium spheres of radius U[3,15] voxels at uniform positions inside it
    (all lengths scale with n / 325 for other grid sizes)."""
    rng = np.random.default_rng(seed)
    f = n / 325.0
    a = np.array([150.0, 125.0, 150.0]) * f * v
    prims = [(_lib.SHAPE_ELLIPSOID, 1, (0.0, 0.0, 0.0), tuple(a))]
    placed = 0
    while placed < 12:
        r = rng.uniform(3.0, 15.0) * f * v
        c = rng.uniform(-1.0, 1.0, 3) * (a - r)
        if np.sum((c / (a - r)) ** 2) <= 1.0:
            prims.append((_lib.SHAPE_ELLIPSOID, 2, tuple(c), (r, r, r)))
            placed += 1
    return LabelPhantom(paint((n, n, n), (v, v, v), prims), (v, v, v), list(MATERIALS))


def cube_phantom(side: int, v: float, seed: int) -> LabelPhantom:
    """Water ellipsoid filling 90% of a cube of `side` voxels with one Ca insert."""
    rng = np.random.default_rng(seed)
    half = 0.45 * side * v
    a = np.array([half, half * rng.uniform(0.8, 1.0), half])
    r = 0.15 * half
    c = rng.uniform(-0.4, 0.4, 3) * a
    prims = [(_lib.SHAPE_ELLIPSOID, 1, (0.0, 0.0, 0.0), tuple(a)),
             (_lib.SHAPE_ELLIPSOID, 2, tuple(c), (r, r, r))]
    return LabelPhantom(paint((side, side, side), (v, v, v), prims), (v, v, v), list(MATERIALS))


def reference_rods_phantom(v: float = 0.1) -> LabelPhantom:
    """Label-map form of the reference's analytic phantom (monte_cu/CBCT_real325.cu:
    916-928): a water cylinder of radius 10 cm along x with |x| <= 10, holding
    eight calcium rods (density 1.55) of radius 1.5 cm parallel to x, their axes
    at radius 5 cm in the y-z plane every 45 degrees.  Voxel centres decide
    membership; at 0.1 cm the grid is 200^3 over [-10, 10]^3."""
    n = int(round(20.0 / v))
    prims = [(_lib.SHAPE_CYLINDER_X, 1, (0.0, 0.0, 0.0), (10.0, 10.0, 10.0))]
    for k in range(8):
        a = k * np.pi / 4.0
        prims.append((_lib.SHAPE_CYLINDER_X, 2, (0.0, 5.0 * np.sin(a), 5.0 * np.cos(a)), (10.0, 1.5, 1.5)))
    return LabelPhantom(paint((n, n, n), (v, v, v), prims), (v, v, v), list(MATERIALS))
