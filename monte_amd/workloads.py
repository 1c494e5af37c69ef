"""Measured workloads that span several entry points.

  batched_jobs()   many uneven independent MC jobs for one persistent launch
                   (64 cubic phantoms of side U[65, 325] voxels at 0.5 mm,
                   log-uniform 1e5..1e8 photons per job, mono 40..120 keV,
                   square detectors of side U[64, 256], 1..36 views)
  e2e_inputs()     the end-to-end scan: 512^3 phantom at 0.4 mm, 720 views of a
                   1024 x 768 detector, 120 kVp spectrum

The reference runs one job per program invocation (monte_cu/CBCT_real325.cu:
84-303: one phantom, one geometry, a fixed 360-view loop per pixel thread
:450-452); batching jobs into one launch has no counterpart there.
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry as G
from . import phantoms, tables
from .mc import McJobSpec, McSettings


def _cube_geometry(side: int, v: float, nd: int, views: int) -> G.Geometry:
    """Standard geometry (SOD 60 cm, SDD 120 cm) whose nd x nd detector covers
    the cube's shadow from every view: in-plane the cube reaches radius
    s / sqrt(2) about the axis, magnified at most SDD / (SOD - s / sqrt(2))."""
    sod, sdd = 60.0, 120.0
    s = side * v
    r = s / math.sqrt(2.0)
    mag = sdd / (sod - r)
    du = 2.0 * r * mag * 1.02 / nd
    dv = 2.0 * 0.5 * s * mag * 1.02 / nd
    return G.Geometry(n_views=views, nu=nd, nv=nd, nx=side, ny=side, nz=side, sod=sod, sdd=sdd, du=du, dv=dv,
                      vx=v, vy=v, vz=v, filter=1)


def batched_jobs(seed: int = 20261202, n_jobs: int = 64, v: float = 0.05, mc_seed: int = 20261202) -> list:
    """Jobs of the batched MC workload, drawn from `seed`; job i uses Philox
    stream id i, so each job draws the same histories alone or batched.
    Photons per job are a total over its views (photons per view = the total
    divided by the views, rounded up)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        side = int(rng.integers(65, 326))
        total = float(10.0 ** rng.uniform(5.0, 8.0))
        e_kev = float(rng.uniform(40.0, 120.0))
        nd = int(rng.integers(64, 257))
        views = int(rng.integers(1, 37))
        ph = phantoms.cube_phantom(side, v, seed=seed + i)
        geom = _cube_geometry(side, v, nd, views)
        ppv = max(1, int(math.ceil(total / views)))
        s = McSettings(photons_per_view=ppv, e_kev=e_kev, seed=mc_seed, job=i)
        jobs.append(McJobSpec(phantom=ph, geom=geom, settings=s))
    return jobs


def e2e_inputs(n_views: int = 720, seed: int = 20261202):
    """Geometry, phantom and spectrum of the end-to-end scan."""
    return G.e2e(n_views=n_views), phantoms.full_phantom(seed=seed, n=512, v=0.04), \
        tables.kramers_spectrum(120.0, 2.5)
