"""Deterministic synthetic inputs shared by the golden-fixture scripts and the tests.

Everything here is plain numpy computed in float64 and rounded to float32 once,
so the same call yields the same arrays here and on any other host (up to the
last bit of numpy's sin/cos, far below every tolerance used with them).
"""
from __future__ import annotations

import numpy as np


def bp3d20_sphere_stack(H: int, pitch: float, n_views: int = 360, step_deg: int = 1,
                        sod: float = 160.0, sdd: float = 220.0,
                        centre=(2.0, 1.0, 0.5), radius: float = 5.0, mu: float = 0.2) -> np.ndarray:
    """Line integrals of a uniform sphere in the reference CBCT geometry.

    Layout [view][r][c] as read by recon/bp3d20_325.cpp:29-30: view k at k*step_deg
    degrees, r the in-plane detector index (y' = H/2*pitch - (r+0.5)*pitch),
    c the axial index (z' = H/2*pitch - (c+0.5)*pitch); source at (-sod,0,0)
    and detector plane x' = sdd - sod in the rotating frame.
    """
    half = 0.5 * H * pitch
    r = np.arange(H, dtype=np.float64)
    yp = half - (r + 0.5) * pitch
    zp = half - (r + 0.5) * pitch
    out = np.zeros((n_views, H, H), dtype=np.float32)
    c = np.asarray(centre, dtype=np.float64)
    for k in range(n_views):
        b = np.deg2rad(k * step_deg)
        cb, sb = np.cos(b), np.sin(b)
        src = np.array([-sod * cb, -sod * sb, 0.0])
        xd = sdd - sod
        # detector points in world coordinates, [r][c][3]
        px = xd * cb - yp * sb
        py = xd * sb + yp * cb
        P = np.empty((H, H, 3))
        P[:, :, 0] = px[:, None]
        P[:, :, 1] = py[:, None]
        P[:, :, 2] = zp[None, :]
        d = P - src
        L = np.linalg.norm(d, axis=-1)
        d /= L[..., None]
        m = src - c
        bq = np.einsum("ijk,k->ij", d, m)
        cq = m @ m - radius * radius
        disc = bq * bq - cq
        chord = np.where(disc > 0, 2.0 * np.sqrt(np.maximum(disc, 0.0)), 0.0)
        out[k] = (mu * chord).astype(np.float32)
    return out


def random_stack(shape, seed: int, scale: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (scale * rng.standard_normal(shape)).astype(np.float32)
