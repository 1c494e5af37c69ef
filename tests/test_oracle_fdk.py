"""Pins for the FDK oracle (CPU only).

bp3d20 preset: against the reference's own programs (recon/bp3d20*.cpp compiled
unmodified and run on the same input, tests/golden/bp3d20_*.npz).
Standard preset: against closed forms and invariants -- the Ram-Lak kernel's
frequency response, a hand-written single-voxel backprojection, the
reconstruction of a uniform cylinder from its exact chords (mu in cm^-1), and
the transpose identity <A x, y> = <x, A^T y> for each stage and the whole
operator (A^T is separate scatter-form code).
"""
import os

import numpy as np
import pytest

from conftest import ROOT, rel_l2, rel_max


def _bp3d20_geom(oracle, H):
    if H == 65:
        return oracle.geom(preset=oracle.PRESET_BP3D20, n_views=360, nu=65, nv=65, nx=256, ny=256, nz=256, sod=160.0,
                           sdd=220.0, du=0.5, dv=0.5, vx=0.1, vy=0.1, vz=0.1, odd=60.0, scale=2.7, step_deg=1,
                           roi_x0=125, roi_x1=130, mask_r=118.0)
    return oracle.geom(preset=oracle.PRESET_BP3D20, n_views=360, nu=325, nv=325, nx=256, ny=256, nz=256, sod=160.0,
                       sdd=220.0, du=0.1, dv=0.1, vx=0.1, vy=0.1, vz=0.1, odd=60.0, scale=13.5, step_deg=1,
                       roi_x0=125, roi_x1=130, mask_r=-1.0)


@pytest.mark.parametrize("H", [65, 325])
def test_bp3d20_oracle_matches_reference_program(oracle, H):
    from synth import bp3d20_sphere_stack

    gold = np.load(os.path.join(ROOT, "tests", "golden", f"bp3d20_{H}.npz"))
    g = _bp3d20_geom(oracle, H)
    proj = bp3d20_sphere_stack(H, g.du)
    vol, filt = oracle.bp3d20_fdk(g, proj)
    # the reference accumulates in FP32; the FP64 restatement sits within a few 1e-6
    assert rel_max(vol[::4, ::4, 125:130], gold["vol_xy"]) < 5e-6
    assert rel_l2(vol[::4, ::4, 125:130], gold["vol_xy"]) < 2e-6
    assert rel_max(filt[list(gold["views"])][:, ::5, ::5], gold["filt"]) < 2e-6
    assert sorted(np.nonzero(np.any(vol != 0, axis=(0, 1)))[0].tolist()) == gold["nonzero_cols"].tolist()


def _std(oracle, **kw):
    base = dict(n_views=8, nu=24, nv=20, nx=9, ny=10, nz=11, sod=20.0, sdd=40.0, du=0.15, dv=0.17,
                vx=0.2, vy=0.21, vz=0.19, u_off=0.05, v_off=-0.03, ox=0.1, oy=-0.05, oz=0.07)
    base.update(kw)
    return oracle.geom(**base)


def test_ramlak_frequency_response(oracle):
    """The unit Ram-Lak kernel h[0]=1/4, h[odd]=-1/(pi n)^2 samples |w|/(2 pi) on
    [-pi, pi]: its DTFT at w=pi equals 1/2 and at w=0 tends to 0."""
    n = np.arange(-2001, 2002)
    h = np.where(n == 0, 0.25, np.where(n % 2 == 0, 0.0, -1.0 / (np.pi * np.maximum(np.abs(n), 1)) ** 2))
    tail = 1.0 / (np.pi ** 2 * 2001)  # sum over |n| > 2001 of the odd taps
    assert abs(np.sum(h * np.cos(np.pi * n)) - 0.5) < 1.01 * tail
    assert abs(np.sum(h)) < 1.01 * tail
    # the oracle's filter applies exactly these taps (scaled by 1/(2 tau))
    g = _std(oracle, n_views=1, nu=15, nv=1)
    x = np.zeros(g.n_views * g.nu * g.nv)
    x[7] = 1.0
    x = x.reshape(1, 15, 1)
    f = oracle.fdk_std_filter(g, x)[0, 0]
    tau = g.du * g.sod / g.sdd
    ref = np.array([0.25 if k == 0 else (0.0 if k % 2 == 0 else -1.0 / (np.pi * k) ** 2) for k in range(-7, 8)])
    assert np.allclose(f, ref / (2 * tau), rtol=1e-14, atol=1e-16)


def test_single_voxel_backprojection_by_hand(oracle):
    """One view, one voxel: bilinear sample at the voxel's projection x (sod/U)^2 x dangle."""
    g = _std(oracle, n_views=1, nx=1, ny=1, nz=1, ox=0.3, oy=-0.2, oz=0.1, angle0=0.6, dangle=0.25)
    rng = np.random.default_rng(0)
    filt = rng.standard_normal((1, g.nv, g.nu))
    vol = oracle.fdk_std_backproject(g, filt)
    X, Y, Z = g.ox, g.oy, g.oz
    b = g.angle0
    xr = X * np.cos(b) + Y * np.sin(b)
    yr = -X * np.sin(b) + Y * np.cos(b)
    U = g.sod + xr
    su = g.nu / 2 + (g.sdd * yr / U - g.u_off) / g.du - 0.5
    sv = g.nv / 2 + (g.sdd * Z / U - g.v_off) / g.dv - 0.5
    iu, iv = int(np.floor(su)), int(np.floor(sv))
    assert 0 <= iu < g.nu - 1 and 0 <= iv < g.nv - 1
    fu, fv = su - iu, sv - iv
    F = filt[0]
    val = ((1 - fv) * ((1 - fu) * F[iv, iu] + fu * F[iv, iu + 1]) + fv * ((1 - fu) * F[iv + 1, iu] + fu * F[iv + 1, iu + 1]))
    assert abs(vol[0, 0, 0] - val * (g.sod / U) ** 2 * g.dangle) < 1e-12


def _cylinder_chords(g, radius, mu):
    """Exact line integrals of an infinite uniform cylinder along z (radius, mu)."""
    out = np.zeros((g.n_views, g.nu, g.nv))
    iu = np.arange(g.nu)
    iv = np.arange(g.nv)
    yp = (iu + 0.5 - g.nu / 2) * g.du + g.u_off
    zp = (iv + 0.5 - g.nv / 2) * g.dv + g.v_off
    for k in range(g.n_views):
        # in-plane distance of the ray from the axis; rotation-invariant for a centred cylinder
        d = g.sod * np.abs(yp) / np.sqrt(g.sdd ** 2 + yp ** 2)
        chord2d = 2 * np.sqrt(np.maximum(radius ** 2 - d ** 2, 0.0))
        ray_len_scale = np.sqrt(g.sdd ** 2 + yp[:, None] ** 2 + zp[None, :] ** 2) / np.sqrt(g.sdd ** 2 + yp[:, None] ** 2)
        out[k] = mu * chord2d[:, None] * ray_len_scale
    return out


def test_uniform_cylinder_reconstructs_mu(oracle):
    """FDK of the exact chords of a uniform cylinder returns mu (cm^-1) inside it."""
    g = oracle.geom(n_views=180, nu=160, nv=16, nx=41, ny=41, nz=3, sod=50.0, sdd=100.0, du=0.1, dv=0.1,
                    vx=0.1, vy=0.1, vz=0.1)
    mu = 0.2
    proj = _cylinder_chords(g, 1.5, mu)
    vol = oracle.fdk_std(g, proj)
    yy, xx = np.mgrid[0:41, 0:41]
    r = np.hypot((xx - 20) * 0.1, (yy - 20) * 0.1)
    inner = vol[1][r < 1.1]
    outer = vol[1][(r > 1.9) & (r < 2.0)]
    assert abs(inner.mean() / mu - 1) < 0.01
    assert inner.std() / mu < 0.02
    assert np.abs(outer).max() < 0.05 * mu


@pytest.mark.parametrize("filt", [0, 1])
def test_adjoint_identity(oracle, filt):
    g = _std(oracle, filter=filt)
    rng = np.random.default_rng(1 + filt)
    x = rng.standard_normal((g.n_views, g.nu, g.nv))
    y = rng.standard_normal((g.nz, g.ny, g.nx))
    # whole operator
    lhs = np.sum(oracle.fdk_std(g, x) * y)
    rhs = np.sum(x * oracle.fdk_std_adjoint(g, y))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs) + 1)
    # stages
    f = rng.standard_normal((g.n_views, g.nv, g.nu))
    assert abs(np.sum(oracle.fdk_std_backproject(g, f) * y) - np.sum(f * oracle.fdk_std_backproject_T(g, y))) < 1e-10
    assert abs(np.sum(oracle.fdk_std_filter(g, x) * f) - np.sum(x * oracle.fdk_std_filter_T(g, f))) < 1e-10


def test_explicit_matrix_transpose(oracle):
    """The adjoint applied to unit vectors equals the transposed explicit matrix."""
    g = _std(oracle, n_views=3, nu=6, nv=5, nx=3, ny=3, nz=3)
    nin = g.n_views * g.nu * g.nv
    nout = g.nx * g.ny * g.nz
    A = np.zeros((nout, nin))
    for j in range(nin):
        e = np.zeros(nin)
        e[j] = 1.0
        A[:, j] = oracle.fdk_std(g, e.reshape(g.n_views, g.nu, g.nv)).ravel()
    AT = np.zeros((nin, nout))
    for i in range(nout):
        e = np.zeros(nout)
        e[i] = 1.0
        AT[:, i] = oracle.fdk_std_adjoint(g, e.reshape(g.nz, g.ny, g.nx)).ravel()
    assert np.abs(A).max() > 0
    assert np.allclose(A.T, AT, rtol=1e-12, atol=1e-14)
