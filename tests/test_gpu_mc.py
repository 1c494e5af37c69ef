"""Monte Carlo transport and the exact primary estimator on the GPU.

* Primary-only projection (scatter off) vs exact Beer-Lambert line integrals
  from the FP64 Siddon oracle: 1e-5 relative.
* Full MC vs the FP64 CPU oracle at equal photon count.  Both draw every
  history from the same Philox stream in the same order, so almost all
  histories coincide; the test asserts per-pixel agreement within 3 sigma of
  the Poisson bound, reports the chi-square statistic and the fraction of
  identical pixels.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, oracle_geom, oracle_phantom

pytestmark = pytest.mark.gpu


def _tables(oracle):
    return oracle.tables(np.stack([oracle.read_xcom_csv(os.path.join(ROOT, "monte_amd", "data", f))[0]
                                   for f in ("xcom2.csv", "Ca.csv")]))


@pytest.mark.parametrize("mode", ["mono60", "mono30", "spectrum"])
def test_primary_exact_matches_siddon(ctx, oracle, mode):
    from monte_amd import geometry as G
    from monte_amd import phantoms, tables
    from monte_amd.mc import primary_exact

    geom = G.small(n_views=8)
    ph = phantoms.small_phantom()
    spec = tables.kramers_spectrum(120.0, 2.5) if mode == "spectrum" else None
    e = 30.0 if mode == "mono30" else 60.0
    got = primary_exact(ph, geom, e_kev=e, spectrum=spec, ctx=ctx).cpu().numpy()
    ref = oracle.primary_lineint(oracle_geom(oracle, geom), oracle_phantom(oracle, ph), _tables(oracle), e, spec)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    print(mode, "max rel", err.max(), "max L", ref.max())
    assert err.max() < 1e-5


def _mc_pair(ctx, oracle, geom, ph, photons, **kw):
    from monte_amd.mc import McSettings, mc_project

    s = McSettings(photons_per_view=photons, **kw)
    prim, scat, st = mc_project(ph, geom, s, ctx=ctx)
    spec_cdf = None
    if s.spectrum is not None:
        from monte_amd.tables import spectrum_cdf_f32

        spec_cdf = spectrum_cdf_f32(s.spectrum)
    p_ref, s_ref, st_ref = oracle.mc_run(oracle_geom(oracle, geom), oracle_phantom(oracle, ph), _tables(oracle),
                                         photons, e_kev=s.e_kev, spectrum_cdf=spec_cdf,
                                         max_interactions=s.max_interactions, coherent_mode=s.coherent,
                                         seed=s.seed, job=s.job, ray_brick=ctx.tuning["mc_ray_brick"])
    return prim.cpu().numpy(), scat.cpu().numpy(), st, p_ref, s_ref, st_ref


@pytest.mark.parametrize("coherent", [0, 1])
def test_mc_matches_oracle_small(ctx, oracle, coherent):
    from monte_amd import geometry as G
    from monte_amd import phantoms

    geom = G.small(n_views=4)
    ph = phantoms.small_phantom()
    N = 200_000
    prim, scat, st, p_ref, s_ref, st_ref = _mc_pair(ctx, oracle, geom, ph, N, e_kev=60.0, coherent=coherent)
    print("gpu", st)
    print("cpu", st_ref)
    assert st["histories"] == st_ref["histories"] == N * geom.n_views
    for a, b in ((prim, p_ref), (scat, s_ref)):
        assert a.sum() > 0
        same = np.mean(a == b)
        var = np.maximum(a + b, 1.0)
        z = (a - b) / np.sqrt(var)
        chi2 = float(np.sum((a - b) ** 2 / var))
        print(f"identical pixels {same:.4f}  max|z| {np.abs(z).max():.2f}  chi2/dof {chi2 / a.size:.4f}")
        assert np.abs(z).max() < 3.0
        assert same > 0.95
    # totals of the per-history statistics agree to the same few divergent histories
    for k in ("tallied_primary", "tallied_scatter", "photo", "compton", "coherent"):
        assert abs(st[k] - st_ref[k]) <= max(10, 1e-3 * st_ref[k]), k


@pytest.mark.parametrize("nu,nv,photons", [(13, 11, 100), (13, 11, 1000), (37, 21, 5000), (1, 1, 1), (1, 1, 50),
                                            (9, 1, 7), (1, 9, 200)])
@pytest.mark.parametrize("coherent", [0, 1])
def test_mc_partial_tiles_and_cycles(ctx, oracle, nu, nv, photons, coherent):
    """Work units are 8 x 8 detector tiles over photon cycles: detectors that end in partial
    tiles, views with fewer photons than pixels (only the tile rows holding photons count) and
    a partial last cycle must still run every photon once, as the oracle does.  The degenerate
    cases are one pixel with one photon or many cycles, and one-row or one-column detectors.
    Forward coherent scattering (coherent = 0) keeps the direction, so scattered photons reach
    even a one-pixel detector and the scatter images are compared too."""
    from monte_amd import geometry as G
    from monte_amd import phantoms

    geom = G.small(n_views=3).replace(nu=nu, nv=nv, du=0.8, dv=0.8)
    ph = phantoms.small_phantom()
    prim, scat, st, p_ref, s_ref, st_ref = _mc_pair(ctx, oracle, geom, ph, photons, e_kev=60.0, coherent=coherent)
    print(nu, nv, photons, coherent, "gpu", st, "cpu", st_ref)
    assert st["histories"] == st_ref["histories"] == photons * geom.n_views
    # every history ends once: tallied, missed, absorbed or capped
    assert st["tallied_primary"] + st["tallied_scatter"] + st["missed"] + st["photo"] + st["capped"] == \
        st["histories"]
    assert prim.sum() == st["tallied_primary"] and scat.sum() == st["tallied_scatter"]
    # the same photons aim at the same pixels with the same draws
    assert np.mean(prim == p_ref) > 0.99 and np.mean(scat == s_ref) > 0.95
    assert abs(prim.sum() - p_ref.sum()) <= max(3, 1e-3 * p_ref.sum())


@pytest.mark.parametrize("label", [1, 2])
@pytest.mark.parametrize("coherent", [0, 1])
def test_mc_one_voxel_phantom(ctx, oracle, label, coherent):
    """A phantom of one voxel: the flight box and the first-flight label sets reduce to that
    voxel, and the detector sees it at a few pixels only.  Forward coherent scattering keeps
    the photon's direction, so that mode also tallies scatter at those pixels."""
    from monte_amd import geometry as G
    from monte_amd.phantoms import MATERIALS, LabelPhantom

    ph = LabelPhantom(np.full((1, 1, 1), label, np.uint8), (0.4, 0.4, 0.4), list(MATERIALS))
    geom = G.small(n_views=3).replace(nu=16, nv=16, du=0.2, dv=0.2, nx=1, ny=1, nz=1, vx=0.4, vy=0.4, vz=0.4)
    prim, scat, st, p_ref, s_ref, st_ref = _mc_pair(ctx, oracle, geom, ph, 20_000, e_kev=30.0, coherent=coherent)
    print(label, coherent, "gpu", st, "cpu", st_ref)
    assert st["histories"] == st_ref["histories"] == 20_000 * geom.n_views
    assert st["photo"] + st["compton"] > 0
    if coherent == 0:
        assert st["tallied_scatter"] > 0
    assert prim.sum() == st["tallied_primary"] and scat.sum() == st["tallied_scatter"]
    assert np.mean(prim == p_ref) > 0.99 and np.mean(scat == s_ref) > 0.99
    for k in ("tallied_primary", "tallied_scatter", "photo", "compton", "coherent"):
        assert abs(st[k] - st_ref[k]) <= 3, k


def test_mc_primary_matches_beer_lambert(ctx, oracle):
    """Primary counts are Binomial(N0, exp(-L)) with L the exact line integral."""
    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.mc import McSettings, mc_project

    geom = G.small(n_views=2)
    ph = phantoms.small_phantom()
    npix = geom.nu * geom.nv
    N = 200 * npix
    prim, _, st = mc_project(ph, geom, McSettings(photons_per_view=N, e_kev=60.0), ctx=ctx)
    L = oracle.primary_lineint(oracle_geom(oracle, geom), oracle_phantom(oracle, ph), _tables(oracle), 60.0)
    p = np.exp(-L)
    n0 = 200
    mean = n0 * p
    sd = np.sqrt(np.maximum(n0 * p * (1 - p), 1e-12))
    z = (prim.cpu().numpy() - mean) / sd
    inside = p < 0.999
    chi2 = float(np.sum(z[inside] ** 2)) / inside.sum()
    print("chi2/dof", chi2, "max|z|", np.abs(z[inside]).max())
    assert 0.9 < chi2 < 1.1
    # photons through air only are never absorbed
    assert np.all(prim.cpu().numpy()[~inside] == n0)


def test_counts_to_lineint(ctx):
    import torch

    from monte_amd import geometry as G
    from monte_amd.mc import counts_to_lineint, photons_per_pixel

    geom = G.small(n_views=2)
    N = 61 * geom.nu * geom.nv + 1000  # uneven: the first 1000 pixels in aim order get 62 photons
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 70, size=geom.proj_shape).astype(np.float64)
    got = counts_to_lineint(torch.from_numpy(counts).cuda(), geom, N, ctx).cpu().numpy()
    n0 = np.full(geom.nu * geom.nv, 61.0)  # aim order: iv * nu + iu
    n0[:1000] = 62.0
    n0 = n0.reshape(geom.nv, geom.nu).T[None]
    assert np.array_equal(n0[0], photons_per_pixel(geom, N))
    ref = np.log(n0) - np.log(np.clip(counts, 1.0, n0))
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6)


def test_counts_to_lineint_flat(ctx):
    import ctypes as C

    import torch

    from monte_amd import _lib
    from monte_amd import geometry as G
    from monte_amd.mc import counts_to_lineint, counts_to_lineint_flat, photons_per_pixel

    geom = G.small(n_views=3)
    N = 61 * geom.nu * geom.nv + 1000
    rng = np.random.default_rng(8)
    counts = rng.integers(0, 70, size=geom.proj_shape).astype(np.float64)
    cd = torch.from_numpy(counts).cuda()
    # the nominal photons per pixel as the flat field gives the plain conversion, bitwise
    n0 = photons_per_pixel(geom, N)
    same = counts_to_lineint_flat(cd, geom, torch.from_numpy(n0).cuda(), ctx)
    assert torch.equal(same, counts_to_lineint(cd, geom, N, ctx))
    # one air scan for every view, and one per view, with dead pixels (F <= 0 -> 0)
    for flat in (rng.uniform(40.0, 80.0, size=(geom.nu, geom.nv)), rng.uniform(40.0, 80.0, size=geom.proj_shape)):
        flat[..., :3, :2] = 0.0
        flat[..., 5, 7] = -1.0
        got = counts_to_lineint_flat(cd, geom, torch.from_numpy(flat).cuda(), ctx).cpu().numpy()
        f = np.broadcast_to(flat, geom.proj_shape)
        with np.errstate(divide="ignore", invalid="ignore"):
            ref = np.where(f > 0, np.log(f) - np.log(np.maximum(np.minimum(counts, f), 1.0)), 0.0)
        assert np.allclose(got, ref.astype(np.float32), rtol=1e-6, atol=1e-6)
        assert np.all(got[..., :3, :2] == 0.0) and np.all(got[..., 5, 7] == 0.0)
    # flat_views must be 1 or the number of views
    L = _lib.lib()
    out = torch.empty(geom.proj_shape, dtype=torch.float32, device="cuda")
    fl = torch.ones((2, geom.nu, geom.nv), dtype=torch.float64, device="cuda")
    g = geom.c()
    rc = L.monte_counts_to_lineint_flat(ctx.handle, C.byref(g), C.c_void_p(cd.data_ptr()), C.c_void_p(fl.data_ptr()),
                                        2, C.c_void_p(out.data_ptr()))
    assert rc == _lib.ERR_INVALID


def _box_chords(geom, lo, hi):
    """Closed-form chord lengths of every source -> pixel-centre ray through the box
    [lo, hi] (cm), standard preset: source at -sod (cos b, sin b, 0), pixel centre at
    sdd along the central ray plus u' = (iu + 1/2 - nu/2) du + u_off across it and
    v' = (iv + 1/2 - nv/2) dv + v_off along z."""
    d = geom.as_dict()
    out = np.zeros(geom.proj_shape)
    yp = (np.arange(geom.nu) + 0.5 - 0.5 * geom.nu) * d["du"] + d["u_off"]
    zp = (np.arange(geom.nv) + 0.5 - 0.5 * geom.nv) * d["dv"] + d["v_off"]
    Y, Z = np.meshgrid(yp, zp, indexing="ij")
    for k in range(geom.n_views):
        b = d["angle0"] + k * d["dangle"]
        c, s = np.cos(b), np.sin(b)
        p0 = np.array([-d["sod"] * c, -d["sod"] * s, 0.0])
        ray = np.stack([d["sdd"] * c - Y * s, d["sdd"] * s + Y * c, Z], axis=-1)
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (lo - p0) / ray
            tb = (hi - p0) / ray
        t0 = np.nanmax(np.minimum(ta, tb), axis=-1)
        t1 = np.nanmin(np.maximum(ta, tb), axis=-1)
        out[k] = np.clip(t1 - t0, 0.0, None) * np.linalg.norm(ray, axis=-1)
    return out


def test_primary_exact_box_closed_form(ctx):
    """Exact primary through an off-centre water box whose faces lie on voxel faces:
    mu x the closed-form chord, pinned without the oracle."""
    from monte_amd import geometry as G
    from monte_amd import tables
    from monte_amd.mc import primary_exact
    from monte_amd.phantoms import LabelPhantom

    geom = G.small(n_views=7).replace(u_off=0.13, v_off=-0.21, angle0=0.3)
    v = (0.1, 0.08, 0.12)
    n = (40, 50, 30)  # nz, ny, nx
    origin = (0.25, -0.4, 0.3)
    labels = np.zeros(n, np.uint8)
    sl = (slice(6, 25), slice(10, 41), slice(4, 27))  # z, y, x voxel ranges
    labels[sl] = 1
    ph = LabelPhantom(labels, v, [(-1, 0.0), (tables.WATER, 1.0)], origin)
    b0 = [origin[a] - 0.5 * n[2 - a] * v[a] for a in range(3)]  # x, y, z corner of the grid
    lo = np.array([b0[0] + sl[2].start * v[0], b0[1] + sl[1].start * v[1], b0[2] + sl[0].start * v[2]])
    hi = np.array([b0[0] + sl[2].stop * v[0], b0[1] + sl[1].stop * v[1], b0[2] + sl[0].stop * v[2]])
    mu = np.float32(tables.linear_mu(ctx.tables, tables.WATER, 1.0, 60.0))
    ref = float(mu) * _box_chords(geom, lo, hi)
    got = primary_exact(ph, geom, e_kev=60.0, ctx=ctx).double().cpu().numpy()
    assert (ref > 0).mean() > 0.1
    err = np.abs(got - ref) / np.maximum(ref, 1e-3 * ref.max())
    print("box closed form: max rel", err.max(), "max L", ref.max())
    assert err.max() < 1e-5


def test_woodcock_majorant_invariance(ctx):
    """A label that no voxel carries, made dense, raises the global Woodcock majorant
    (more virtual steps per history) without changing the physics: both runs
    sample the same distributions, so their tallies agree pixel by pixel within
    Poisson noise.  (Run with the global majorant for every flight: the
    first-flight label sets would leave the unused label out.)"""
    import torch

    from monte_amd import geometry as G
    from monte_amd import phantoms, tables
    from monte_amd.mc import McSettings, mc_project

    geom = G.small(n_views=4)
    base = phantoms.small_phantom()
    heavy = phantoms.LabelPhantom(base.labels, base.voxel, list(base.materials) + [(tables.CALCIUM, 5.0)],
                                  base.origin)
    assert 3 not in np.unique(base.labels)
    s = McSettings(photons_per_view=4_000_000, e_kev=60.0)
    brick = ctx.tuning["mc_ray_brick"]
    ctx.set_tuning(mc_ray_brick=0)
    try:
        pa, sa, sta = mc_project(base, geom, s, ctx=ctx)
        pb, sb, stb = mc_project(heavy, geom, s, ctx=ctx)
    finally:
        ctx.set_tuning(mc_ray_brick=brick)
    print("steps per history", sta["steps"] / sta["histories"], "->", stb["steps"] / stb["histories"])
    assert stb["steps"] > 1.5 * sta["steps"]
    print("tallied scatter", sta["tallied_scatter"], stb["tallied_scatter"])

    def bins(t, k):  # k x k pixel blocks, so the sparse scatter image has counts to compare
        x = t.cpu().numpy()
        return x.reshape(x.shape[0], x.shape[1] // k, k, x.shape[2] // k, k).sum(axis=(2, 4))

    for name, a, b in (("primary", bins(pa, 8), bins(pb, 8)), ("scatter", bins(sa, 16), bins(sb, 16))):
        keep = a + b >= 20
        chi2 = float(np.mean((a - b)[keep] ** 2 / (a + b)[keep]))
        dof = int(keep.sum())
        tot = (a.sum() - b.sum()) / np.sqrt(a.sum() + b.sum())
        print(f"{name}: bins {dof}  chi2/dof {chi2:.4f}  total difference {tot:.2f} sigma")
        assert dof > 100
        tol = 5 * np.sqrt(2.0 / dof)
        if name == "scatter":  # one count per history at most: Poisson bins, chi2/dof near 1
            assert abs(chi2 - 1.0) < tol
        else:  # binomial bins: a + b over-estimates their variance, so only an excess counts
            assert chi2 < 1.0 + tol
        assert abs(tot) < 5.0
    for k in ("tallied_primary", "tallied_scatter", "photo", "compton", "coherent"):
        assert abs(sta[k] - stb[k]) < 6 * np.sqrt(max(sta[k] + stb[k], 1)), k
