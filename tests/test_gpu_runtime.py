"""Runtime features on the GPU: hipGraph replay, autotuning invariance, the
persistent multi-job MC launch, host-pointer mode and error handling."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_graph_replay_bitwise_equals_eager(ctx):
    import torch

    from monte_amd import geometry as G
    from monte_amd.fdk import fdk, fdk_graph

    for geom in (G.small(n_views=24), G.bp3d20_65(n_views=30)):
        rng = np.random.default_rng(1)
        proj = _t(rng.standard_normal((2, *geom.proj_shape)).astype(np.float32))
        eager = fdk(proj, geom, ctx)
        out = torch.full_like(eager, 7.0)
        for rep in range(3):
            fdk_graph(proj, geom, out, ctx)
            torch.cuda.synchronize()
            assert torch.equal(out, eager), rep
            # unrelated eager work between replays, then overwrite the output
            fdk(proj[:1].contiguous(), geom, ctx)
            out.fill_(-1.0)
            torch.cuda.synchronize()
        # new input data, same buffers: the replay sees it
        proj.mul_(0.5)
        fdk_graph(proj, geom, out, ctx)
        torch.cuda.synchronize()
        assert torch.equal(out, fdk(proj, geom, ctx))


def test_backprojection_invariant_under_tuning(ctx):
    import torch

    from monte_amd import geometry as G
    from monte_amd.fdk import fdk

    geom = G.small(n_views=16)
    rng = np.random.default_rng(2)
    proj = _t(rng.standard_normal(geom.proj_shape).astype(np.float32))
    saved = ctx.tuning
    try:
        ref = fdk(proj, geom, ctx)
        for zc in (4, 8, 16):
            for tx in (8, 16, 32, 64):
                ctx.set_tuning(bp_zchunk=zc, bp_tile_x=tx)
                assert torch.equal(fdk(proj, geom, ctx), ref), (zc, tx)
    finally:
        ctx.set_tuning(**saved)


def test_mc_invariant_under_tuning(ctx):
    import torch

    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.mc import McSettings, mc_project

    geom = G.small(n_views=3)
    ph = phantoms.small_phantom()
    s = McSettings(photons_per_view=150_000, e_kev=50.0)
    saved = ctx.tuning
    try:
        ref_p, ref_s, ref_st = mc_project(ph, geom, s, ctx=ctx)
        # 128^2 pixels and 150000 photons per view make 10 photon cycles: units of 4, 8, 3 and
        # 1 cycles of one tile, or (32 per unit) 10 cycles of 3 tiles, the last group partial
        for blk, bt, bpc in ((64, 4, 2), (128, 32, 8), (256, 8, 16), (256, 1, 1), (256, 3, 6)):
            ctx.set_tuning(mc_block=blk, mc_batch=bt, mc_blocks_per_cu=bpc)
            p, sc, st = mc_project(ph, geom, s, ctx=ctx)
            assert torch.equal(p, ref_p) and torch.equal(sc, ref_s), (blk, bt, bpc)
            assert st["histories"] == 3 * 150_000
            # the per-wave statistics, misses and interactions derived, do not depend on the launch
            assert {k: v for k, v in st.items() if k != "kernel_ms"} == \
                {k: v for k, v in ref_st.items() if k != "kernel_ms"}, (blk, bt, bpc)
    finally:
        ctx.set_tuning(**saved)


def test_autotune_records_choice(ctx):
    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.mc import McSettings, autotune

    saved = ctx.tuning
    try:
        t = autotune(phantoms.small_phantom(), G.small(n_views=4), McSettings(photons_per_view=100_000), what=7,
                     ctx=ctx)
        assert t == ctx.tuning
        assert t["bp_zchunk"] in (4, 8, 16) and t["mc_block"] in (64, 128, 256) and t["pj_zchunk"] in (0, 2, 4, 8)
    finally:
        ctx.set_tuning(**saved)


def test_batched_jobs_equal_single_runs(ctx):
    """64 uneven jobs in one persistent launch give exactly the tallies of the
    same jobs run one by one (RNG streams are keyed by job, view and photon)."""
    import torch

    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.mc import McJobSpec, McSettings, mc_batch, mc_project

    rng = np.random.default_rng(4)
    jobs = []
    for j in range(64):
        side = int(rng.integers(9, 40))
        v = 0.1
        det = int(rng.integers(8, 48))
        views = int(rng.integers(1, 5))
        ph = phantoms.cube_phantom(side, v, seed=j)
        half = side * v * 0.75
        du = 2.2 * half * 2.0 / det
        geom = G.Geometry(n_views=views, nu=det, nv=det, nx=side, ny=side, nz=side, sod=20.0, sdd=40.0, du=du, dv=du,
                          vx=v, vy=v, vz=v)
        photons = int(np.exp(rng.uniform(np.log(2e3), np.log(3e4))))
        s = McSettings(photons_per_view=photons, e_kev=float(rng.uniform(40, 120)), job=j, seed=77)
        jobs.append(McJobSpec(ph, geom, s))
    stats = mc_batch(jobs, ctx=ctx)
    for j, spec in enumerate(jobs):
        p, sc, st = mc_project(spec.phantom, spec.geom, spec.settings, ctx=ctx)
        assert torch.equal(p, spec.primary) and torch.equal(sc, spec.scatter), j
        assert stats[j]["histories"] == st["histories"] == spec.settings.photons_per_view * spec.geom.n_views
        for k in ("steps", "interactions", "tallied_primary", "tallied_scatter"):
            assert stats[j][k] == st[k], (j, k)


def test_host_pointer_mode_matches_device(ctx):
    import torch

    from monte_amd import _lib
    from monte_amd import geometry as G

    geom = G.small(n_views=6)
    rng = np.random.default_rng(6)
    proj = rng.standard_normal(geom.proj_shape).astype(np.float32)
    from monte_amd.fdk import fdk

    dev = fdk(_t(proj), geom, ctx).cpu().numpy()
    torch.cuda.synchronize()
    L = _lib.lib()
    out = np.zeros(geom.vol_shape, dtype=np.float32)
    g = geom.c()
    assert L.monte_set_pointer_mode(ctx.handle, 1) == 0
    try:
        rc = L.monte_fdk(ctx.handle, C.byref(g), proj.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), 1)
        assert rc == 0, L.monte_last_error(ctx.handle)
    finally:
        L.monte_set_pointer_mode(ctx.handle, 0)
    assert np.array_equal(out, dev)


def test_invalid_arguments_are_reported(ctx):
    import torch

    from monte_amd import _lib
    from monte_amd import geometry as G
    from monte_amd.fdk import fdk
    from monte_amd.mc import McSettings, mc_project
    from monte_amd import phantoms

    bad = G.small(n_views=4).replace(sdd=10.0, sod=50.0)
    with pytest.raises(_lib.MonteError) as e:
        fdk(_t(np.zeros(bad.proj_shape, np.float32)), bad, ctx)
    assert e.value.code == _lib.ERR_INVALID
    with pytest.raises(_lib.MonteError):
        mc_project(phantoms.small_phantom(), G.small(n_views=1), McSettings(photons_per_view=10, e_kev=500.0), ctx=ctx)
    g = G.small(n_views=1)
    with pytest.raises(_lib.MonteError):
        mc_project(phantoms.small_phantom(), g, McSettings(photons_per_view=10, views=(0, 5)), ctx=ctx)
    # the context keeps working after errors
    out = fdk(_t(np.ones(g.proj_shape, np.float32)), g, ctx)
    torch.cuda.synchronize()
    assert torch.isfinite(out).all()
