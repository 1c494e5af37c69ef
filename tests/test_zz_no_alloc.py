"""monte_mc_project_noalloc: the MC projection without device allocation.

Called through the C ABI and checked against monte_mc_project on the same
inputs: every tally adds 1.0 along the same Philox streams, so the two must be
bitwise equal.  The allocation-free variant gives every flight the global
Woodcock majorant (the first-flight label sets need device memory in
proportion to views x pixels), so monte_mc_project runs with the same setting
(tuning mc_ray_brick = 0) for the comparison.  Two calls in a row, the second with a smaller phantom,
detector, view count and photon count, so the reused scratch array must be
rewritten and its queue and counters cleared between calls.  Requests that do
not fit the scratch array, and host pointers, are refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, ph, geom, settings, labels, primary, scatter):
    from monte_amd import _lib

    g = geom.c()
    p = ph.c(labels.data_ptr())
    cfg, keep = settings.c()
    st = _lib.McStats()
    ctx.call("monte_mc_project_noalloc", C.byref(g), C.byref(p), C.byref(cfg), C.c_void_p(primary.data_ptr()),
             C.c_void_p(scatter.data_ptr()), C.byref(st))
    del keep
    return st.as_dict()


def test_noalloc_matches_mc_project_twice(ctx):
    import torch

    from monte_amd import geometry as G
    from monte_amd import phantoms, tables
    from monte_amd.mc import McSettings, device_labels, mc_project

    cases = [
        (phantoms.small_phantom(), G.small(n_views=12),
         McSettings(photons_per_view=400_000, spectrum=tables.kramers_spectrum(120.0, 2.5), coherent=1, seed=7)),
        (phantoms.cube_phantom(33, 0.1, seed=3), G.small(n_views=3).replace(nu=40, nv=24, du=0.2, dv=0.2),
         McSettings(photons_per_view=50_000, e_kev=45.0, coherent=0, seed=8, job=5)),
    ]
    for i, (ph, geom, s) in enumerate(cases):
        print(f"case {i}: phantom {ph.labels.shape} detector {geom.nu}x{geom.nv} views {geom.n_views}", flush=True)
        labels = device_labels(ph, "cuda")
        brick = ctx.tuning["mc_ray_brick"]
        ctx.set_tuning(mc_ray_brick=0)
        try:
            pa, sa, sta = mc_project(ph, geom, s, labels_dev=labels, ctx=ctx)
        finally:
            ctx.set_tuning(mc_ray_brick=brick)
        pb = torch.full(geom.proj_shape, 0.0, dtype=torch.float64, device="cuda")
        sb = torch.zeros_like(pb)
        stb = _noalloc(ctx, ph, geom, s, labels, pb, sb)
        torch.cuda.synchronize()
        print("mc_project", sta)
        print("noalloc   ", stb)
        assert sta["histories"] == s.photons_per_view * geom.n_views
        for k in sta:
            if k != "kernel_ms":
                assert sta[k] == stb[k], k
        assert torch.equal(pa, pb) and torch.equal(sa, sb)
        assert float(pa.sum()) > 0 and float(sa.sum()) > 0


def test_noalloc_refuses_what_does_not_fit(ctx):
    import torch

    from monte_amd import _lib
    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.mc import McSettings, device_labels

    ph = phantoms.cube_phantom(17, 0.1, seed=1)
    labels = device_labels(ph, "cuda")
    geom = G.small(n_views=20000).replace(nu=4, nv=4)
    t = torch.zeros(geom.proj_shape, dtype=torch.float64, device="cuda")
    g = geom.c()
    p = ph.c(labels.data_ptr())
    cfg, _keep = McSettings(photons_per_view=64).c()
    L = _lib.lib()
    ctx.sync_stream()
    rc = L.monte_mc_project_noalloc(ctx.handle, C.byref(g), C.byref(p), C.byref(cfg), C.c_void_p(t.data_ptr()),
                                    C.c_void_p(t.data_ptr()), None)
    assert rc == _lib.ERR_INVALID
    assert b"scratch" in L.monte_last_error(ctx.handle)
    assert not t.any()
    # host pointers would need staging buffers: refused
    assert L.monte_set_pointer_mode(ctx.handle, 1) == 0
    try:
        h = np.zeros(geom.proj_shape, np.float64)
        rc = L.monte_mc_project_noalloc(ctx.handle, C.byref(g), C.byref(p), C.byref(cfg),
                                        h.ctypes.data_as(C.c_void_p), h.ctypes.data_as(C.c_void_p), None)
        assert rc == _lib.ERR_UNSUPPORTED
    finally:
        L.monte_set_pointer_mode(ctx.handle, 0)
