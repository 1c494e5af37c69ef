"""The end-to-end differentiable scan (monte_e2e_step): MC with scatter ->
line integrals -> FDK -> L2 loss against the primary-only FDK -> gradient to
the line integrals through the FDK adjoint, eager and hipGraph-captured.

* Small scan: the one call equals, bitwise, the same chain run through the
  separate entry points (mc_project, counts_to_lineint of primary + scatter,
  fdk, the residual, fdk_adjoint); the loss equals 1/2 ||residual||^2 in FP64.
* The end-to-end workload at its full size (512^3 phantom at 0.4 mm, 720 views
  of a 1024 x 768 detector, 120 kVp, 5e7 photons per view): the captured graph,
  replayed with other GPU work and synchronisations in between, equals the
  eager step bitwise, every replay; <fdk(lineint), residual> =
  <lineint, fdk^T(residual)> to 1e-5 relative; every history ends exactly once
  and the tallies sum to the counters.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("primary", "scatter", "lineint", "vol", "residual", "grad", "loss")


def _snapshot(E):
    return {k: getattr(E, k).clone() for k in FIELDS}


def _assert_same(E, ref, what):
    import torch

    for k in FIELDS:
        assert torch.equal(getattr(E, k), ref[k]), f"{what}: {k} differs"


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def _conservation(E, st):
    assert st["histories"] == E.settings.photons_per_view * E.geom.n_views
    assert st["histories"] == st["tallied_primary"] + st["tallied_scatter"] + st["missed"] + st["photo"] + \
        st["capped"]
    assert float(E.primary.sum()) == st["tallied_primary"] and float(E.scatter.sum()) == st["tallied_scatter"]


def test_e2e_small_equals_separate_calls(ctx):
    import torch

    from monte_amd import geometry as G
    from monte_amd import phantoms, tables
    from monte_amd.e2e import EndToEnd
    from monte_amd.fdk import fdk, fdk_adjoint
    from monte_amd.mc import McSettings, counts_to_lineint, mc_project, primary_exact

    geom = G.small(n_views=36)
    ph = phantoms.small_phantom()
    spec = tables.kramers_spectrum(120.0, 2.5)
    s = McSettings(photons_per_view=400_000, spectrum=spec, coherent=1)
    vol_ref = fdk(primary_exact(ph, geom, spectrum=spec, ctx=ctx), geom, ctx).contiguous()
    E = EndToEnd(ph, geom, s, vol_ref, ctx=ctx)
    st = E.step(graph=False)
    torch.cuda.synchronize()
    _conservation(E, st)
    p, sc, _ = mc_project(ph, geom, s, ctx=ctx)
    L = counts_to_lineint(p + sc, geom, s.photons_per_view, ctx)
    v = fdk(L, geom, ctx)
    r = (v - vol_ref).contiguous()
    gr = fdk_adjoint(r, geom, ctx)
    torch.cuda.synchronize()
    for name, got, ref in (("primary", E.primary, p), ("scatter", E.scatter, sc), ("lineint", E.lineint, L),
                           ("vol", E.vol, v), ("residual", E.residual, r), ("grad", E.grad, gr)):
        assert torch.equal(got, ref), name
    loss = 0.5 * float((E.residual.double() ** 2).sum())
    print("loss", float(E.loss), loss)
    assert abs(float(E.loss) - loss) <= 1e-12 * loss and loss > 0
    ref = _snapshot(E)
    for i in range(3):
        E.step(graph=True)
        torch.cuda.synchronize()
        _assert_same(E, ref, f"graph replay {i}")
        fdk(torch.randn_like(E.lineint), geom, ctx)  # other work on the same workspaces between replays
        torch.cuda.synchronize()


def test_e2e_full_size_graph_equals_eager(ctx):
    import torch

    from monte_amd import workloads
    from monte_amd.e2e import EndToEnd
    from monte_amd.fdk import fdk
    from monte_amd.mc import McSettings, device_labels, primary_exact

    geom, ph, spec = workloads.e2e_inputs()
    print(f"phantom {ph.labels.shape}, detector {geom.nu}x{geom.nv}, {geom.n_views} views", flush=True)
    labels = device_labels(ph, "cuda")
    s = McSettings(photons_per_view=50_000_000, spectrum=spec, coherent=1)
    L = primary_exact(ph, geom, spectrum=spec, labels_dev=labels, ctx=ctx)
    vol_ref = fdk(L, geom, ctx).contiguous()
    del L
    E = EndToEnd(ph, geom, s, vol_ref, labels_dev=labels, ctx=ctx)
    st = E.step(graph=False)
    torch.cuda.synchronize()
    print("eager", st, flush=True)
    _conservation(E, st)
    ref = _snapshot(E)
    # the adjoint identity on the step's own operands, FP64 dot products
    lhs, rhs = _dot(E.vol, E.residual), _dot(E.lineint, E.grad)
    print(f"<A x, r> {lhs:.9e}  <x, A^T r> {rhs:.9e}  rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    loss = 0.5 * float((E.residual.double() ** 2).sum())
    assert abs(float(E.loss) - loss) <= 1e-12 * loss
    for i in range(2):
        stg = E.step(graph=True)
        torch.cuda.synchronize()
        print(f"graph {i}: {stg['kernel_ms']:.1f} ms", flush=True)
        _assert_same(E, ref, f"graph launch {i}")
        assert {k: v for k, v in stg.items() if k != "kernel_ms"} == {k: v for k, v in st.items() if k != "kernel_ms"}
        E.vol.fill_(0.0)  # eager work between replays
        torch.cuda.synchronize()
    assert np.isfinite(float(E.loss)) and float(E.loss) > 0
