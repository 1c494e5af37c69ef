"""The FDK adjoint and its torch.autograd binding on the GPU.

* <A x, y> = <x, A^T y> to 1e-5 relative (sums in FP64) for the whole operator
  and for each stage, in both presets;
* A^T against the oracle's explicit FP64 adjoint;
* gradients of a loss on the FDK volume w.r.t. the projections against FP64
  central finite differences of the oracle's FDK, at 33^3 / 64^2 / 30 views.
"""
import numpy as np
import pytest

from conftest import oracle_geom, rel_l2, rel_max

pytestmark = pytest.mark.gpu


def _t(x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _dot(a, b):
    return float(np.sum(np.asarray(a, np.float64) * np.asarray(b, np.float64)))


def _geoms():
    from monte_amd import geometry as G

    std = G.fd_check().replace(n_views=20, nu=48, nv=40, u_off=0.05, v_off=-0.04, ox=0.03, oz=0.05)
    shepp = std.replace(filter=1, n_views=16)
    compat = G.bp3d20_65(n_views=24).replace(nx=64, ny=64, nz=64, roi_x0=20, roi_x1=44, vx=0.4, vy=0.4, vz=0.4,
                                             mask_r=-1.0)
    compat_mask = compat.replace(mask_r=25.0)
    # thin slices under a coarse detector: the v-slope per voxel is 0.4 and 0.2
    # rows, so the projector interleaves 2 and 3 lane phases per column
    thin = std.replace(nz=60, vz=0.02, n_views=12)
    thinner = std.replace(nz=90, vz=0.01, n_views=8)
    # a v-slope of 0.8 rows per voxel, as in the 325^3 / 512^2 workload: two
    # voxels per lane fold into three row updates
    mid = std.replace(nz=45, vz=0.04, n_views=12)
    # tall columns: the projector walks several 64-lane passes per column and
    # runs their LDS updates round by round together, folded (v-slope 0.8) and
    # unfolded (1.1 rows per voxel)
    tall = std.replace(nx=24, ny=24, nz=300, vz=0.04, nv=256, n_views=8)
    tall1 = std.replace(nx=24, ny=24, nz=400, vz=0.055, nv=480, n_views=8)
    return {"std": std, "shepp": shepp, "bp3d20": compat, "bp3d20_mask": compat_mask, "thin": thin,
            "thinner": thinner, "mid": mid, "tall": tall, "tall1": tall1}


@pytest.mark.parametrize("name", ["std", "shepp", "bp3d20", "bp3d20_mask", "thin", "thinner", "mid", "tall", "tall1"])
def test_dot_product_identity(ctx, name):
    from monte_amd.fdk import (fdk, fdk_adjoint, fdk_backproject, fdk_filter, fdk_filter_adjoint,
                               fdk_project)

    g = _geoms()[name]
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.proj_shape).astype(np.float32)
    y = rng.standard_normal(g.vol_shape).astype(np.float32)
    if g.preset == 1:  # only the ROI is reconstructed
        y[:, :, : g.roi_x0] = 0
        y[:, :, g.roi_x1:] = 0
    ax = fdk(_t(x), g, ctx).cpu().numpy()
    aty = fdk_adjoint(_t(y), g, ctx).cpu().numpy()
    lhs, rhs = _dot(ax, y), _dot(x, aty)
    print(name, lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # stages
    f = rng.standard_normal(g.filt_shape).astype(np.float32)
    b1, b2 = _dot(fdk_backproject(_t(f), g, ctx).cpu().numpy(), y), _dot(f, fdk_project(_t(y), g, ctx).cpu().numpy())
    assert abs(b1 - b2) <= 1e-5 * abs(b1)
    f1, f2 = _dot(fdk_filter(_t(x), g, ctx).cpu().numpy(), f), _dot(x, fdk_filter_adjoint(_t(f), g, ctx).cpu().numpy())
    assert abs(f1 - f2) <= 1e-5 * abs(f1)


@pytest.mark.parametrize("name", ["std", "thinner", "mid", "tall", "tall1"])
def test_adjoint_matches_oracle(ctx, oracle, name):
    from monte_amd.fdk import fdk_adjoint

    g = _geoms()[name]
    rng = np.random.default_rng(5)
    y = rng.standard_normal(g.vol_shape).astype(np.float32)
    got = fdk_adjoint(_t(y), g, ctx).cpu().numpy()
    ref = oracle.fdk_std_adjoint(oracle_geom(oracle, g), y)
    print("adjoint rel_max", rel_max(got, ref), "rel_l2", rel_l2(got, ref))
    assert rel_max(got, ref) < 1e-5
    assert rel_l2(got, ref) < 1e-5


@pytest.mark.parametrize("n_views", [1, 5, 13, 27])
def test_projector_view_counts(ctx, oracle, n_views):
    """The projector hands views to the 8 XCDs in contiguous ranges of groups of 8; view counts
    that leave XCDs empty (1, 5), split unevenly (13) or end in a partial group (27) must still
    project every (view, strip) once."""
    from monte_amd.fdk import fdk_adjoint, fdk_backproject, fdk_project

    g = _geoms()["mid"].replace(n_views=n_views)
    rng = np.random.default_rng(11 + n_views)
    y = rng.standard_normal(g.vol_shape).astype(np.float32)
    got = fdk_adjoint(_t(y), g, ctx).cpu().numpy()
    ref = oracle.fdk_std_adjoint(oracle_geom(oracle, g), y)
    print(n_views, "adjoint rel_max", rel_max(got, ref), "rel_l2", rel_l2(got, ref))
    assert rel_max(got, ref) < 1e-5
    assert rel_l2(got, ref) < 1e-5
    f = rng.standard_normal(g.filt_shape).astype(np.float32)
    p = fdk_project(_t(y), g, ctx).cpu().numpy()
    assert np.all(np.any(p != 0, axis=tuple(range(1, p.ndim)))), "a view was left unprojected"
    b1, b2 = _dot(fdk_backproject(_t(f), g, ctx).cpu().numpy(), y), _dot(f, p)
    assert abs(b1 - b2) <= 1e-5 * abs(b1)


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 33, 1), (40, 3, 17), (13, 29, 70)])
def test_odd_volume_shapes(ctx, oracle, shape):
    """Volumes (nx, ny, nz) of one voxel, one slice, a few rows, and taller than the detector's
    field of view: partial backprojection tiles and z-chunks in every direction, and projector
    column lists over slabs of one to 40 voxels, against the oracle's FP64 operators."""
    from monte_amd.fdk import fdk_adjoint, fdk_backproject, fdk_project

    g = _geoms()["std"].replace(nx=shape[0], ny=shape[1], nz=shape[2], n_views=12)
    og = oracle_geom(oracle, g)
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal(g.filt_shape).astype(np.float32)
    v = fdk_backproject(_t(f), g, ctx).cpu().numpy()
    v_ref = oracle.fdk_std_backproject(og, f)
    y = rng.standard_normal(g.vol_shape).astype(np.float32)
    a = fdk_adjoint(_t(y), g, ctx).cpu().numpy()
    a_ref = oracle.fdk_std_adjoint(og, y)
    print(shape, "backprojection", rel_max(v, v_ref), "adjoint", rel_max(a, a_ref), rel_l2(a, a_ref))
    assert v.shape == v_ref.shape == g.vol_shape
    assert rel_max(v, v_ref) < 1e-5
    assert rel_max(a, a_ref) < 1e-5 and rel_l2(a, a_ref) < 1e-5
    b1, b2 = _dot(v, y), _dot(f, fdk_project(_t(y), g, ctx).cpu().numpy())
    assert abs(b1 - b2) <= 1e-5 * abs(b1)


def test_autograd_matches_fp64_finite_differences(ctx, oracle):
    """d/dp of L = sum(g * FDK(p)) at 33^3 / 64^2 / 30 views."""
    import torch

    from monte_amd import geometry as G
    from monte_amd.fdk import fdk_autograd

    geom = G.fd_check()
    og = oracle_geom(oracle, geom)
    rng = np.random.default_rng(8)
    p = rng.standard_normal(geom.proj_shape)
    gw = rng.standard_normal(geom.vol_shape)
    x = _t(p.astype(np.float32)).requires_grad_(True)
    vol = fdk_autograd(x, geom, ctx)
    (vol * _t(gw.astype(np.float32))).sum().backward()
    grad = x.grad.cpu().numpy()
    # FP64 central differences of the oracle FDK at a sample of entries
    eps = 1e-3
    idx = [tuple(int(rng.integers(0, s)) for s in geom.proj_shape) for _ in range(12)]
    # entries whose ray meets the volume, so the derivative is not trivially zero
    idx += [(k, 32, 32) for k in (0, 7, 29)]
    fd = []
    for i in idx:
        e = np.zeros(geom.proj_shape)
        e[i] = eps
        lp = np.sum(gw * oracle.fdk_std(og, p + e))
        lm = np.sum(gw * oracle.fdk_std(og, p - e))
        fd.append((lp - lm) / (2 * eps))
    fd = np.array(fd)
    got = np.array([grad[i] for i in idx])
    scale = np.abs(grad).max()
    print("fd", fd[-3:], "gpu", got[-3:], "max|diff|/max|grad|", np.abs(got - fd).max() / scale)
    assert np.abs(got - fd).max() <= 1e-5 * scale
    # and the whole gradient against the explicit FP64 adjoint
    ref = oracle.fdk_std_adjoint(og, gw)
    assert rel_max(grad, ref) < 1e-5


def test_autograd_batch_and_graph_of_ops(ctx):
    """Gradients flow through a batch and through torch ops around the FDK."""
    import torch

    from monte_amd import geometry as G
    from monte_amd.fdk import fdk_adjoint, fdk_autograd

    geom = G.fd_check().replace(n_views=10)
    rng = np.random.default_rng(9)
    x = _t(rng.standard_normal((3, *geom.proj_shape)).astype(np.float32)).requires_grad_(True)
    target = _t(rng.standard_normal((3, *geom.vol_shape)).astype(np.float32))
    vol = fdk_autograd(x * 2.0, geom, ctx)
    loss = 0.5 * ((vol - target) ** 2).sum()
    loss.backward()
    expect = 2.0 * fdk_adjoint((vol - target).detach().contiguous(), geom, ctx)
    assert torch.allclose(x.grad, expect, rtol=1e-6, atol=1e-6 * expect.abs().max().item())


@pytest.mark.parametrize("name", ["std", "thin", "mid", "bp3d20_mask", "tall", "tall1"])
def test_projector_agrees_across_lane_chunks(ctx, name):
    """Every voxels-per-lane variant of the projector is the same transpose:
    each passes the dot-product identity, and they agree with each other up
    to FP32 summation order."""
    from monte_amd.fdk import fdk_backproject, fdk_project

    g = _geoms()[name]
    rng = np.random.default_rng(11)
    y = rng.standard_normal(g.vol_shape).astype(np.float32)
    f = rng.standard_normal(g.filt_shape).astype(np.float32)
    bf = _dot(fdk_backproject(_t(f), g, ctx).cpu().numpy(), y)
    saved = ctx.tuning
    outs = {}
    try:
        for cz in (2, 4, 8):
            ctx.set_tuning(pj_zchunk=cz)
            outs[cz] = fdk_project(_t(y), g, ctx).cpu().numpy()
            pf = _dot(f, outs[cz])
            assert abs(bf - pf) <= 1e-5 * abs(bf), (cz, bf, pf)
    finally:
        ctx.set_tuning(**saved)
    for cz in (4, 8):
        assert rel_max(outs[cz], outs[2]) < 1e-6, cz


@pytest.mark.parametrize("name", ["std", "bp3d20", "tall1"])
def test_projector_is_deterministic(ctx, name):
    """Repeated projections are bitwise equal: the strip edge columns take two
    addends onto zero, and the bp3d20 guard cells are folded back by a gather
    in a fixed order (three guard cells can alias one interior cell)."""
    from monte_amd.fdk import fdk_adjoint, fdk_project

    g = _geoms()[name]
    y = _t(np.random.default_rng(5).standard_normal(g.vol_shape).astype(np.float32))
    ref = fdk_project(y, g, ctx).cpu().numpy()
    for _ in range(4):
        assert np.array_equal(fdk_project(y, g, ctx).cpu().numpy(), ref)
    a = fdk_adjoint(y, g, ctx).cpu().numpy()
    assert np.array_equal(fdk_adjoint(y, g, ctx).cpu().numpy(), a)
