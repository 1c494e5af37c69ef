"""FDK reconstruction and its adjoint on torch device tensors.

  fdk(proj, geom)                    [B][V][nu][nv] or [V][nu][nv] -> [B][nz][ny][nx]
  fdk_adjoint(dvol, geom)            exact transpose of fdk
  fdk_filter / fdk_backproject / fdk_project / fdk_filter_adjoint    the stages
  FDKFunction / fdk_autograd         torch.autograd.Function: backward = fdk_adjoint
  fdk_graph(proj, geom, out)         the same FDK replayed from a captured hipGraph

Replaces recon/bp3d20_325.cpp:main (and bp3d20.cpp): weight (:39-46), ramp
filter (:51-81), backprojection (:91-175).  The adjoint has no reference
counterpart.
"""
from __future__ import annotations

import ctypes as C

import torch

from .context import Context, default_context
from .geometry import Geometry


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _batched(x: torch.Tensor, inner: tuple, what: str):
    if x.dtype != torch.float32 or not x.is_cuda:
        raise TypeError(f"{what}: expected a float32 CUDA tensor, got {x.dtype} on {x.device}")
    if tuple(x.shape[-3:]) != tuple(inner):
        raise ValueError(f"{what}: expected trailing shape {tuple(inner)}, got {tuple(x.shape)}")
    squeeze = x.dim() == 3
    xb = x.unsqueeze(0) if squeeze else x.reshape(-1, *inner)
    return xb.contiguous(), squeeze


def _run(name, x, geom, inner_in, inner_out, ctx, what, extra=()):
    ctx = ctx or default_context(x.device.index or 0)
    xb, squeeze = _batched(x, inner_in, what)
    out = torch.empty((xb.shape[0], *inner_out), dtype=torch.float32, device=x.device)
    g = geom.c()
    ctx.call(name, C.byref(g), _ptr(xb), _ptr(out), xb.shape[0], *extra)
    return out[0] if squeeze else out


def fdk(proj: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    return _run("monte_fdk", proj, geom, geom.proj_shape, geom.vol_shape, ctx, "fdk")


def fdk_filter(proj: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    return _run("monte_fdk_filter", proj, geom, geom.proj_shape, geom.filt_shape, ctx, "fdk_filter")


def fdk_backproject(filt: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    return _run("monte_fdk_backproject", filt, geom, geom.filt_shape, geom.vol_shape, ctx, "fdk_backproject",
                extra=(0,))


def fdk_project(vol: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    return _run("monte_fdk_project", vol, geom, geom.vol_shape, geom.filt_shape, ctx, "fdk_project")


def fdk_filter_adjoint(filt: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    return _run("monte_fdk_filter_adjoint", filt, geom, geom.filt_shape, geom.proj_shape, ctx,
                "fdk_filter_adjoint")


def fdk_adjoint(dvol: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    return _run("monte_fdk_adjoint", dvol, geom, geom.vol_shape, geom.proj_shape, ctx, "fdk_adjoint")


def fdk_graph(proj: torch.Tensor, geom: Geometry, out: torch.Tensor, ctx: Context | None = None) -> torch.Tensor:
    """FDK replayed from a hipGraph captured on first use (re-captured when the
    geometry, the buffers or the tuning change).  `out` must stay allocated."""
    ctx = ctx or default_context(proj.device.index or 0)
    xb, _ = _batched(proj, geom.proj_shape, "fdk_graph")
    if xb.data_ptr() != proj.data_ptr():
        raise ValueError("fdk_graph needs a contiguous input")
    if out.dtype != torch.float32 or not out.is_contiguous() or out.numel() != xb.shape[0] * geom.nx * geom.ny * geom.nz:
        raise ValueError("fdk_graph: out must be a contiguous float32 tensor of the volume batch size")
    g = geom.c()
    # capture needs a created stream: run on a side stream ordered against the caller's
    cur = torch.cuda.current_stream(proj.device)
    side = _graph_stream(proj.device)
    side.wait_stream(cur)
    with torch.cuda.stream(side):
        ctx.call("monte_fdk_graph", C.byref(g), _ptr(xb), _ptr(out), xb.shape[0])
    cur.wait_stream(side)
    return out


_graph_streams: dict = {}


def _graph_stream(device):
    key = device.index or 0
    if key not in _graph_streams:
        _graph_streams[key] = torch.cuda.Stream(device)
    return _graph_streams[key]


class FDKFunction(torch.autograd.Function):
    """vol = FDK(proj); d loss / d proj = FDK^T(d loss / d vol).

    FDK is linear, so the backward is exactly the adjoint operator (the
    matched projector, the transposed filter and the cosine weight)."""

    @staticmethod
    def forward(ctx, proj, geom, mctx):
        ctx.geom = geom
        ctx.mctx = mctx
        return fdk(proj, geom, mctx)

    @staticmethod
    def backward(ctx, dvol):
        return fdk_adjoint(dvol.contiguous(), ctx.geom, ctx.mctx), None, None


def fdk_autograd(proj: torch.Tensor, geom: Geometry, ctx: Context | None = None) -> torch.Tensor:
    ctx = ctx or default_context(proj.device.index or 0)
    return FDKFunction.apply(proj, geom, ctx)
