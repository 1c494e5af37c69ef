"""The differentiable scan end to end, as one call (monte_e2e_step):

  MC projection (primary + scatter tallies) -> line integrals of their sum ->
  FDK -> L2 loss against a target volume -> gradient of the loss with respect
  to the line integrals (the FDK adjoint applied to the residual)

optionally captured once into a hipGraph and replayed.  The reference runs the
first three as separate programs joined by raw files (monte_cu/CBCT_real325.cu:
241-300, then recon/bp3d20_325.cpp:39-175) and has no loss or gradient.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .context import Context, default_context
from .fdk import _graph_stream
from .geometry import Geometry
from .mc import McSettings, _check_labels, device_labels
from .phantoms import LabelPhantom


class EndToEnd:
    """Device buffers of the step, allocated once so graph replays reuse them.

    After step(): primary, scatter (FP64 tallies), lineint, vol, residual,
    grad (FP32) and loss (a one-element FP64 tensor) hold the step's results."""

    def __init__(self, ph: LabelPhantom, geom: Geometry, settings: McSettings, vol_ref: torch.Tensor,
                 labels_dev: torch.Tensor | None = None, ctx: Context | None = None):
        self.ctx = ctx or default_context(0)
        dev = torch.device("cuda", self.ctx.device)
        if vol_ref.dtype != torch.float32 or tuple(vol_ref.shape) != geom.vol_shape or not vol_ref.is_contiguous() \
                or vol_ref.device != dev:
            raise TypeError(f"vol_ref must be a contiguous float32 tensor {geom.vol_shape} on {dev}")
        if settings.views is not None and tuple(settings.views) != (0, geom.n_views):
            raise ValueError("the end-to-end step simulates every view")
        self.ph, self.geom, self.settings = ph, geom, settings
        self.labels = device_labels(ph, dev) if labels_dev is None else labels_dev
        _check_labels(self.labels, ph, self.ctx)
        self.vol_ref = vol_ref
        self.primary = torch.zeros(geom.proj_shape, dtype=torch.float64, device=dev)
        self.scatter = torch.zeros_like(self.primary)
        self.lineint = torch.zeros(geom.proj_shape, dtype=torch.float32, device=dev)
        self.grad = torch.zeros_like(self.lineint)
        self.vol = torch.zeros(geom.vol_shape, dtype=torch.float32, device=dev)
        self.residual = torch.zeros_like(self.vol)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)

    def step(self, graph: bool = True) -> dict:
        b = _lib.E2EBuffers()
        for name in ("primary", "scatter", "lineint", "vol", "vol_ref", "residual", "grad", "loss"):
            setattr(b, name, getattr(self, name).data_ptr())
        g = self.geom.c()
        p = self.ph.c(self.labels.data_ptr())
        cfg, keep = self.settings.c()
        st = _lib.McStats()
        if graph:
            # capture needs a created stream: run on a side stream ordered against the caller's
            dev = self.vol.device
            cur = torch.cuda.current_stream(dev)
            side = _graph_stream(dev)
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                self.ctx.call("monte_e2e_step", C.byref(g), C.byref(p), C.byref(cfg), C.byref(b), 1, C.byref(st))
            cur.wait_stream(side)
        else:
            self.ctx.call("monte_e2e_step", C.byref(g), C.byref(p), C.byref(cfg), C.byref(b), 0, C.byref(st))
        del keep
        return st.as_dict()
