"""Monte Carlo projection ("phantom in, projections out") on torch device tensors.

  mc_project(phantom, geom, photons_per_view, ...)  -> (primary, scatter, stats)
  mc_batch(jobs)                                    many uneven jobs, one persistent launch
  primary_exact(phantom, geom, e_kev | spectrum)    exact Beer-Lambert line integrals (Siddon)
  counts_to_lineint(counts, geom, photons_per_view) ln N0 - ln clamp(N, 1, N0)

Replaces the projection kernel of monte_cu/CBCT_real325.cu:419-975 and its
count post-process :264-294.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from .context import Context, default_context
from .geometry import Geometry
from .phantoms import LabelPhantom


@dataclasses.dataclass
class McSettings:
    photons_per_view: int
    e_kev: float = 60.0
    spectrum: np.ndarray | None = None     # weights per keV bin [201] (host)
    max_interactions: int = 5              # ScatterNUM, CBCT_real325.cu:9
    coherent: int = _lib.COH_RAYLEIGH
    seed: int = 20261202
    job: int = 0
    views: tuple | None = None             # (begin, end)

    def c(self):
        cfg = _lib.McCfg()
        cfg.photons_per_view = int(self.photons_per_view)
        cfg.e_kev = float(self.e_kev)
        keep = None
        if self.spectrum is not None:
            keep = np.ascontiguousarray(self.spectrum, dtype=np.float64)
            cfg.spectrum = keep.ctypes.data
        cfg.max_interactions = int(self.max_interactions)
        cfg.coherent_mode = int(self.coherent)
        cfg.seed = int(self.seed)
        cfg.job = int(self.job)
        if self.views is not None:
            cfg.view_begin, cfg.view_end = self.views
        return cfg, keep


def photons_per_pixel(geom: Geometry, photons_per_view: int) -> np.ndarray:
    """Photons aimed at each pixel [nu][nv]: photon n aims at pixel n mod
    (nu nv) in the order iv * nu + iu, so the first N mod (nu nv) pixels of that
    order get one more than the rest."""
    npix = geom.nu * geom.nv
    base, rem = divmod(int(photons_per_view), npix)
    order = np.arange(npix, dtype=np.int64).reshape(geom.nv, geom.nu).T  # [iu][iv] -> iv * nu + iu
    return (base + (order < rem)).astype(np.float64)


def device_labels(ph: LabelPhantom, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(ph.labels)).to(device)


def _check_labels(labels: torch.Tensor, ph: LabelPhantom, ctx: Context) -> None:
    """The kernels read nx * ny * nz bytes of labels on the context's device."""
    if not isinstance(labels, torch.Tensor) or labels.dtype != torch.uint8 or not labels.is_contiguous():
        raise TypeError("labels must be a contiguous uint8 tensor")
    if not labels.is_cuda or labels.device.index != ctx.device:
        raise TypeError(f"labels must be on cuda:{ctx.device}")
    if labels.numel() != ph.labels.size:
        raise ValueError(f"labels hold {labels.numel()} voxels, the phantom {ph.labels.size}")


def _check_tally(t: torch.Tensor, geom: Geometry, ctx: Context, what: str) -> None:
    """Tallies take FP64 atomics over [views][nu][nv] on the context's device."""
    shape = geom.proj_shape
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or tuple(t.shape) != shape \
            or not t.is_contiguous():
        raise TypeError(f"{what} must be a contiguous float64 tensor of shape {shape}")
    if not t.is_cuda or t.device.index != ctx.device:
        raise TypeError(f"{what} must be on cuda:{ctx.device}")


def mc_project(ph: LabelPhantom, geom: Geometry, settings: McSettings, labels_dev: torch.Tensor | None = None,
               primary: torch.Tensor | None = None, scatter: torch.Tensor | None = None,
               ctx: Context | None = None, device: int = 0):
    """Tallies are accumulated into `primary`/`scatter` (zeros if not given)."""
    ctx = ctx or default_context(device)
    dev = torch.device("cuda", ctx.device)
    if labels_dev is None:
        labels_dev = device_labels(ph, dev)
    _check_labels(labels_dev, ph, ctx)
    shape = geom.proj_shape
    if primary is None:
        primary = torch.zeros(shape, dtype=torch.float64, device=dev)
    if scatter is None:
        scatter = torch.zeros(shape, dtype=torch.float64, device=dev)
    _check_tally(primary, geom, ctx, "primary")
    _check_tally(scatter, geom, ctx, "scatter")
    g = geom.c()
    p = ph.c(labels_dev.data_ptr())
    cfg, _keep = settings.c()
    st = _lib.McStats()
    ctx.call("monte_mc_project", C.byref(g), C.byref(p), C.byref(cfg), C.c_void_p(primary.data_ptr()),
             C.c_void_p(scatter.data_ptr()), C.byref(st))
    return primary, scatter, st.as_dict()


@dataclasses.dataclass
class McJobSpec:
    phantom: LabelPhantom
    geom: Geometry
    settings: McSettings
    labels_dev: torch.Tensor | None = None
    primary: torch.Tensor | None = None
    scatter: torch.Tensor | None = None


def mc_batch(jobs: list, ctx: Context | None = None, device: int = 0):
    """Runs all jobs in one persistent launch; returns per-job stats."""
    ctx = ctx or default_context(device)
    dev = torch.device("cuda", ctx.device)
    keep = []
    arr = (_lib.McJob * len(jobs))()
    for i, j in enumerate(jobs):
        if j.labels_dev is None:
            j.labels_dev = device_labels(j.phantom, dev)
        if j.primary is None:
            j.primary = torch.zeros(j.geom.proj_shape, dtype=torch.float64, device=dev)
        if j.scatter is None:
            j.scatter = torch.zeros(j.geom.proj_shape, dtype=torch.float64, device=dev)
        _check_labels(j.labels_dev, j.phantom, ctx)
        _check_tally(j.primary, j.geom, ctx, f"job {i} primary")
        _check_tally(j.scatter, j.geom, ctx, f"job {i} scatter")
        g = j.geom.c()
        p = j.phantom.c(j.labels_dev.data_ptr())
        cfg, k = j.settings.c()
        keep += [g, p, cfg, k]
        arr[i].geom = C.pointer(g)
        arr[i].phantom = C.pointer(p)
        arr[i].cfg = C.pointer(cfg)
        arr[i].primary = j.primary.data_ptr()
        arr[i].scatter = j.scatter.data_ptr()
    stats = (_lib.McStats * len(jobs))()
    ctx.call("monte_mc_batch", arr, len(jobs), stats)
    return [s.as_dict() for s in stats]


def first_flight_sets(ph: LabelPhantom, geom: Geometry, brick: int, labels_dev: torch.Tensor | None = None,
                      ctx: Context | None = None, device: int = 0) -> torch.Tensor:
    """The label sets bounding each first flight's Woodcock majorant, uint8
    [views][nv * nu] (pixel iv * nu + iu): bit l for label l < 4, 16 for any
    label >= 4 (or a line missing the flight box)."""
    ctx = ctx or default_context(device)
    dev = torch.device("cuda", ctx.device)
    if labels_dev is None:
        labels_dev = device_labels(ph, dev)
    _check_labels(labels_dev, ph, ctx)
    out = torch.zeros((geom.n_views, geom.nu * geom.nv), dtype=torch.uint8, device=dev)
    g = geom.c()
    p = ph.c(labels_dev.data_ptr())
    ctx.call("monte_mc_first_flight_sets", C.byref(g), C.byref(p), int(brick), C.c_void_p(out.data_ptr()))
    return out


def primary_exact(ph: LabelPhantom, geom: Geometry, e_kev: float = 60.0, spectrum: np.ndarray | None = None,
                  labels_dev: torch.Tensor | None = None, ctx: Context | None = None, device: int = 0):
    ctx = ctx or default_context(device)
    dev = torch.device("cuda", ctx.device)
    if labels_dev is None:
        labels_dev = device_labels(ph, dev)
    _check_labels(labels_dev, ph, ctx)
    out = torch.empty(geom.proj_shape, dtype=torch.float32, device=dev)
    g = geom.c()
    p = ph.c(labels_dev.data_ptr())
    sp = None
    if spectrum is not None:
        sp = np.ascontiguousarray(spectrum, dtype=np.float64)
    ctx.call("monte_primary_exact", C.byref(g), C.byref(p), float(e_kev),
             sp.ctypes.data_as(C.c_void_p) if sp is not None else None, C.c_void_p(out.data_ptr()))
    return out


def counts_to_lineint(counts: torch.Tensor, geom: Geometry, photons_per_view: int, ctx: Context | None = None):
    ctx = ctx or default_context(counts.device.index or 0)
    if counts.dtype != torch.float64 or tuple(counts.shape) != geom.proj_shape:
        raise TypeError(f"counts must be float64 {geom.proj_shape}")
    if not counts.is_cuda or counts.device.index != ctx.device:
        raise TypeError(f"counts must be on cuda:{ctx.device}")
    counts = counts.contiguous()
    out = torch.empty(geom.proj_shape, dtype=torch.float32, device=counts.device)
    g = geom.c()
    ctx.call("monte_counts_to_lineint", C.byref(g), C.c_void_p(counts.data_ptr()), int(photons_per_view),
             C.c_void_p(out.data_ptr()))
    return out


def counts_to_lineint_flat(counts: torch.Tensor, geom: Geometry, flat: torch.Tensor, ctx: Context | None = None):
    """ln(F) - ln(clamp(N, 1, F)) against a measured flat field F: float64 [nu][nv]
    (one air scan for every view) or [views][nu][nv]; 0 where F <= 0."""
    ctx = ctx or default_context(counts.device.index or 0)
    if counts.dtype != torch.float64 or tuple(counts.shape) != geom.proj_shape:
        raise TypeError(f"counts must be float64 {geom.proj_shape}")
    pix = (geom.nu, geom.nv)
    if flat.dtype != torch.float64 or tuple(flat.shape) not in (pix, (1, *pix), geom.proj_shape):
        raise TypeError(f"flat must be float64 {pix} or {geom.proj_shape}")
    if not counts.is_cuda or flat.device != counts.device:
        raise TypeError("counts and flat must be on the same CUDA device")
    counts, flat = counts.contiguous(), flat.contiguous()
    flat_views = 1 if flat.numel() == geom.nu * geom.nv else geom.n_views
    out = torch.empty(geom.proj_shape, dtype=torch.float32, device=counts.device)
    g = geom.c()
    ctx.call("monte_counts_to_lineint_flat", C.byref(g), C.c_void_p(counts.data_ptr()), C.c_void_p(flat.data_ptr()),
             flat_views, C.c_void_p(out.data_ptr()))
    return out


def autotune(ph: LabelPhantom | None, geom: Geometry, settings: McSettings | None, what: int = 3,
             labels_dev: torch.Tensor | None = None, ctx: Context | None = None, device: int = 0) -> dict:
    """Times candidate launch shapes and keeps the fastest in the context.
    what: bit 1 MC, bit 2 backprojection, bit 4 matched projector."""
    ctx = ctx or default_context(device)
    dev = torch.device("cuda", ctx.device)
    g = geom.c()
    pp = cc = None
    keep = None
    if what & 1:
        if labels_dev is None:
            labels_dev = device_labels(ph, dev)
        _check_labels(labels_dev, ph, ctx)
        pp = ph.c(labels_dev.data_ptr())
        cc, keep = settings.c()
    t = _lib.Tuning()
    ctx.call("monte_autotune", int(what), C.byref(g), C.byref(pp) if pp is not None else None,
             C.byref(cc) if cc is not None else None, C.byref(t))
    del keep
    return t.as_dict()
