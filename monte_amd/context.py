"""Context: one monte_ctx (device, stream, tables, tuning, workspaces)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .tables import reference_tables


class Context:
    """Owns a monte_ctx on one device.  Calls run on torch's current stream of
    that device (set before every call), so they order with surrounding torch work."""

    def __init__(self, device: int = 0, tables: np.ndarray | None = None):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("monte_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = device
        h = C.c_void_p()
        _lib.check(_lib.lib().monte_create(C.byref(h), device), what="monte_create")
        self._h = h
        self.tables = reference_tables() if tables is None else np.ascontiguousarray(tables, dtype=np.float64)
        self._call("monte_load_tables", self.tables.ctypes.data_as(C.c_void_p), int(self.tables.shape[0]))

    @property
    def handle(self):
        return self._h

    def sync_stream(self):
        import torch

        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.lib().monte_set_stream(self._h, C.c_void_p(s)), self._h, "monte_set_stream")

    def _call(self, name: str, *args):
        fn = getattr(_lib.lib(), name)
        return _lib.check(fn(self._h, *args), self._h, name)

    def call(self, name: str, *args):
        """Run a monte_* entry point on torch's current stream."""
        self.sync_stream()
        return self._call(name, *args)

    @property
    def tuning(self) -> dict:
        t = _lib.Tuning()
        self._call("monte_get_tuning", C.byref(t))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _lib.Tuning()
        self._call("monte_get_tuning", C.byref(t))
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._call("monte_set_tuning", C.byref(t))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().monte_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]
