"""Raw files in the reference's layouts, each with a JSON sidecar.

The reference programs exchange headerless little-endian C-order arrays:

  phantom   uint8   [z][y][x]              (monte_cpp/make_image01.cpp:29, CBCT_real325im.cu)
  proj*     int32   [view][u][v] counts    (monte_cu/CBCT_real325.cu:296,299)
  map*      float32 [view][u][v] ln N0/N   (CBCT_real325.cu:297,300; read by recon/bp3d20_325.cpp:29-30)
  xy*       float32 [z][y][x] volume       (recon/bp3d20_325.cpp:180)
  zy*       float32 [x][y][z] volume       (recon/bp3d20_325.cpp:186)
  filtered  float32 [view][v][u]           (recon/bp3d20_325.cpp:196)

The raw bytes are exactly those layouts, so files interchange with the
reference programs.  Next to each `name.raw` a `name.raw.json` records dtype,
shape, layout and the geometry or phantom description, so a file can be read
back without knowing its producer's constants.
"""
from __future__ import annotations

import json
import os

import numpy as np

LAYOUTS = {
    "phantom": ("uint8", "[z][y][x]"),
    "counts": ("int32", "[view][u][v]"),
    "lineint": ("float32", "[view][u][v]"),
    "volume": ("float32", "[z][y][x]"),
    "volume_zy": ("float32", "[x][y][z]"),
    "filtered": ("float32", "[view][v][u]"),
}


def sidecar_path(path: str) -> str:
    return path + ".json"


def write_raw(path: str, arr: np.ndarray, kind: str, meta: dict | None = None) -> None:
    """Write `arr` as headerless little-endian raw bytes plus its sidecar."""
    if kind not in LAYOUTS:
        raise ValueError(f"unknown raw kind {kind!r}")
    dtype, layout = LAYOUTS[kind]
    a = np.ascontiguousarray(arr, dtype=np.dtype(dtype).newbyteorder("<"))
    a.tofile(path)
    side = {"kind": kind, "dtype": dtype, "endian": "little", "layout": layout, "shape": list(a.shape)}
    if meta:
        side["meta"] = meta
    with open(sidecar_path(path), "w") as f:
        json.dump(side, f, indent=1, sort_keys=True)


def read_raw(path: str, kind: str | None = None, shape=None) -> tuple[np.ndarray, dict]:
    """Read a raw file.  The sidecar gives dtype and shape; without one, `kind`
    and `shape` must be given (as for a file the reference wrote)."""
    side = {}
    if os.path.exists(sidecar_path(path)):
        with open(sidecar_path(path)) as f:
            side = json.load(f)
        if kind is not None and side.get("kind") != kind:
            raise ValueError(f"{path}: sidecar says {side.get('kind')!r}, expected {kind!r}")
        kind = side["kind"]
        shape = tuple(side["shape"])
    if kind is None or shape is None:
        raise ValueError(f"{path}: no sidecar, so kind and shape are required")
    dtype = np.dtype(LAYOUTS[kind][0]).newbyteorder("<")
    n = int(np.prod(shape))
    size = os.path.getsize(path)
    if size != n * dtype.itemsize:
        raise ValueError(f"{path}: {size} bytes, expected {n * dtype.itemsize} for {kind} {tuple(shape)}")
    arr = np.fromfile(path, dtype=dtype).reshape(shape)
    return arr.astype(dtype.newbyteorder("="), copy=False), side.get("meta", {})
