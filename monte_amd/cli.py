"""File-level entry points with the reference programs' inputs and outputs.

    python -m monte_amd.cli phantom --preset full --out phantom.raw
    python -m monte_amd.cli mc --phantom phantom.raw --geometry full --photons 1e8 --out run
        -> run_proj0.raw / run_proj5.raw (int32 counts: primary / primary+scatter)
           run_map0.raw  / run_map5.raw  (float32 line integrals)
    python -m monte_amd.cli fdk --map run_map5.raw --geometry full --out vol.raw [--zy zy.raw] [--filtered f.raw]

`mc` stands in for monte_cu/CBCT_real325.cu (phantom .raw in, proj*/map*
.raw out, :264-300); `fdk` for recon/bp3d20_325.cpp (map .raw in, volumes and
the filtered sinogram out, :29-30, :179-202).  Files are headerless raw arrays
in the reference's layouts with JSON sidecars (monte_amd.rawio).
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import geometry as G
from . import phantoms, rawio, tables

GEOMETRIES = {"full": G.full, "small": G.small, "e2e": G.e2e, "bp3d20_325": G.bp3d20_325, "bp3d20_65": G.bp3d20_65}
PHANTOMS = {"full": phantoms.full_phantom, "small": phantoms.small_phantom}


def _geometry(name: str, views: int | None) -> G.Geometry:
    if name not in GEOMETRIES:
        raise SystemExit(f"unknown geometry {name!r}; choose from {sorted(GEOMETRIES)}")
    return GEOMETRIES[name](n_views=views) if views else GEOMETRIES[name]()


def _phantom_meta(ph: phantoms.LabelPhantom) -> dict:
    return {"voxel": list(ph.voxel), "origin": list(ph.origin), "materials": [list(m) for m in ph.materials]}


def read_phantom(path: str) -> phantoms.LabelPhantom:
    labels, meta = rawio.read_raw(path, "phantom")
    if not meta:
        raise SystemExit(f"{path}: the phantom needs its sidecar (voxel size and materials)")
    return phantoms.LabelPhantom(labels=labels, voxel=tuple(meta["voxel"]), origin=tuple(meta["origin"]),
                                 materials=[tuple(m) for m in meta["materials"]])


def cmd_phantom(a) -> int:
    ph = PHANTOMS[a.preset]()
    rawio.write_raw(a.out, ph.labels, "phantom", _phantom_meta(ph))
    print(f"wrote {a.out}: {ph.labels.shape} labels, voxel {ph.voxel} cm")
    return 0


def cmd_mc(a) -> int:
    import torch

    from .context import Context
    from .mc import McSettings, counts_to_lineint, device_labels, mc_project, photons_per_pixel

    geom = _geometry(a.geometry, a.views)
    ph = read_phantom(a.phantom)
    spec = None
    if a.spectrum:
        kvp, al = (float(x) for x in a.spectrum.split(","))
        spec = tables.kramers_spectrum(kvp, al)
    s = McSettings(photons_per_view=int(float(a.photons)), e_kev=a.e_kev, spectrum=spec, coherent=a.coherent,
                   seed=a.seed, max_interactions=a.max_interactions)
    ctx = Context(0)
    prim, scat, st = mc_project(ph, geom, s, labels_dev=device_labels(ph, torch.device("cuda", 0)), ctx=ctx)
    meta = {"geometry": geom.as_dict(), "photons_per_view": s.photons_per_view, "seed": s.seed,
            "stats": {k: v for k, v in st.items() if k != "kernel_ms"}}
    # CBCT_real325.cu:264-294 clamps the counts to [1, N0] and writes them with
    # the line integrals ln N0 - ln N; image0 is primary, image5 primary + scatter
    n0 = torch.from_numpy(photons_per_pixel(geom, s.photons_per_view)).to(prim.device)
    for tag, counts in (("0", prim), ("5", prim + scat)):
        lineint = counts_to_lineint(counts.contiguous(), geom, s.photons_per_view, ctx)
        clamped = torch.minimum(torch.clamp(counts, min=1.0), n0)
        rawio.write_raw(f"{a.out}_proj{tag}.raw", clamped.round().to(torch.int32).cpu().numpy(), "counts", meta)
        rawio.write_raw(f"{a.out}_map{tag}.raw", lineint.cpu().numpy(), "lineint", meta)
    print(f"wrote {a.out}_proj0/5.raw and {a.out}_map0/5.raw: {geom.proj_shape}, "
          f"{st['histories']} histories in {st['kernel_ms']:.1f} ms")
    return 0


def cmd_fdk(a) -> int:
    import torch

    from .context import Context
    from .fdk import fdk, fdk_filter

    geom = _geometry(a.geometry, a.views)
    m, _ = rawio.read_raw(a.map, "lineint", shape=geom.proj_shape)
    if tuple(m.shape) != geom.proj_shape:
        raise SystemExit(f"{a.map}: shape {m.shape} does not match geometry {geom.proj_shape}")
    ctx = Context(0)
    proj = torch.from_numpy(np.ascontiguousarray(m, dtype=np.float32)).cuda()
    vol = fdk(proj, geom, ctx).cpu().numpy()
    meta = {"geometry": geom.as_dict(), "source": a.map}
    rawio.write_raw(a.out, vol, "volume", meta)
    if a.zy:
        rawio.write_raw(a.zy, np.ascontiguousarray(vol.transpose(2, 1, 0)), "volume_zy", meta)
    if a.filtered:
        rawio.write_raw(a.filtered, fdk_filter(proj, geom, ctx).cpu().numpy(), "filtered", meta)
    print(f"wrote {a.out}: {vol.shape} volume")
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m monte_amd.cli")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("phantom", help="write a label phantom")
    p.add_argument("--preset", choices=sorted(PHANTOMS), default="full")
    p.add_argument("--out", required=True)
    p = sub.add_parser("mc", help="phantom .raw -> proj/map .raw (Monte Carlo)")
    p.add_argument("--phantom", required=True)
    p.add_argument("--geometry", default="full")
    p.add_argument("--views", type=int)
    p.add_argument("--photons", default="1e6", help="photons per view")
    p.add_argument("--e-kev", type=float, default=60.0)
    p.add_argument("--spectrum", help="KVP,AL_MM: Kramers spectrum instead of mono energy")
    p.add_argument("--coherent", type=int, default=1, help="0 forward (reference), 1 Rayleigh")
    p.add_argument("--max-interactions", type=int, default=5)
    p.add_argument("--seed", type=int, default=20261202)
    p.add_argument("--out", required=True, help="output prefix")
    p = sub.add_parser("fdk", help="map .raw -> volume .raw (FDK)")
    p.add_argument("--map", required=True)
    p.add_argument("--geometry", default="full")
    p.add_argument("--views", type=int)
    p.add_argument("--out", required=True)
    p.add_argument("--zy")
    p.add_argument("--filtered")
    a = ap.parse_args(argv)
    return {"phantom": cmd_phantom, "mc": cmd_mc, "fdk": cmd_fdk}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
