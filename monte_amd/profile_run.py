"""Fixed workloads for profiling under rocprofv3 (one launch set per run).

    python3 -m monte_amd.profile_run copy [--mib M]
    python3 -m monte_amd.profile_run mc [--views V] [--photons P]
    python3 -m monte_amd.profile_run fdk [--views V]
    python3 -m monte_amd.profile_run adjoint [--views V]

`copy` copies M MiB device-to-device once (a known byte count, used to scale
the FETCH_SIZE / WRITE_SIZE counters).  The others run one pass of a hot path
(`adjoint`: matched projector + transposed filter) on the full CBCT workload (325^3 phantom at 0.5 mm, 512^2 detector at 1.2 mm,
SOD 600 / SDD 1200 mm, 120 kVp spectrum), or with `--geom e2e` on the
end-to-end geometry (512^3 at 0.4 mm, 720 views of 1024 x 768), after one
untimed warm-up pass, and print what ran, so the kernel of interest appears in
the trace twice with the same arguments.
"""
from __future__ import annotations

import argparse
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=("copy", "mc", "fdk", "adjoint"))
    ap.add_argument("--views", type=int, default=360)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--geom", choices=("full", "e2e"), default="full",
                    help="geometry (and MC phantom): full (325^3, 512^2) or e2e (512^3 at 0.4 mm, 1024 x 768, 720 views)")
    a = ap.parse_args(argv)

    import torch

    from . import Context, geometry, phantoms, tables

    dev = torch.device("cuda", 0)
    if a.what == "copy":
        n = a.mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy: {a.mib} MiB read + {a.mib} MiB written", flush=True)
        return 0

    ctx = Context(0)
    geom = geometry.full(n_views=a.views) if a.geom == "full" else geometry.e2e()
    if a.what == "mc":
        from .mc import McSettings, device_labels, mc_project

        ph = phantoms.full_phantom() if a.geom == "full" else phantoms.full_phantom(n=512, v=0.04)
        labels = device_labels(ph, dev)
        s = McSettings(photons_per_view=int(a.photons), spectrum=tables.kramers_spectrum(120.0, 2.5), coherent=1,
                       seed=20261202)
        for rep in range(2):
            t = time.perf_counter()
            _, _, st = mc_project(ph, geom, s, labels_dev=labels, ctx=ctx)
            torch.cuda.synchronize()
            print(f"mc pass {rep}: {geom.n_views} views x {int(a.photons)} photons, kernel {st['kernel_ms']:.1f} ms, "
                  f"wall {(time.perf_counter() - t) * 1e3:.1f} ms", flush=True)
        return 0

    from .fdk import fdk, fdk_adjoint

    g = torch.Generator(device=dev).manual_seed(0)
    if a.what == "fdk":
        x = torch.randn(geom.proj_shape, device=dev, generator=g)
        fn = fdk
    else:
        x = torch.randn(geom.vol_shape, device=dev, generator=g)
        fn = fdk_adjoint
    for rep in range(2):
        t = time.perf_counter()
        fn(x, geom, ctx)
        torch.cuda.synchronize()
        print(f"{a.what} pass {rep}: {geom.n_views} views, wall {(time.perf_counter() - t) * 1e3:.1f} ms", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
