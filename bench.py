"""Benchmark of the two hot paths on one MI355X.

    python bench.py [--steps K] [--warmup W] [--photons P] [--views V] [--cpu-seconds S]

Workload (full CBCT): 325^3 label phantom at 0.5 mm (water ellipsoid 300 x 250 x
300 voxels, 12 seeded Ca inserts r in [3,15] voxels), 512 x 512 detector at
1.2 mm, SOD 600 / SDD 1200 mm, 360 views, 120 kVp Kramers spectrum with
2.5 mm Al in 1 keV bins, 1e8 photons per view, FP32 tracking with FP64
tallies; FDK (Shepp-Logan) of the resulting line integrals to 325^3 FP32;
and a differentiable-FDK step (forward FDK + adjoint through
torch.autograd) on a batch of 8 stacks of the same geometry.

One step = MC projection of all views + counts->line integrals + FDK + the
autograd forward/backward.  Inputs (phantom, stacks) are resident in HBM
before timing.  Prints one JSON line; `value` is MC histories per second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_LANE_OPS = 256 * 64 * 2.4e9  # MI355X FP32 VALU: 256 CUs x 64 lanes x 2.4 GHz, one op per lane per clock
PEAK_HBM = 8.0e12
BP_OPS = 18       # FP32 VALU ops per backprojection voxel-update (algorithmic)
MC_OPS_STEP, MC_OPS_INT, MC_OPS_HIST = 40, 150, 60  # per Woodcock step, per interaction, per history


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8, help="photons per view")
    ap.add_argument("--views", type=int, default=360)
    ap.add_argument("--batch", type=int, default=8, help="stacks in the differentiable-FDK step")
    ap.add_argument("--cpu-seconds", type=float, default=24.0, help="CPU-baseline budget (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU-baseline threads (0 = every CPU this process may use, capped by OMP_NUM_THREADS)")
    ap.add_argument("--batched", type=int, default=1, help="1: also time the 64-job batched MC workload")
    ap.add_argument("--e2e", type=int, default=1, help="1: also time the end-to-end 512^3 / 720-view step")
    ap.add_argument("--e2e-photons", type=float, default=5e7, help="photons per view of the end-to-end step")
    ap.add_argument("--e2e-steps", type=int, default=1, help="timed graph replays of the end-to-end step")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: time candidate launch shapes (MC, backprojection, projector) first; the defaults are "
                         "this workload's best by earlier sweeps")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: every path here runs on a "
              "single GPU (MC would shard by views, FDK by z-slabs)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from monte_amd import Context, geometry, phantoms, tables
    from monte_amd.fdk import fdk, fdk_autograd
    from monte_amd.mc import McSettings, autotune, counts_to_lineint, device_labels, mc_project

    dev = torch.device("cuda", 0)
    ctx = Context(0)
    geom = geometry.full(n_views=a.views)
    ph = phantoms.full_phantom()
    labels = device_labels(ph, dev)
    spec = tables.kramers_spectrum(120.0, 2.5)
    N = int(a.photons)
    settings = McSettings(photons_per_view=N, spectrum=spec, coherent=1, seed=20261202)
    primary = torch.zeros(geom.proj_shape, dtype=torch.float64, device=dev)
    scatter = torch.zeros_like(primary)
    rng = np.random.default_rng(20261202)
    # differentiable-FDK input: noisy stacks (Poisson at 1e5 counts per pixel) from the exact primary
    from monte_amd.mc import primary_exact

    L = primary_exact(ph, geom, spectrum=spec, labels_dev=labels, ctx=ctx)
    stacks = []
    for b in range(a.batch):
        counts = torch.poisson(1e5 * torch.exp(-L.double()), generator=torch.Generator(device=dev).manual_seed(b))
        stacks.append((np.log(1e5) - torch.log(torch.clamp(counts, min=1.0))).float())
    stacks = torch.stack(stacks).contiguous()
    del L
    tuning_default = ctx.tuning
    if a.autotune:
        autotune(ph, geom, settings, what=7, labels_dev=labels, ctx=ctx)
    tuning = ctx.tuning
    torch.cuda.synchronize()

    def step():
        rec = {}
        primary.zero_()
        scatter.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _, _, st = mc_project(ph, geom, settings, labels_dev=labels, primary=primary, scatter=scatter, ctx=ctx)
        e1.record()
        lineint = counts_to_lineint(primary + scatter, geom, N, ctx)
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        f0.record()
        vol = fdk(lineint, geom, ctx)
        f1.record()
        g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        x = stacks.clone().requires_grad_(True)
        g0.record()
        v = fdk_autograd(x, geom, ctx)
        v.backward(v.detach())  # gradient of 0.5 ||v||^2
        g1.record()
        h0, h1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        h0.record()
        fdk(stacks, geom, ctx)
        h1.record()
        torch.cuda.synchronize()
        rec["mc_ms"] = e0.elapsed_time(e1)
        rec["mc_kernel_ms"] = st["kernel_ms"]
        rec["mc_stats"] = st
        rec["fdk_ms"] = f0.elapsed_time(f1)
        rec["fwd_adj_ms"] = g0.elapsed_time(g1)
        rec["fdk_batch_ms"] = h0.elapsed_time(h1)
        rec["vol_mean"] = float(vol.mean())
        return rec

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = [step() for _ in range(a.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    hist = N * a.views
    mc_s = np.mean([r["mc_ms"] for r in recs]) / 1e3
    mc_k = np.mean([r["mc_kernel_ms"] for r in recs]) / 1e3
    st = recs[-1]["mc_stats"]
    S_kernel = st["steps"] / st["histories"]
    # The work model counts Woodcock steps under the global majorant; the kernel
    # takes fewer (first flights use the majorant of the labels their line
    # crosses), so S is measured on a subsample run with the global majorant.
    S, Cn = global_majorant_steps(ctx, ph, geom, settings, labels)
    w_mc = MC_OPS_STEP * S + MC_OPS_INT * Cn + MC_OPS_HIST
    fdk_s = np.mean([r["fdk_ms"] for r in recs]) / 1e3
    updates = geom.nx * geom.ny * geom.nz * geom.n_views
    fwd_adj = np.mean([r["fwd_adj_ms"] for r in recs])
    fdk_batch = np.mean([r["fdk_batch_ms"] for r in recs])
    out = {
        "metric": "MC histories/s; FDK GUPS; fwd+adjoint ms at 325³, vs plain same GPU",
        "value": hist / mc_s,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 tracking, f64 tallies",
        "data": "synthetic (seeded phantom, generated 120 kVp spectrum)",
        "config": {"workload": "full CBCT 325^3 / 512^2 / %d views / %.0e photons per view / FDK Shepp-Logan; "
                               "diff-FDK batch %d" % (a.views, N, a.batch),
                   "phantom": "325^3 u8 at 0.5 mm", "detector": "512x512 at 1.2 mm", "views": a.views,
                   "photons_per_view": N, "spectrum": "120 kVp Kramers + 2.5 mm Al", "parallelism": "single GPU"},
        "fdk_gups": updates / fdk_s / 1e9,
        "fdk_ms": fdk_s * 1e3,
        "fwd_adjoint_ms": fwd_adj,
        "fdk_batch_ms": fdk_batch,
        "fwd_adjoint_over_fdk": fwd_adj / fdk_batch,
        "mc_steps_per_history": S,
        "mc_kernel_steps_per_history": S_kernel,
        "mc_interactions_per_history": Cn,
        "roofline": {
            "bound": "valu", "kernel": "mc_transport_kernel",
            "achieved": w_mc * hist / mc_k / 1e12, "peak": PEAK_LANE_OPS / 1e12, "unit": "Tlane-op/s",
            "frac": w_mc * hist / mc_k / PEAK_LANE_OPS, "traffic": profiled_traffic("mc_transport_kernel", a, N),
            "work_per_unit": "W = 40 S + 150 C + 60 lane-ops per history (S Woodcock steps under the global "
                             "majorant, C interactions)",
            "kernel_ms": mc_k * 1e3,
        },
        "roofline_fdk_backproject": {
            "bound": "valu", "achieved": BP_OPS * updates / fdk_s / 1e12, "peak": PEAK_LANE_OPS / 1e12,
            "unit": "Tlane-op/s", "frac": BP_OPS * updates / fdk_s / PEAK_LANE_OPS,
            "note": "whole FDK time (filter + backprojection) against the backprojection roof",
        },
        "tuning": tuning,
        "tuning_default": tuning_default,
        "autotuned": bool(a.autotune),
    }
    del primary, scatter, stacks, labels
    torch.cuda.empty_cache()
    if a.batched:
        out["batched_mc"] = bench_batched(a, ctx)
    if a.e2e:
        out["e2e"] = bench_e2e(a, ctx)
    if a.cpu_seconds > 0:
        a.ray_brick = tuning["mc_ray_brick"]  # the CPU port runs the same algorithm
        out["cpu_baseline"], out["cpu_baseline_fdk"] = cpu_baselines(a, geom, ph, spec, hist / mc_s, updates / fdk_s)
    print(json.dumps(out))


def global_majorant_steps(ctx, ph, geom, settings, labels, photons=2_000_000) -> tuple:
    """Woodcock steps S and interactions C per history with every flight on the
    global majorant, on `photons` per view of every view."""
    import dataclasses

    from monte_amd.mc import mc_project

    brick = ctx.tuning["mc_ray_brick"]
    ctx.set_tuning(mc_ray_brick=0)
    try:
        s = dataclasses.replace(settings, photons_per_view=min(photons, settings.photons_per_view))
        _, _, st = mc_project(ph, geom, s, labels_dev=labels, ctx=ctx)
    finally:
        ctx.set_tuning(mc_ray_brick=brick)
    return st["steps"] / st["histories"], st["interactions"] / st["histories"]


def mc_work(stats) -> tuple:
    """Steps S and interactions C per history, and W = 40 S + 150 C + 60."""
    hist = sum(s["histories"] for s in stats)
    S = sum(s["steps"] for s in stats) / hist
    Cn = sum(s["interactions"] for s in stats) / hist
    return S, Cn, MC_OPS_STEP * S + MC_OPS_INT * Cn + MC_OPS_HIST


def bench_batched(a, ctx) -> dict:
    """64 uneven MC jobs in one persistent launch (cubic phantoms of side
    U[65, 325] voxels, 1e5..1e8 photons per job, mono 40..120 keV, 64..256^2
    detectors, 1..36 views): the launch timed with its own events, against the
    same jobs launched one by one."""
    import numpy as np
    import torch

    from monte_amd import workloads
    from monte_amd.mc import device_labels, mc_batch, mc_project

    jobs = workloads.batched_jobs()
    for j in jobs:
        j.labels_dev = device_labels(j.phantom, "cuda")
    hist = sum(j.settings.photons_per_view * j.geom.n_views for j in jobs)

    def zero():
        for j in jobs:
            if j.primary is not None:
                j.primary.zero_()
                j.scatter.zero_()

    mc_batch(jobs, ctx=ctx)  # warm-up (allocates the tallies)
    times = []
    for _ in range(max(1, a.steps)):
        zero()
        stats = mc_batch(jobs, ctx=ctx)
        times.append(stats[0]["kernel_ms"])
    ms = float(np.mean(times))
    singles = 0.0
    for j in jobs:
        singles += mc_project(j.phantom, j.geom, j.settings, labels_dev=j.labels_dev, ctx=ctx)[2]["kernel_ms"]
    torch.cuda.synchronize()
    S_kernel = sum(s["steps"] for s in stats) / sum(s["histories"] for s in stats)
    brick = ctx.tuning["mc_ray_brick"]
    ctx.set_tuning(mc_ray_brick=0)  # the work model's steps: every flight on the global majorant
    try:
        zero()
        S, Cn, w = mc_work(mc_batch(jobs, ctx=ctx))
    finally:
        ctx.set_tuning(mc_ray_brick=brick)
    res = {"workload": "64 jobs: cubes of side U[65,325] at 0.5 mm, 1e5..1e8 photons per job, 40..120 keV mono, "
                       "64..256^2 detectors, 1..36 views (seed 20261202)",
           "jobs": len(jobs), "histories": hist, "kernel_ms": ms, "histories_per_s": hist / (ms / 1e3),
           "singles_kernel_ms": singles, "batch_speedup_over_singles": singles / ms,
           "mc_steps_per_history": S, "mc_kernel_steps_per_history": S_kernel, "mc_interactions_per_history": Cn,
           "roofline": {"bound": "valu", "kernel": "mc_transport_kernel", "unit": "Tlane-op/s",
                        "achieved": w * hist / (ms / 1e3) / 1e12, "peak": PEAK_LANE_OPS / 1e12,
                        "frac": w * hist / (ms / 1e3) / PEAK_LANE_OPS}}
    del jobs
    torch.cuda.empty_cache()
    return res


def bench_e2e(a, ctx) -> dict:
    """The end-to-end scan (512^3 phantom at 0.4 mm, 720 views of a 1024 x 768
    detector, 120 kVp spectrum, 5e7 photons per view): one captured step (MC
    with scatter, line integrals, FDK, L2 loss against the primary-only FDK,
    gradient through the adjoint) replayed and timed, plus the FDK and its
    adjoint timed alone on the step's own operands."""
    import torch

    from monte_amd import workloads
    from monte_amd.e2e import EndToEnd
    from monte_amd.fdk import fdk, fdk_adjoint
    from monte_amd.mc import McSettings, device_labels, primary_exact

    geom, ph, spec = workloads.e2e_inputs()
    labels = device_labels(ph, "cuda")
    N = int(a.e2e_photons)
    s = McSettings(photons_per_view=N, spectrum=spec, coherent=1, seed=20261202)
    L = primary_exact(ph, geom, spectrum=spec, labels_dev=labels, ctx=ctx)
    vol_ref = fdk(L, geom, ctx).contiguous()
    del L
    E = EndToEnd(ph, geom, s, vol_ref, labels_dev=labels, ctx=ctx)
    E.step(graph=False)
    E.step(graph=True)  # capture
    torch.cuda.synchronize()
    t = []
    for _ in range(max(1, a.e2e_steps)):
        st = E.step(graph=True)
        t.append(st["kernel_ms"])
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    fdk(E.lineint, geom, ctx)
    fdk_adjoint(E.residual, geom, ctx)
    ev[0].record()
    fdk(E.lineint, geom, ctx)
    ev[1].record()
    ev[2].record()
    fdk_adjoint(E.residual, geom, ctx)
    ev[3].record()
    torch.cuda.synchronize()
    fdk_ms, adj_ms = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    hist = N * geom.n_views
    step_ms = sum(t) / len(t)
    res = {"workload": "512^3 at 0.4 mm, 720 views, 1024x768 detector, 120 kVp, %.0e photons per view, coherent "
                       "Rayleigh; FDK Shepp-Logan 512^3; loss vs primary-only FDK; hipGraph replay" % N,
           "step_ms": step_ms, "histories": hist, "histories_per_s_in_step": hist / (step_ms / 1e3),
           "fdk_ms": fdk_ms, "adjoint_ms": adj_ms, "fdk_gups": geom.nx * geom.ny * geom.nz * geom.n_views / fdk_ms / 1e6,
           "fwd_adjoint_over_fdk": (fdk_ms + adj_ms) / fdk_ms, "loss": float(E.loss),
           "mc_kernel_steps_per_history": st["steps"] / st["histories"]}
    del E, vol_ref, labels
    torch.cuda.empty_cache()
    return res


def profiled_traffic(kernel, a, photons):
    """Memory traffic per launch (bytes) of `kernel` on this workload, from the
    FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/ (scaled by a
    copy of known size); None for any other workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not (os.path.exists(path) and a.views == 360 and photons == int(1e8)):
        return None
    with open(path) as f:
        rec = json.load(f).get(kernel)
    return rec["bytes_per_launch"] if rec else None


def host_cpus() -> dict:
    """The host CPUs this run may use: every CPU of this process's affinity
    set, capped by OMP_NUM_THREADS where the environment sets it (a GPU box
    gives each GPU a share of its host CPUs and says so there), with the
    machine's own count and CPU model for the record."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    share = affinity
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"share": share, "affinity": affinity, "machine": os.cpu_count(), "omp_num_threads": env or None,
            "model": model}


def cpu_baselines(a, geom, ph, spec, gpu_hps, gpu_ups):
    """Oracle (FP64 C restatement) on the host cores, on a bounded sample of
    about half the budget each for MC and FDK: a probe run, then up to two runs
    resized from the last one's time until it lands within 1.5x of the target."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    from monte_amd import tables

    host = host_cpus()
    nthreads = a.cpu_threads or host["share"]
    ray_brick = a.ray_brick
    og = O.geom(**geom.as_dict())
    oph = O.phantom(ph.labels, *ph.voxel, materials=ph.materials)
    tb = O.tables(tables.reference_tables())
    cdf = tables.spectrum_cdf_f32(spec)
    target = a.cpu_seconds / 2

    def mc_time(n, brick):
        t0 = time.perf_counter()
        O.mc_run(og, oph, tb, n, spectrum_cdf=cdf, coherent_mode=1, views=(0, 2), nthreads=nthreads,
                 ray_brick=brick)
        return time.perf_counter() - t0

    def mc_rate(brick, budget):
        # size the sample from a probe, then correct once if the probe (which pays
        # the fixed per-run costs) mispredicted by more than a factor of 1.5
        n = 2_000_000
        dt = mc_time(n, brick)
        for _ in range(2):
            if budget / 1.5 <= dt <= budget * 1.5:
                break
            n = max(n, int(n * budget / max(dt, 1e-3)))
            dt = mc_time(n, brick)
        return n, dt

    # The reference's CPU path: Woodcock steps under the global majorant
    # (monte_cu/CBCT_real325.cu:852 takes max(mu_water, mu_Ca) for every step).
    n, dt = mc_rate(0, target * 0.6)
    mc = {"value": 2 * n / dt, "unit": "histories/s", "cores": nthreads, "kind": "port", "host": host,
          "sample": f"oracle MC (FP64, same physics and Philox streams, the reference's global Woodcock "
                    f"majorant) on 2 views x {n} photons of the workload, {dt:.1f} s",
          "gpu_over_cpu": gpu_hps / (2 * n / dt)}
    # the same oracle with the GPU path's first-flight majorants, for the record
    n2, dt2 = mc_rate(ray_brick, target * 0.4)
    mc["same_majorants"] = {"value": 2 * n2 / dt2, "ray_brick": ray_brick,
                            "sample": f"2 views x {n2} photons, {dt2:.1f} s", "gpu_over_cpu": gpu_hps / (2 * n2 / dt2)}

    def fdk_time(nv):
        g = O.geom(**{**geom.as_dict(), "n_views": nv, "dangle": geom.c().dangle})
        proj = np.random.default_rng(0).standard_normal((nv, geom.nu, geom.nv)).astype(np.float32)
        t0 = time.perf_counter()
        O.fdk_std(g, proj, nthreads)
        return time.perf_counter() - t0

    nv = 2
    dt = fdk_time(nv)
    for _ in range(2):
        if target / 1.5 <= dt <= target * 1.5 or nv >= geom.n_views:
            break
        nv = int(min(geom.n_views, max(nv + 1, nv * target / max(dt, 1e-3))))
        dt = fdk_time(nv)
    ups = nv * geom.nx * geom.ny * geom.nz / dt
    fdk = {"value": ups / 1e9, "unit": "GUPS", "cores": nthreads, "kind": "port", "host": host,
           "sample": f"oracle FDK (FP64: weight, direct-convolution filter, backprojection) of {nv} views "
                     f"onto the full {geom.nx}^3 volume, {dt:.1f} s", "gpu_over_cpu": gpu_ups / ups}
    return mc, fdk


if __name__ == "__main__":
    main()
