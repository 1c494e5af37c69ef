"""The batched uneven MC workload at its stated sizes: 64 independent jobs in
one persistent launch (cubic phantoms of side U[65, 325] voxels, 1e5..1e8
photons per job, mono 40..120 keV, 64..256^2 detectors, 1..36 views).

* Every job's tallies equal, bitwise, the same job run alone: a job's
  histories draw from its own Philox stream id whatever else shares the launch,
  and every tally adds 1.0.
* Per-job statistics equal too, and every history ends exactly once.
* The smallest jobs match the FP64 oracle on the same streams (per-pixel
  |z| < 3, >= 95% identical pixels).

The reference runs one job per program (monte_cu/CBCT_real325.cu:84-303), its
per-pixel serial loop (:450-452) being what the work queue replaces.
"""
import numpy as np
import pytest

from conftest import oracle_geom, oracle_phantom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_run(ctx):
    import torch

    from monte_amd import workloads
    from monte_amd.mc import device_labels, mc_batch

    jobs = workloads.batched_jobs()
    for j in jobs:
        j.labels_dev = device_labels(j.phantom, "cuda")
    hist = sum(j.settings.photons_per_view * j.geom.n_views for j in jobs)
    print(f"64 jobs, {hist:.3e} histories, sides {min(j.geom.nx for j in jobs)}..{max(j.geom.nx for j in jobs)}, "
          f"detectors {min(j.geom.nu for j in jobs)}..{max(j.geom.nu for j in jobs)}", flush=True)
    stats = mc_batch(jobs, ctx=ctx)
    torch.cuda.synchronize()
    return jobs, stats


def test_batch_covers_the_stated_ranges(batch_run):
    jobs, stats = batch_run
    assert len(jobs) == 64
    sides = [j.geom.nx for j in jobs]
    dets = [j.geom.nu for j in jobs]
    views = [j.geom.n_views for j in jobs]
    tot = [j.settings.photons_per_view * j.geom.n_views for j in jobs]
    es = [j.settings.e_kev for j in jobs]
    assert 65 <= min(sides) and max(sides) <= 325 and max(sides) - min(sides) > 150
    assert 64 <= min(dets) and max(dets) <= 256
    assert 1 <= min(views) and max(views) <= 36
    assert min(tot) >= 1e5 and max(tot) <= 1e8 + 36 and max(tot) / min(tot) > 100
    assert 40 <= min(es) and max(es) <= 120
    assert len({j.settings.job for j in jobs}) == 64


def test_batch_equals_single_runs(ctx, batch_run):
    import torch

    from monte_amd.mc import mc_project

    jobs, stats = batch_run
    for i, (j, st) in enumerate(zip(jobs, stats)):
        p, s, st1 = mc_project(j.phantom, j.geom, j.settings, labels_dev=j.labels_dev, ctx=ctx)
        torch.cuda.synchronize()
        assert torch.equal(p, j.primary) and torch.equal(s, j.scatter), i
        for k in st:
            if k != "kernel_ms":
                assert st[k] == st1[k], (i, k)
        assert st["histories"] == j.settings.photons_per_view * j.geom.n_views
        assert st["histories"] == st["tallied_primary"] + st["tallied_scatter"] + st["missed"] + st["photo"] + \
            st["capped"]
        assert float(p.sum()) == st["tallied_primary"] and float(s.sum()) == st["tallied_scatter"]


def test_smallest_jobs_match_oracle(ctx, oracle, batch_run):
    from monte_amd import tables

    jobs, stats = batch_run
    order = sorted(range(len(jobs)), key=lambda i: jobs[i].settings.photons_per_view * jobs[i].geom.n_views)
    tb = oracle.tables(tables.reference_tables())
    for i in order[:3]:
        j = jobs[i]
        s = j.settings
        p_ref, s_ref, st_ref = oracle.mc_run(oracle_geom(oracle, j.geom), oracle_phantom(oracle, j.phantom), tb,
                                             s.photons_per_view, e_kev=s.e_kev, max_interactions=s.max_interactions,
                                             coherent_mode=s.coherent, seed=s.seed, job=s.job,
                                             ray_brick=ctx.tuning["mc_ray_brick"])
        st = stats[i]
        print(f"job {i}: side {j.geom.nx}, detector {j.geom.nu}, views {j.geom.n_views}, "
              f"{st['histories']} histories, {s.e_kev:.1f} keV")
        assert st["histories"] == st_ref["histories"]
        for a, b in ((j.primary.cpu().numpy(), p_ref), (j.scatter.cpu().numpy(), s_ref)):
            same = np.mean(a == b)
            z = (a - b) / np.sqrt(np.maximum(a + b, 1.0))
            print(f"  identical pixels {same:.4f}  max|z| {np.abs(z).max():.2f}")
            assert np.abs(z).max() < 3.0 and same > 0.95
        for k in ("tallied_primary", "tallied_scatter", "photo", "compton", "coherent"):
            assert abs(st[k] - st_ref[k]) <= max(10, 1e-3 * st_ref[k]), (i, k)
