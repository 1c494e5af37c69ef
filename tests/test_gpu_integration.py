"""INTEGRATION.md's C++ calls, compiled as programs and run against libmonte.so.

tests/integration/mc_host.cpp and fdk_host.cpp hold the calls a maintainer
would put into monte_cu/CBCT_real325.cu (in place of the projection<<<>>>
launch, :241-250) and recon/bp3d20_325.cpp (in place of the loops :39-175),
with host arrays; the document quotes them verbatim (checked on the CPU by
test_integration_doc.py).  Here they run on the GPU:

* mc_host on the label map of the reference's rod phantom in the reference
  geometry gives, bitwise, the tallies of the same call through the Python
  binding with device pointers;
* fdk_host on the synthetic sphere stack reproduces the reference program's
  own output (the golden made by compiling recon/bp3d20_325.cpp) at 1e-5.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2, rel_max

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "integration", "bin")


def _program(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: build with `make -C tests/integration` (part of __graft_entry__.build)")
    return path


def test_mc_host_program_matches_device_call(ctx, tmp_path):
    import torch

    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.mc import McSettings, mc_project

    rods = phantoms.reference_rods_phantom(0.1)
    n = rods.labels.shape[0]
    lab = tmp_path / "labels.raw"
    rods.labels.tofile(lab)
    views, step, per = 4, 45, 2
    print("running mc_host", flush=True)
    r = subprocess.run([_program("mc_host"), os.path.join(ROOT, "monte_amd", "data"), str(lab), str(n), str(views),
                        str(step), str(per), str(tmp_path / "out_")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    shape = (views, 325, 325)
    prim = np.fromfile(tmp_path / "out_primary.raw", dtype=np.float64).reshape(shape)
    scat = np.fromfile(tmp_path / "out_scatter.raw", dtype=np.float64).reshape(shape)
    geom = G.bp3d20_325(n_views=views).replace(step_deg=step)
    s = McSettings(photons_per_view=per * 325 * 325, e_kev=140.0, coherent=0, max_interactions=5, seed=20261202)
    p, sc, st = mc_project(rods, geom, s, ctx=ctx)
    torch.cuda.synchronize()
    assert f"histories {st['histories']} " in r.stdout
    assert np.array_equal(prim, p.cpu().numpy()) and np.array_equal(scat, sc.cpu().numpy())
    assert prim.sum() > 0 and scat.sum() > 0


def test_fdk_host_program_matches_reference_golden(tmp_path):
    from synth import bp3d20_sphere_stack

    gold = np.load(os.path.join(ROOT, "tests", "golden", "bp3d20_325.npz"))
    proj = bp3d20_sphere_stack(325, 0.1)
    inp = tmp_path / "map.raw"
    proj.tofile(inp)
    print("running fdk_host", flush=True)
    r = subprocess.run([_program("fdk_host"), str(inp), str(tmp_path / "xy.raw")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    vol = np.fromfile(tmp_path / "xy.raw", dtype=np.float32).reshape(256, 256, 256)
    got = vol[::4, ::4, 125:130]
    print(f"rel_max {rel_max(got, gold['vol_xy']):.3e} rel_l2 {rel_l2(got, gold['vol_xy']):.3e}")
    assert rel_max(got, gold["vol_xy"]) < 1e-5 and rel_l2(got, gold["vol_xy"]) < 1e-5
    outside = np.ones(256, bool)
    outside[125:130] = False
    assert np.all(vol[:, :, outside] == 0)
