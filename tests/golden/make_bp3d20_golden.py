"""Run the reference's own FDK programs on a synthetic input and keep a small golden.

The programs are recon/bp3d20_325.cpp and recon/bp3d20.cpp, compiled unmodified
from their sources by `make -C oracle ref` into oracle/_ref/ (this container
only: the reference checkout is not present elsewhere).  Each reads a fixed
input file name from its working directory and writes its outputs there, so
this script runs it in a scratch directory.

Input: tests/golden/synth.bp3d20_sphere_stack (line integrals of a sphere in
the reference geometry), regenerated identically by the tests.

Kept (tests/golden/bp3d20_<H>.npz):
  vol_xy   the reconstructed slab image_xy[z][t][s] at s in [125,130), every
           4th z and t  -> [64][64][5] float32
  filt     the filtered sinogram map_out[view][v][u] at views 0,90,180,270,
           every 5th v and u
  sha256_* digests of the program's full output files
Usage: python tests/golden/make_bp3d20_golden.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from synth import bp3d20_sphere_stack  # noqa: E402

CASES = {
    325: dict(binary="bp3d20_325", infile="mapg325_0_2tca.raw", pitch=0.1),
    65: dict(binary="bp3d20", infile="mapg_0_20000h2oCa.raw", pitch=0.5),
}
NOUT = 256
VIEWS_KEPT = (0, 90, 180, 270)


def run_case(H: int) -> dict:
    case = CASES[H]
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    exe = os.path.join(REPO, "oracle", "_ref", case["binary"])
    proj = bp3d20_sphere_stack(H, case["pitch"])
    with tempfile.TemporaryDirectory() as tmp:
        proj.tofile(os.path.join(tmp, case["infile"]))
        subprocess.run([exe], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        outs = {}
        for name in os.listdir(tmp):
            if name == case["infile"]:
                continue
            data = open(os.path.join(tmp, name), "rb").read()
            outs[name] = data
    vol_name = next(n for n in outs if "xy" in n and len(outs[n]) == 4 * NOUT ** 3)
    zy_name = next(n for n in outs if "zy" in n and len(outs[n]) == 4 * NOUT ** 3)
    map_name = next(n for n in outs if "map" in n and len(outs[n]) == 4 * 360 * H * H)
    xy = np.frombuffer(outs[vol_name], dtype=np.float32).reshape(NOUT, NOUT, NOUT)
    zy = np.frombuffer(outs[zy_name], dtype=np.float32).reshape(NOUT, NOUT, NOUT)
    filt = np.frombuffer(outs[map_name], dtype=np.float32).reshape(360, H, H)
    assert np.array_equal(xy[:, :, 125:130], np.transpose(zy[125:130], (2, 1, 0))), "xy/zy disagree"
    return dict(
        vol_xy=np.ascontiguousarray(xy[::4, ::4, 125:130]),
        filt=np.ascontiguousarray(filt[list(VIEWS_KEPT)][:, ::5, ::5]),
        views=np.array(VIEWS_KEPT),
        sha256_xy=hashlib.sha256(outs[vol_name]).hexdigest(),
        sha256_map=hashlib.sha256(outs[map_name]).hexdigest(),
        sha256_input=hashlib.sha256(proj.tobytes()).hexdigest(),
        nonzero_cols=np.array(sorted(set(np.nonzero(np.any(xy != 0, axis=(0, 1)))[0].tolist()))),
    )


def main():
    for H in sorted(CASES):
        res = run_case(H)
        path = os.path.join(HERE, f"bp3d20_{H}.npz")
        np.savez_compressed(path, **res)
        print(path, {k: (v.shape if hasattr(v, "shape") else v) for k, v in res.items()})


if __name__ == "__main__":
    main()
