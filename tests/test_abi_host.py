"""The C ABI (include/monte.h / libmonte.so) and host-side logic, no GPU needed.

* libmonte.so loads and exports every function the header declares;
* the ctypes mirror (monte_amd._lib) has the exact size and field offsets of
  each C struct (checked against a probe compiled with gcc from the header);
* host-only entry points: CSV parsing, phantom painting, status strings and
  their error returns.
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "monte.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    return sorted(set(re.findall(r"\b(monte_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from monte_amd import _lib

    names = _declared_functions()
    assert len(names) >= 20
    L = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes mirror binds exactly those
    assert sorted(_lib.SIGNATURES) == names


STRUCTS = {
    "monte_geom": "Geom", "monte_phantom": "Phantom", "monte_mc_cfg": "McCfg", "monte_mc_stats": "McStats",
    "monte_mc_job": "McJob", "monte_tuning": "Tuning", "monte_primitive": "Primitive",
    "monte_e2e_buffers": "E2EBuffers",
}


def test_struct_layouts_match_c():
    from monte_amd import _lib

    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in getattr(_lib, pyname)._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0; }")
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "probe.c")
        exe = os.path.join(tmp, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        st = getattr(_lib, STRUCTS[cname])
        if field == "size":
            assert C.sizeof(st) == int(val), cname
        else:
            assert getattr(st, field).offset == int(val), (cname, field)


def test_status_strings():
    from monte_amd import _lib

    L = _lib.lib()
    for code in (0, -1, -2, -3, -4, -5, -6):
        s = L.monte_status_str(code).decode()
        assert s and s != "unknown status"
    assert L.monte_status_str(-99).decode() == "unknown status"


def test_csv_errors():
    from monte_amd import _lib, tables

    with pytest.raises(_lib.MonteError) as e:
        tables.read_xcom_csv("/nonexistent/table.csv")
    assert e.value.code == _lib.ERR_IO
    with tempfile.NamedTemporaryFile("w", suffix=".csv", delete=False) as f:
        f.write("not,a,table\n")
        name = f.name
    try:
        with pytest.raises(_lib.MonteError):
            tables.read_xcom_csv(name)
    finally:
        os.unlink(name)
    assert _lib.lib().monte_read_xcom_csv(None, None) == _lib.ERR_INVALID


def _paint_numpy(shape, v, prims):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*[(np.arange(n) - 0.5 * (n - 1)) * v for n in shape], indexing="ij")
    lab = np.zeros(shape, dtype=np.uint8)
    for s, l, c, r in prims:
        a, b, cc = (x - c[0]) / r[0], (y - c[1]) / r[1], (z - c[2]) / r[2]
        if s == 0:
            m = a * a + b * b + cc * cc <= 1
        elif s == 1:
            m = (a * a + b * b <= 1) & (np.abs(cc) <= 1)
        elif s == 3:
            m = (b * b + cc * cc <= 1) & (np.abs(a) <= 1)
        else:
            m = (np.abs(a) <= 1) & (np.abs(b) <= 1) & (np.abs(cc) <= 1)
        lab[m] = l
    return lab


def test_phantom_paint_matches_numpy():
    from monte_amd import phantoms

    prims = [(0, 1, (0.1, -0.2, 0.05), (1.2, 0.9, 1.1)), (1, 2, (0.4, 0.3, 0.0), (0.3, 0.25, 0.8)),
             (2, 3, (-0.5, 0.2, -0.3), (0.2, 0.35, 0.15)), (3, 4, (0.2, -0.6, 0.5), (0.8, 0.3, 0.45))]
    shape = (23, 29, 31)
    got = phantoms.paint(shape, (0.1, 0.1, 0.1), prims)
    ref = _paint_numpy(shape, 0.1, prims)
    # voxels exactly on a surface may round either way in the two evaluations
    assert np.mean(got != ref) < 1e-3
    assert np.bincount(got.ravel(), minlength=5)[1:].min() > 0


def test_reference_rods_phantom():
    """The label map of the reference's analytic medium (CBCT_real325.cu:916-928),
    checked voxel by voxel against the reference's membership tests evaluated
    at the voxel centres with the same 45-degree rod placement."""
    from monte_amd import phantoms

    ph = phantoms.reference_rods_phantom(v=0.25)
    n = ph.labels.shape[0]
    assert ph.labels.shape == (80, 80, 80) and ph.materials[2][1] == 1.55
    c = (np.arange(n) - 0.5 * (n - 1)) * 0.25
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    water = (y * y + z * z <= 100) & (np.abs(x) <= 10)
    rods = np.zeros_like(water)
    for yc, zc in ((5, 0), (-5, 0), (0, 5), (0, -5)):
        rods |= (y - yc) ** 2 + (z - zc) ** 2 <= 1.5 ** 2
    for ang in (np.pi / 4, 3 * np.pi / 4):
        for sgn in (1, -1):
            rods |= (y - sgn * 5 * np.sin(ang)) ** 2 + (z - sgn * 5 * np.cos(ang)) ** 2 <= 1.5 ** 2
    ref = np.where(water, np.where(rods, 2, 1), 0)
    assert np.mean(ph.labels != ref) < 1e-3
    assert (ph.labels == 2).sum() > 0.9 * rods[water].sum()


def test_phantom_paint_rejects_bad_input():
    from monte_amd import _lib

    L = _lib.lib()
    buf = np.zeros(8, dtype=np.uint8)
    assert L.monte_phantom_paint(buf.ctypes.data_as(C.c_void_p), 2, 2, 2, 0.1, 0.1, 0.1, None, 0) == 0
    prim = _lib.Primitive(shape=7, label=1, rx=1, ry=1, rz=1)
    assert L.monte_phantom_paint(buf.ctypes.data_as(C.c_void_p), 2, 2, 2, 0.1, 0.1, 0.1, C.byref(prim), 1) == \
        _lib.ERR_INVALID
    assert L.monte_phantom_paint(None, 2, 2, 2, 0.1, 0.1, 0.1, None, 0) == _lib.ERR_INVALID


def test_workload_phantoms():
    from monte_amd import phantoms

    ph = phantoms.small_phantom()
    assert ph.labels.shape == (65, 65, 65)
    counts = np.bincount(ph.labels[32].ravel(), minlength=3)
    # slice through the water cylinder (r = 25 voxels) and two Ca rods (r = 4)
    assert abs(counts[1] + counts[2] - np.pi * 25 ** 2) < 0.03 * np.pi * 25 ** 2
    assert abs(counts[2] - 2 * np.pi * 16) < 0.25 * 2 * np.pi * 16
    fp = phantoms.full_phantom(seed=1, n=65, v=0.25)
    assert fp.labels.max() == 2 and (fp.labels == 1).sum() > 0
    a = phantoms.full_phantom(seed=5, n=41, v=0.4).labels
    b = phantoms.full_phantom(seed=5, n=41, v=0.4).labels
    assert np.array_equal(a, b)


def test_geometry_presets():
    from monte_amd import geometry as G

    g = G.bp3d20_325()
    c = g.c()
    assert (c.nu, c.nv, c.sod, c.sdd, c.du, c.roi_x0, c.roi_x1, c.scale) == (325, 325, 160.0, 220.0, 0.1, 125, 130, 13.5)
    f = G.full()
    # the magnified field of view covers the whole cube at every view
    half_diag = 0.5 * f.nx * f.vx * np.sqrt(2)
    assert f.sdd * half_diag / (f.sod - half_diag) < 0.5 * f.nu * f.du
    assert abs(f.c().dangle - 2 * np.pi / 360) < 1e-15


def test_create_without_gpu_reports_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from monte_amd import _lib

    h = C.c_void_p()
    assert _lib.lib().monte_create(C.byref(h), 0) < 0
    assert not h.value


def test_one_hip_runtime_with_torch():
    """Importing monte_amd (even before torch) leaves one HIP runtime in the
    process: libmonte.so binds to the libamdhip64 that torch loaded."""
    import subprocess
    import sys

    code = ("import re, monte_amd, torch\n"
            "maps = open('/proc/self/maps').read()\n"
            "print(len(set(re.findall(r'\\S*libamdhip64\\S*', maps))))\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout + out.stderr
