"""Pins for the oracle's RNG and the attenuation tables (CPU only)."""
import hashlib
import json
import os

import numpy as np

from conftest import ROOT

DATA = os.path.join(ROOT, "monte_amd", "data")


def test_philox_known_answers(oracle):
    """Philox4x32-10 against vectors from rocRAND's header (make_philox_kat.py);
    the first three are also the published Random123 known-answer vectors."""
    doc = json.load(open(os.path.join(ROOT, "tests", "golden", "philox_kat.json")))
    assert len(doc["vectors"]) >= 10
    for v in doc["vectors"]:
        assert oracle.philox(v["ctr"], v["key"]).tolist() == v["out"]
    assert doc["vectors"][0]["out"] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


def test_uniform_mapping(oracle):
    assert oracle.u01(0) == 0.5 / 2 ** 24
    assert oracle.u01(0xFFFFFFFF) == (2 ** 24 - 0.5) / 2 ** 24
    assert 0.0 < oracle.u01(0) and oracle.u01(0xFFFFFFFF) < 1.0
    # equidistribution of the 24-bit mapping over a Philox stream
    xs = np.array([oracle.philox([i, 1, 2, 3], [5, 6])[j] for i in range(4000) for j in range(4)], dtype=np.uint64)
    u = ((xs >> 8).astype(np.float64) + 0.5) / 2 ** 24
    h, _ = np.histogram(u, bins=20, range=(0, 1))
    chi2 = np.sum((h - len(u) / 20) ** 2 / (len(u) / 20))
    assert chi2 < 45.0  # 19 dof, p ~ 1e-3


def test_table_files_are_the_reference_files():
    digests = {
        "xcom2.csv": "89f2e8d8cde65a32349a2792b6a3583b5ebaea42454a06cf71ad9e55a3a57f4e",
        "Ca.csv": "9444bd376f22e6a71ad53c617d6cef08e807b44d7227c2e874d4f4ec62f48d05",
    }
    for name, h in digests.items():
        assert hashlib.sha256(open(os.path.join(DATA, name), "rb").read()).hexdigest() == h


def test_table_parse_semantics(oracle, ref_tables):
    """readcsv (monte_cpp/CBCT_real2.cpp:633-668): row k = k keV, 200 rows, BOM and CR tolerated."""
    for f in ("xcom2.csv", "Ca.csv"):
        tab, n = oracle.read_xcom_csv(os.path.join(DATA, f))
        assert n == 200
        first = open(os.path.join(DATA, f), "rb").read().split(b"\r\n")[0].decode("utf-8-sig").split(",")
        assert tab[0, 1] == float(first[0])  # the BOM row parses (the original hard-coded 1.372 here)
        assert tab[3, 1] == float(first[3])
    w, c = ref_tables
    assert w[3, 60] == 0.20585 and w[2, 60] == 0.0149 and w[0, 60] == 0.0139 and w[1, 60] == 0.177


def test_table_physics_invariants(ref_tables):
    for tab in ref_tables:
        tot = tab[3, 1:]
        parts = tab[0, 1:] + tab[1, 1:] + tab[2, 1:]
        assert np.max(np.abs(parts - tot) / tot) < 0.004  # components sum to the total (XCOM rounding)
        assert np.all(tot > 0)
    ca = ref_tables[1]
    assert ca[2, 5] / ca[2, 4] > 4.5  # calcium K edge (4.04 keV) between the 4 and 5 keV rows
    # above the K edge the photoelectric cross-section falls monotonically
    assert np.all(np.diff(ca[2, 5:]) < 0)


def test_library_parser_matches_oracle(oracle):
    """monte_read_xcom_csv (C ABI, host only) == the oracle's parser."""
    from monte_amd import tables

    for f in ("xcom2.csv", "Ca.csv"):
        a = tables.read_xcom_csv(os.path.join(DATA, f))
        b, _ = oracle.read_xcom_csv(os.path.join(DATA, f))
        assert np.array_equal(a, b)


def test_spectrum(ref_tables):
    from monte_amd import tables

    w = tables.kramers_spectrum(120.0, 2.5)
    assert abs(w.sum() - 1.0) < 1e-12
    assert w[0] == 0 and np.all(w[120:] == 0) and np.all(w[1:120] >= 0)
    e = np.arange(201)
    mean = float((w * e).sum())
    assert 50.0 < mean < 65.0  # 120 kVp with 2.5 mm Al: mean photon energy ~ 55-60 keV
    # NIST aluminium anchor values are reproduced exactly at the nodes
    assert abs(tables.al_mu_rho(60.0) - 0.2778) < 1e-12
    cdf = tables.spectrum_cdf_f32(w)
    assert cdf[-1] == 1.0 and np.all(np.diff(cdf) >= 0)
