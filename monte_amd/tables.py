"""Attenuation tables and source spectra.

* Water and calcium mass attenuation: the reference's own XCOM tables
  (monte_cpp/xcom2.csv, monte_cpp/Ca.csv), packaged under monte_amd/data and
  parsed by the library's monte_read_xcom_csv (restating readcsv,
  monte_cpp/CBCT_real2.cpp:633-668).
* Aluminium (filtration only): NIST X-ray mass attenuation coefficients
  (Hubbell & Seltzer, NISTIR 5632), log-log interpolated.  Not in the reference.
* Kramers bremsstrahlung photon spectrum N(E) ~ (kVp - E) / E on 1 keV bins,
  attenuated by the filtration.  The reference reads a measured 125 kV
  spectrum from files it does not ship (monte_cu/CBCT_real325.cu:349-389), so
  the spectrum here is generated.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
ROWS = _lib.TABLE_ROWS

WATER, CALCIUM = 0, 1  # table indices of reference_tables()
COH, INCOH, PHOTO, TOTAL = 0, 1, 2, 3

# NIST (Hubbell & Seltzer) mass attenuation of aluminium, total with coherent:
# (keV, cm^2/g)
AL_MU_RHO = (
    (5.0, 193.4), (6.0, 115.3), (8.0, 50.33), (10.0, 26.23), (15.0, 7.955), (20.0, 3.441),
    (30.0, 1.128), (40.0, 0.5685), (50.0, 0.3681), (60.0, 0.2778), (80.0, 0.2018),
    (100.0, 0.1704), (150.0, 0.1378), (200.0, 0.1223),
)
AL_DENSITY = 2.699  # g/cm^3


def read_xcom_csv(path: str) -> np.ndarray:
    """Parse a 4-column XCOM CSV into [4][201] (row k = k keV) with the C ABI parser."""
    tab = np.zeros((4, ROWS), dtype=np.float64)
    rc = _lib.lib().monte_read_xcom_csv(path.encode(), tab.ctypes.data_as(C.c_void_p))
    if rc < 0:
        raise _lib.MonteError(rc, f"cannot read {path}")
    return tab


def reference_tables() -> np.ndarray:
    """[2][4][201]: water (WATER) and calcium (CALCIUM) from the reference's CSVs."""
    return np.stack([read_xcom_csv(os.path.join(DATA, "xcom2.csv")),
                     read_xcom_csv(os.path.join(DATA, "Ca.csv"))])


def al_mu_rho(e_kev) -> np.ndarray:
    e = np.asarray(e_kev, dtype=np.float64)
    xs = np.log([a for a, _ in AL_MU_RHO])
    ys = np.log([b for _, b in AL_MU_RHO])
    return np.exp(np.interp(np.log(np.clip(e, AL_MU_RHO[0][0], AL_MU_RHO[-1][0])), xs, ys))


def kramers_spectrum(kvp: float = 120.0, al_mm: float = 2.5) -> np.ndarray:
    """Photon weights per 1 keV bin, index = keV (0..200); bins k in [1, kVp)."""
    w = np.zeros(ROWS, dtype=np.float64)
    k = np.arange(1, ROWS, dtype=np.float64)
    n = np.where(k < kvp, (kvp - k) / k, 0.0)
    n *= np.exp(-al_mu_rho(k) * AL_DENSITY * (al_mm / 10.0))
    w[1:] = n
    return w / w.sum()


def spectrum_cdf_f32(weights: np.ndarray) -> np.ndarray:
    """The cumulative table the transport kernel samples, rounded to FP32 the
    same way the library rounds it (so a CPU replay samples identical bins)."""
    w = np.asarray(weights, dtype=np.float64)
    tot = w[1:].sum()
    cdf = np.zeros(ROWS, dtype=np.float64)
    cdf[1:] = np.cumsum(w[1:]) / tot
    cdf = cdf.astype(np.float32).astype(np.float64)
    cdf[-1] = 1.0
    return cdf


def linear_mu(tables: np.ndarray, table: int, density: float, e_kev: float) -> float:
    """mu (cm^-1) at the table row the transport uses, (int)(E + 0.5)."""
    return float(tables[table, TOTAL, int(e_kev + 0.5)] * density)
