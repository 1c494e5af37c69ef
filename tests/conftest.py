import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libmonte.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def ref_tables():
    """[2][4][201] water / calcium tables parsed by the oracle (FP64)."""
    import oracle as O

    O.build()
    w, _ = O.read_xcom_csv(os.path.join(ROOT, "monte_amd", "data", "xcom2.csv"))
    c, _ = O.read_xcom_csv(os.path.join(ROOT, "monte_amd", "data", "Ca.csv"))
    return np.stack([w, c])


@pytest.fixture(scope="session")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from monte_amd import Context

    c = Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


def oracle_geom(O, geom):
    """The oracle's Geom for a monte_amd Geometry (same fields)."""
    return O.geom(**geom.as_dict())


def oracle_phantom(O, ph):
    return O.phantom(ph.labels, *ph.voxel, materials=ph.materials, origin=ph.origin)


def rel_max(a, ref):
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))


def rel_l2(a, ref):
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))
