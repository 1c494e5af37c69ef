"""Pins for the oracle's Siddon traversal and Monte Carlo physics (CPU only).

Siddon: chord lengths through boxes (closed form), total length conservation,
brute-force sampling on a tiny label grid.
MC sampling: Kahn/Klein-Nishina against the closed-form cross-section
(chi-square), the Thomson inverse CDF against its closed form, Beer-Lambert
transmission of a homogeneous slab (binomial), interaction branching ratios
(multinomial), and the Compton energy-angle relation.
"""
import numpy as np
import pytest
from scipy import stats as sps


def _phantom(oracle, labels, v, materials):
    return oracle.phantom(labels, v, materials=materials)


def test_siddon_box_chord(oracle):
    labels = np.ones((7, 9, 11), dtype=np.uint8)
    ph = _phantom(oracle, labels, 0.3, [(-1, 0.0), (0, 1.0)])
    rng = np.random.default_rng(0)
    for _ in range(50):
        p0 = rng.uniform(-10, 10, 3)
        p0[0] = -15.0
        p1 = rng.uniform(-2, 2, 3)
        p1[0] = 15.0
        L = oracle.siddon_lengths(ph, p0, p1, 4)
        # closed-form chord through the box [-1.65,1.65]x[-1.35,1.35]x[-1.05,1.05]
        d = p1 - p0
        b = np.array([11, 9, 7]) * 0.3 / 2
        with np.errstate(divide="ignore"):
            t0 = np.max(np.minimum((-b - p0) / d, (b - p0) / d))
            t1 = np.min(np.maximum((-b - p0) / d, (b - p0) / d))
        chord = max(0.0, min(t1, 1) - max(t0, 0)) * np.linalg.norm(d)
        assert abs(L[1] - chord) < 1e-12
        assert L[0] == 0


def test_siddon_brute_force(oracle):
    rng = np.random.default_rng(1)
    labels = rng.integers(0, 4, size=(4, 5, 6)).astype(np.uint8)
    v = 0.5
    ph = _phantom(oracle, labels, v, [(-1, 0.0), (0, 1.0), (0, 1.0), (0, 1.0)])
    b = np.array([6, 5, 4]) * v / 2
    for _ in range(20):
        p0 = rng.uniform(-1, 1, 3) * 8
        p1 = -p0 + rng.uniform(-0.5, 0.5, 3)
        L = oracle.siddon_lengths(ph, p0, p1, 4)
        n = 200000
        t = (np.arange(n) + 0.5) / n
        P = p0[None] + t[:, None] * (p1 - p0)[None]
        inside = np.all(np.abs(P) < b, axis=1)
        idx = np.floor((P[inside] + b) / v).astype(int)
        lab = labels[idx[:, 2], idx[:, 1], idx[:, 0]]
        seg = np.linalg.norm(p1 - p0) / n
        bf = np.bincount(lab, minlength=4) * seg
        assert np.allclose(L, bf, atol=4 * seg)
        assert abs(L.sum() - inside.sum() * seg) < 2 * seg


def kn_pdf_cos(mu, e_kev):
    """Klein-Nishina d sigma / d cos(theta) (unnormalised), k = E / 511."""
    k = e_kev / 511.0
    r = 1.0 / (1.0 + k * (1.0 - mu))
    return r * r * (r + 1.0 / r - (1.0 - mu * mu))


@pytest.mark.parametrize("e_kev", [30.0, 60.0, 140.0])
def test_kahn_samples_klein_nishina(oracle, e_kev):
    rng = np.random.default_rng(int(e_kev))
    n = 40000
    u = rng.random(n * 40)
    mus = np.empty(n)
    pos = 0
    for i in range(n):
        r, used = oracle.kahn_ratio(e_kev, u[pos:pos + 300])
        pos += used
        mus[i] = 1.0 - (r - 1.0) * (511.0 / e_kev)  # cos = 1 - (lambda' - lambda)
    assert np.all(mus >= -1 - 1e-12) and np.all(mus <= 1 + 1e-12)
    edges = np.linspace(-1, 1, 41)
    h, _ = np.histogram(mus, edges)
    fine = np.linspace(-1, 1, 40001)
    pdf = kn_pdf_cos(fine, e_kev)
    cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2 * np.diff(fine))])
    p = np.diff(np.interp(edges, fine, cdf)) / cdf[-1]
    chi2 = np.sum((h - n * p) ** 2 / (n * p))
    assert sps.chi2.sf(chi2, 39) > 1e-3, chi2


def test_thomson_inverse_cdf(oracle):
    u = np.linspace(1e-6, 1 - 1e-6, 2001)
    mu = np.array([oracle.thomson_mu(x) for x in u])
    assert np.allclose((mu ** 3 + 3 * mu + 4) / 8, u, atol=1e-12)  # F(mu) = (mu^3 + 3 mu + 4)/8
    assert oracle.thomson_mu(0.5) == 0.0
    assert abs(oracle.thomson_mu(1 - 1e-15) - 1.0) < 1e-9


def _slab_geom(oracle):
    return oracle.geom(n_views=1, nu=8, nv=8, nx=8, ny=8, nz=8, sod=30.0, sdd=60.0, du=0.02, dv=0.02)


def test_slab_transmission(oracle, ref_tables):
    """Primary transmission through a homogeneous box is Binomial(N, exp(-mu L))."""
    g = _slab_geom(oracle)
    labels = np.ones((8, 8, 8), dtype=np.uint8)
    v = 0.5
    ph = _phantom(oracle, labels, v, [(-1, 0.0), (0, 1.0)])
    tb = oracle.tables(ref_tables)
    N = 64 * 5000
    prim, scat, st = oracle.mc_run(g, ph, tb, N, e_kev=60.0, max_interactions=0, seed=3)
    L = oracle.primary_lineint(g, ph, tb, 60.0)
    mu = ref_tables[0, 3, 60]
    # the central rays cross 4 cm of water
    assert abs(L[0, 3, 3] - mu * 4.0 * np.sqrt(1 + 2 * (0.01 / 60) ** 2)) < 1e-3
    p = np.exp(-L[0])
    z = (prim[0] - 5000 * p) / np.sqrt(5000 * p * (1 - p))
    assert np.abs(z).max() < 4.5
    assert abs(np.mean(z)) < 4.5 / 8
    assert scat.sum() == 0  # max_interactions = 0: every collision ends the history
    assert st["capped"] + st["tallied_primary"] == N


def test_branching_ratios(oracle, ref_tables):
    """Interaction types follow photo/total, coherent/total, rest (multinomial)."""
    g = _slab_geom(oracle)
    labels = np.ones((8, 8, 8), dtype=np.uint8)
    ph = _phantom(oracle, labels, 0.5, [(-1, 0.0), (1, 1.55)])  # calcium
    tb = oracle.tables(ref_tables)
    _, _, st = oracle.mc_run(g, ph, tb, 64 * 4000, e_kev=40.0, max_interactions=1, seed=4)
    first = st["photo"] + st["coherent"] + st["compton"]
    assert first == st["interactions"]
    t = ref_tables[1]
    fa = np.float32(t[2, 40] / t[3, 40])
    fc = np.float32((t[2, 40] + t[0, 40]) / t[3, 40]) - fa
    # max_interactions = 1: a second collision ends the history unsampled, so every
    # sampled interaction is a first one at 40 keV and the counts are multinomial
    assert st["capped"] > 0
    n = first
    obs = np.array([st["photo"], st["coherent"], st["compton"]], dtype=float)
    exp = n * np.array([fa, fc, 1 - fa - fc])
    chi2 = float(np.sum((obs - exp) ** 2 / exp))
    assert sps.chi2.sf(chi2, 2) > 1e-3, (obs, exp)


def test_compton_energy_angle(oracle):
    """E' = E / (1 + (E/511)(1 - cos)) for every Kahn sample."""
    rng = np.random.default_rng(9)
    for e in (20.0, 80.0, 150.0):
        for _ in range(200):
            r, _ = oracle.kahn_ratio(e, rng.random(300))
            lam = 511.0 / e
            mu = 1.0 - (r * lam - lam)
            e_new = 511.0 / (r * lam)
            assert abs(e_new - e / (1 + (e / 511.0) * (1 - mu))) < 1e-9 * e
            assert 1.0 <= r <= 1.0 + 2.0 / lam + 1e-12
