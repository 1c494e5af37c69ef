"""Pins for the oracle's post-interaction kinematics and source draw (CPU only).

* Direction update (monte_cu/CBCT_real325.cu:725-766 restated as a rotation of
  direction cosines): over random unit directions, the |dz| -> 1 branch
  included, the new direction is a unit vector at exactly the requested polar
  cosine to the old one; its azimuth about the old direction, measured in a
  frame built here by Gram-Schmidt (not the oracle's), advances by exactly the
  requested azimuth, so a uniform azimuth draw gives a uniform azimuth (checked
  by chi-square as well).  A dropped 1/r, a sign flipped in the z term or a
  swapped pair of cosines breaks the polar-cosine identity.
* Spectrum draw (CBCT_real325.cu:479-486): the keV bin of a uniform u is the
  first bin whose cumulative weight reaches u, i.e. numpy.searchsorted on the
  CDF, for draws at, between and beyond the CDF values.
* Labels past the table (16..255) read as label 15 in the exact primary and in
  transport, so a phantom holding label 200 gives the same results as the
  same phantom holding 15.
"""
import numpy as np
from scipy import stats as sps


def _frame(d):
    """An orthonormal right-handed pair (a, b) perpendicular to d, built from a
    fixed helper axis by Gram-Schmidt."""
    h = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    a = h - np.dot(h, d) * d
    a /= np.linalg.norm(a)
    b = np.cross(d, a)
    return a, b


def _unit(rng, n, near_pole=False):
    v = rng.standard_normal((n, 3))
    if near_pole:
        v[:, :2] *= 10.0 ** rng.uniform(-9, -3, (n, 1))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_rotate_dir_polar_cosine_and_norm(oracle):
    rng = np.random.default_rng(11)
    dirs = np.concatenate([_unit(rng, 3000), _unit(rng, 500, near_pole=True),
                           np.array([[0, 0, 1.0], [0, 0, -1.0], [1.0, 0, 0], [0, -1.0, 0]])])
    worst_dot = worst_norm = 0.0
    for d in dirs:
        for c in (rng.uniform(-1, 1), -1.0, 1.0, 0.0):
            phi = rng.uniform(0, 2 * np.pi)
            n = oracle.rotate_dir(d, c, phi)
            worst_norm = max(worst_norm, abs(np.linalg.norm(n) - 1.0))
            worst_dot = max(worst_dot, abs(np.dot(n, d) - c))
    print("max |dot - c|", worst_dot, "max |norm - 1|", worst_norm)
    assert worst_norm < 1e-12
    assert worst_dot < 1e-12


def test_rotate_dir_azimuth_advances_by_phi(oracle):
    rng = np.random.default_rng(12)
    for d in np.concatenate([_unit(rng, 40), _unit(rng, 10, near_pole=True)]):
        a, b = _frame(d)
        c = rng.uniform(-0.95, 0.95)
        phis = np.sort(rng.uniform(0, 2 * np.pi, 16))
        psi = []
        for phi in phis:
            n = oracle.rotate_dir(d, c, phi)
            psi.append(np.arctan2(np.dot(n, b), np.dot(n, a)))
        off = np.angle(np.exp(1j * (np.array(psi) - phis)))
        # one fixed offset for every phi: the azimuth turns with phi, same sense
        spread = np.abs(np.angle(np.exp(1j * (off - off[0])))).max()
        assert spread < 1e-9, (d, spread)


def test_rotate_dir_uniform_azimuth_chi2(oracle):
    rng = np.random.default_rng(13)
    d = _unit(rng, 1)[0]
    a, b = _frame(d)
    n_draw, bins = 36000, 36
    psi = np.empty(n_draw)
    for i, (c, phi) in enumerate(zip(rng.uniform(-1, 1, n_draw), rng.uniform(0, 2 * np.pi, n_draw))):
        n = oracle.rotate_dir(d, c, phi)
        psi[i] = np.arctan2(np.dot(n, b), np.dot(n, a))
    h, _ = np.histogram(psi, bins=bins, range=(-np.pi, np.pi))
    chi2, p = sps.chisquare(h)
    print("azimuth chi2", chi2, "p", p)
    assert p > 1e-4


def test_spectrum_bin_is_searchsorted(oracle):
    from monte_amd import tables

    cdf = tables.spectrum_cdf_f32(tables.kramers_spectrum(120.0, 2.5))
    rng = np.random.default_rng(14)
    us = np.concatenate([rng.uniform(0, 1, 20000), cdf[1:], np.nextafter(cdf[1:], 0), [0.0, 1e-12, 1.0]])
    got = np.array([oracle.spectrum_bin(cdf, u) for u in us])
    ref = np.minimum(np.searchsorted(cdf[1:], us, side="left") + 1, 200)
    assert np.array_equal(got, ref)
    # bins above the tube voltage never come out of a draw below 1
    assert got[us < 1.0].max() <= 120


def test_labels_past_table_read_as_last(oracle, ref_tables):
    from monte_amd import geometry as G

    rng = np.random.default_rng(15)
    labels = rng.choice([0, 1, 2, 15], size=(20, 22, 24)).astype(np.uint8)
    high = labels.copy()
    high[labels == 15] = rng.choice([15, 16, 200, 255], size=int((labels == 15).sum()))
    assert (high > 15).sum() > 1000
    mats = [(-1, 0.0), (0, 1.0), (1, 1.55)] + [(-1, 0.0)] * 12 + [(1, 0.8)]
    geom = G.small(n_views=3).replace(nu=32, nv=32, du=0.2, dv=0.2)
    og = oracle.geom(**geom.as_dict())
    tb = oracle.tables(ref_tables)
    ph_a = oracle.phantom(labels, 0.1, materials=mats)
    ph_b = oracle.phantom(high, 0.1, materials=mats)
    la = oracle.primary_lineint(og, ph_a, tb, 60.0)
    lb = oracle.primary_lineint(og, ph_b, tb, 60.0)
    assert la.max() > 1.0 and np.array_equal(la, lb)
    pa, sa, _ = oracle.mc_run(og, ph_a, tb, 20000, e_kev=60.0)
    pb, sb, _ = oracle.mc_run(og, ph_b, tb, 20000, e_kev=60.0)
    assert np.array_equal(pa, pb) and np.array_equal(sa, sb)
    # with fewer labels than 16, label 15 and everything above it are vacuum
    few = [(-1, 0.0), (0, 1.0), (1, 1.55)]
    lc = oracle.primary_lineint(og, oracle.phantom(high, 0.1, materials=few), tb, 60.0)
    ld = oracle.primary_lineint(og, oracle.phantom(np.where(labels == 15, 0, labels).astype(np.uint8), 0.1,
                                                   materials=few), tb, 60.0)
    assert np.array_equal(lc, ld)
