"""The first flight's Woodcock majorant on the GPU (tuning mc_ray_brick).

A photon's first flight runs along the line from the source to its pixel
centre; the largest mu among the labels that line crosses is a valid majorant
for it.  The label sets are formed per view and pixel in FP64 with the CPU
oracle's operations:

* GPU sets equal the oracle's bit for bit (config 1, the reference geometry
  with the rod phantom, an off-centre phantom with labels past the table and
  detector offsets), for several brick sizes;
* MC with the first-flight majorant and with the global one (independent
  seeds) sample the same physics: binned tallies agree within Poisson noise
  and interaction totals within 6 sigma, while the steps per history drop;
* at the full 325^3 / 512^2 workload the steps per history fall by more than
  half.
"""
import numpy as np
import pytest

from conftest import oracle_geom, oracle_phantom

pytestmark = pytest.mark.gpu


def _cases():
    from monte_amd import geometry as G
    from monte_amd import phantoms
    from monte_amd.phantoms import LabelPhantom

    rng = np.random.default_rng(31)
    lab = np.zeros((30, 34, 38), np.uint8)
    lab[5:25, 6:30, 4:34] = 1
    lab[10:16, 12:20, 8:14] = 2
    lab[18:22, 20:26, 20:30] = rng.choice([3, 16, 200], size=(4, 6, 10))
    odd = LabelPhantom(lab, (0.1, 0.12, 0.09), [(-1, 0.0), (0, 1.0), (1, 1.55), (1, 0.8)] + [(0, 0.5)] * 12,
                       (0.3, -0.2, 0.15))
    return [
        ("config1", phantoms.small_phantom(), G.small(n_views=36)),
        ("bp3d20_rods", phantoms.reference_rods_phantom(0.1), G.bp3d20_325(n_views=8).replace(step_deg=45)),
        ("offsets", odd, G.small(n_views=5).replace(nu=72, nv=40, du=0.15, dv=0.2, u_off=0.11, v_off=-0.3,
                                                  angle0=0.7)),
    ]


@pytest.mark.parametrize("brick", [2, 4, 8])
def test_sets_match_oracle_bitwise(ctx, oracle, ref_tables, brick):
    from monte_amd.mc import first_flight_sets

    tb = oracle.tables(ref_tables)
    for name, ph, geom in _cases():
        got = first_flight_sets(ph, geom, brick, ctx=ctx).cpu().numpy()
        ref = oracle.ray_masks(oracle_geom(oracle, geom), oracle_phantom(oracle, ph), tb, brick)
        diff = int((got != ref).sum())
        print(f"{name} brick {brick}: sets {got.size}, differing {diff}, values {np.unique(got).tolist()}")
        assert diff == 0, name


def test_first_flight_majorant_same_physics(ctx):
    from monte_amd import geometry as G
    from monte_amd import phantoms, tables
    from monte_amd.mc import McSettings, mc_project

    geom = G.small(n_views=8)
    ph = phantoms.small_phantom()
    spec = tables.kramers_spectrum(120.0, 2.5)
    brick = ctx.tuning["mc_ray_brick"]
    assert brick > 0
    ctx.set_tuning(mc_ray_brick=0)
    try:
        pa, sa, sta = mc_project(ph, geom, McSettings(photons_per_view=4_000_000, spectrum=spec, seed=1), ctx=ctx)
    finally:
        ctx.set_tuning(mc_ray_brick=brick)
    pb, sb, stb = mc_project(ph, geom, McSettings(photons_per_view=4_000_000, spectrum=spec, seed=2), ctx=ctx)
    print("steps per history", sta["steps"] / sta["histories"], "->", stb["steps"] / stb["histories"])
    assert stb["steps"] < 0.85 * sta["steps"]

    def bins(t, k):
        x = t.cpu().numpy()
        return x.reshape(x.shape[0], x.shape[1] // k, k, x.shape[2] // k, k).sum(axis=(2, 4))

    for name, a, c in (("primary", bins(pa, 8), bins(pb, 8)), ("scatter", bins(sa, 16), bins(sb, 16))):
        keep = a + c >= 20
        chi2 = float(np.mean((a - c)[keep] ** 2 / (a + c)[keep]))
        dof = int(keep.sum())
        tot = (a.sum() - c.sum()) / np.sqrt(a.sum() + c.sum())
        print(f"{name}: bins {dof}  chi2/dof {chi2:.4f}  total difference {tot:.2f} sigma")
        assert dof > 100 and abs(tot) < 5.0
        tol = 5 * np.sqrt(2.0 / dof)
        if name == "scatter":
            assert abs(chi2 - 1.0) < tol
        else:  # binomial bins: a + c over-estimates their variance
            assert chi2 < 1.0 + tol
    for k in ("tallied_primary", "tallied_scatter", "photo", "compton", "coherent"):
        assert abs(sta[k] - stb[k]) < 6 * np.sqrt(max(sta[k] + stb[k], 1)), k


def test_full_workload_steps_halved(ctx):
    from monte_amd import geometry as G
    from monte_amd import phantoms, tables
    from monte_amd.mc import McSettings, device_labels, mc_project

    geom = G.full(n_views=36)
    ph = phantoms.full_phantom()
    labels = device_labels(ph, "cuda")
    s = McSettings(photons_per_view=10_000_000, spectrum=tables.kramers_spectrum(120.0, 2.5), coherent=1)
    brick = ctx.tuning["mc_ray_brick"]
    ctx.set_tuning(mc_ray_brick=0)
    try:
        _, _, sta = mc_project(ph, geom, s, labels_dev=labels, ctx=ctx)
    finally:
        ctx.set_tuning(mc_ray_brick=brick)
    _, _, stb = mc_project(ph, geom, s, labels_dev=labels, ctx=ctx)
    print(f"steps per history {sta['steps'] / sta['histories']:.3f} -> {stb['steps'] / stb['histories']:.3f}; "
          f"kernel {sta['kernel_ms']:.1f} -> {stb['kernel_ms']:.1f} ms")
    assert stb["steps"] < 0.5 * sta["steps"]
    for k in ("tallied_primary", "photo", "compton"):
        assert abs(sta[k] - stb[k]) < 0.01 * sta[k], k
