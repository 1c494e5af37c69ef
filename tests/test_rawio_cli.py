"""Raw files in the reference layouts (CPU) and the file-level entry points
(GPU): phantom .raw -> proj/map .raw -> volume .raw."""
import os

import numpy as np
import pytest


def test_raw_round_trip_and_sidecar(tmp_path):
    from monte_amd import rawio

    rng = np.random.default_rng(0)
    for kind, dtype, shape in (("phantom", np.uint8, (3, 4, 5)), ("counts", np.int32, (2, 6, 7)),
                               ("lineint", np.float32, (2, 6, 7)), ("volume", np.float32, (4, 3, 2))):
        a = (rng.random(shape) * 100).astype(dtype)
        p = str(tmp_path / f"{kind}.raw")
        rawio.write_raw(p, a, kind, {"note": kind})
        # headerless little-endian C order, as the reference writes it
        assert os.path.getsize(p) == a.nbytes
        assert np.array_equal(np.fromfile(p, dtype=np.dtype(dtype).newbyteorder("<")).reshape(shape), a)
        b, meta = rawio.read_raw(p, kind)
        assert b.dtype == dtype and np.array_equal(a, b) and meta == {"note": kind}


def test_raw_without_sidecar_needs_shape(tmp_path):
    from monte_amd import rawio

    p = str(tmp_path / "map.raw")
    np.arange(12, dtype="<f4").tofile(p)  # as the reference's writeRawFile leaves it
    with pytest.raises(ValueError):
        rawio.read_raw(p)
    a, _ = rawio.read_raw(p, "lineint", shape=(1, 3, 4))
    assert a.shape == (1, 3, 4) and a[0, 2, 3] == 11
    with pytest.raises(ValueError):
        rawio.read_raw(p, "lineint", shape=(1, 3, 5))


def test_photons_per_pixel():
    from monte_amd import geometry as G
    from monte_amd.mc import photons_per_pixel

    g = G.small()
    n = photons_per_pixel(g, 3 * g.nu * g.nv + 5)
    assert n.shape == (g.nu, g.nv) and n.sum() == 3 * g.nu * g.nv + 5
    # aim order iv * nu + iu: the extra photons go to iv = 0, iu = 0..4
    assert np.all(n[:5, 0] == 4) and n[5, 0] == 3 and n[0, 1] == 3


@pytest.mark.gpu
def test_cli_phantom_mc_fdk(tmp_path):
    from monte_amd import cli, rawio
    from monte_amd import geometry as G
    from monte_amd.mc import photons_per_pixel

    ph = str(tmp_path / "ph.raw")
    assert cli.main(["phantom", "--preset", "small", "--out", ph]) == 0
    out = str(tmp_path / "run")
    assert cli.main(["mc", "--phantom", ph, "--geometry", "small", "--views", "4", "--photons", "200000",
                     "--e-kev", "60", "--out", out]) == 0
    c0, meta = rawio.read_raw(out + "_proj0.raw", "counts")
    c5, _ = rawio.read_raw(out + "_proj5.raw", "counts")
    m5, _ = rawio.read_raw(out + "_map5.raw", "lineint")
    n0 = photons_per_pixel(G.small(n_views=4), 200000)[None]
    assert c0.shape == (4, 128, 128) and np.all(c5 >= c0) and np.all(c0 >= 1) and np.all(c5 <= n0)
    assert meta["photons_per_view"] == 200000
    assert np.allclose(m5, np.log(n0) - np.log(c5), rtol=0, atol=1e-5)
    vol = str(tmp_path / "vol.raw")
    assert cli.main(["fdk", "--map", out + "_map5.raw", "--geometry", "small", "--views", "4", "--out", vol,
                     "--zy", str(tmp_path / "zy.raw")]) == 0
    v, _ = rawio.read_raw(vol, "volume")
    zy, _ = rawio.read_raw(str(tmp_path / "zy.raw"), "volume_zy")
    assert v.shape == (65, 65, 65) and np.isfinite(v).all() and np.array_equal(zy, v.transpose(2, 1, 0))
