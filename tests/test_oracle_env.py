"""The oracle's environment light (oracle/cpu_ref.cpp: env_pixel,
inverted_cdf, env_init, env_sample and its MIS terms) and the P6 reader and
writer (chakumrender_amd/host.py: read_ppm6, write_ppm6) against the
reference's own EnvSource, MIS and readImage / writeImage
(tests/golden/env.npz):

  * normalizationArea (polar axes x, y, z) and getRadiance equal the
    reference's bit for bit;
  * every one of the reference's first 20,000 importance samples is a
    (theta, phi) cell the oracle's inverse CDFs can return, and the two
    samplers' theta and phi histograms agree (chi-square);
  * MIS::getSamples' weights with the environment technique, on the
    reference's own samples, equal the reference's;
  * getLightFromAllSources_MIS means agree within 4.5 sigma;
  * a constant environment: area 4 pi grey, cos(theta) of the samples
    uniform on [-1, 1] (closed forms);
  * readImage of a P6 file with comments, and writeImage's bytes (values
    out of [0, 1] included)."""
import numpy as np
import pytest

import oracle
from chakumrender_amd import abi, host, scenes
from direct_cases import agree, oracle_means


@pytest.fixture(scope="module")
def G(golden):
    return golden("env.npz")


def env_scene(G, axis=None):
    sc = scenes.sources_room()
    sc.env = scenes.EnvMap(G["env_image"], int(G["env_axis"]) if axis is None else axis,
                           tuple(G["env_col"]), float(G["env_intensity"]))
    return sc


@pytest.fixture(scope="module")
def D(G):
    return oracle.DirectScene(env_scene(G))


@pytest.mark.parametrize("axis,key", [(1, "area"), (0, "area_axis0"), (2, "area_axis2")])
def test_area_equals_reference(G, D, axis, key):
    Dx = D if axis == 1 else oracle.DirectScene(env_scene(G, axis))
    assert Dx.env_tables()[0] == float(G[key])


def test_radiance_equals_reference(G, D):
    np.testing.assert_array_equal(D.env_radiance(G["dirs"]), G["rad"])


def test_reference_samples_reachable_and_distributed_alike(G, D):
    _, im, ic = D.env_tables()
    ti, pj = G["sample_theta_idx"].astype(int), G["sample_phi_idx"].astype(int)
    assert np.isin(ti, im).all()
    rows = {int(t): set(ic[t].tolist()) for t in np.unique(ti)}
    assert all(int(p) in rows[int(t)] for t, p in zip(ti, pj))
    u = np.random.default_rng(7).random((20000, 2))
    s = D.env_sample(u)
    t2 = np.rint(np.arccos(np.clip(s[:, 2], -1, 1)) / (np.pi / 3000)).astype(int)
    p2 = np.rint((np.arctan2(s[:, 1], s[:, 0]) % (2 * np.pi)) / (2 * np.pi / 4000)).astype(int) % 4000
    for a, b, hi in ((ti, t2, 3000), (pj, p2, 4000)):
        h1, e = np.histogram(a, 24, (0, hi))
        h2, _ = np.histogram(b, e)
        keep = h1 + h2 > 0
        chi = ((h1 - h2)[keep] ** 2 / (h1 + h2)[keep]).sum()
        dof = keep.sum()
        assert chi < dof + 5 * np.sqrt(2 * dof), (chi, dof)


def test_mis_weights_equal_reference_getSamples(G, D):
    assert G["samples"].shape[1] == D.n_samples == 10 * (1 + 2 + 1 + 1) + 60
    w = D.mis_weights(G["x"], G["n"], G["wo_ref"], G["wo"], G["prim"], G["samples"])
    np.testing.assert_allclose(w, G["weights"], rtol=1e-12, atol=0)


def test_mis_means_match_reference(G, D):
    m, v = oracle_means(D, abi.DIRECT_MIS, G["x"], G["n"], G["d"], G["prim"], 2048)
    ok, bad = agree(m, v, G["mis_mean"], G["mis_var"])
    assert ok, bad
    # the environment lights the points: more than the scene's own sources
    D0 = oracle.DirectScene(scenes.sources_room())
    m0, _ = oracle_means(D0, abi.DIRECT_MIS, G["x"], G["n"], G["d"], G["prim"], 256)
    assert m.sum() > 1.2 * m0.sum()


def test_constant_environment_closed_forms():
    sc = scenes.direct_room()
    sc.env = scenes.EnvMap(np.full((8, 16, 3), 0.5), 2, (1.0, 1.0, 1.0), 1.0)
    D = oracle.DirectScene(sc)
    area, im, _ = D.env_tables()
    assert abs(area / (4 * np.pi * 0.5) - 1) < 1e-3  # the Riemann sum of sin(theta)
    s = D.env_sample(np.random.default_rng(3).random((40000, 2)))
    c = np.sort(s[:, 2])
    ks = np.abs(c - np.linspace(-1, 1, len(c))).max() / 2  # KS distance to U(-1, 1)
    assert ks < 1.63 / np.sqrt(len(c)) + 2e-3, ks  # 1 % level, plus the 3000-row grid


def test_p6_reader_equals_readImage(G, tmp_path):
    p = tmp_path / "in.ppm"
    p.write_bytes(G["ppm_in"].tobytes())
    np.testing.assert_array_equal(host.read_ppm6(p), G["ppm_in_ref"])


def test_p6_writer_matches_writeImage(G, tmp_path):
    ref = G["ppm_out_ref"].tobytes()
    img = G["ppm_out_img"]
    p = tmp_path / "out.ppm"
    host.write_ppm6(p, img)
    mine = p.read_bytes()
    H, W, _ = img.shape
    assert mine[-3 * W * H:] == ref[-3 * W * H:]  # the pixel bytes
    assert mine.startswith(b"P6\n") and b"%d %d\n255\n" % (W, H) in mine
    q = tmp_path / "ref.ppm"
    q.write_bytes(ref)
    np.testing.assert_array_equal(host.read_ppm6(q), host.read_ppm6(p))
    with pytest.raises(ValueError, match="P6"):
        (tmp_path / "p3.ppm").write_bytes(b"P3\n1 1\n255\n0 0 0 ")
        host.read_ppm6(tmp_path / "p3.ppm")
    with pytest.raises(ValueError, match="8-bits"):
        (tmp_path / "16.ppm").write_bytes(b"P6\n1 1\n65535\n" + bytes(6))
        host.read_ppm6(tmp_path / "16.ppm")
