"""The oracle's spheres and its triangle, sphere and point sources
(oracle/cpu_ref.cpp: sph_isect, sph_cone, cone_dir, mis_weight,
direct_estimate, closest) against the reference's own compiled
LightSource / MIS objects on a scene holding every kind
(tests/golden/sources.npz):

  * MIS::getSamples' weights on the reference's own samples, with the
    sphere source's cone technique and the triangle light's area technique
    in the denominator;
  * getLightFromAllSources_MIS on a scene lit by point sources only, where
    no random draw matters: the values are the reference's;
  * the means of getLightFromAllSources_MIS and _AnalyticCV, and the value of
    _Analytic, at points on triangles and on spheres (spheres shadow);
  * the cone sampler: every reference cone sample lies in the cone, and the
    oracle's fill it uniformly (mean cos theta = (1 + cos theta_max) / 2);
  * the shaders' closest hit: a camera ray whose nearest primitive is the
    sphere source returns its radiance.maxnorm()."""
import numpy as np
import pytest

import oracle
from chakumrender_amd import abi, host, scenes
from direct_cases import agree, oracle_means, sources_scene


@pytest.fixture(scope="module")
def G(golden):
    return golden("sources.npz")


@pytest.fixture(scope="module")
def D(G):
    return oracle.DirectScene(sources_scene(G))


def test_fixture_holds_every_kind(G):
    sc = sources_scene(G)
    assert len(sc.lights) == 2 and len(sc.tri_lights) == 1 and len(sc.sphere_sources) == 1
    assert len(sc.point_sources) == 2 and len(sc.spheres) == 3
    assert (G["prim"] >= sc.n_tris).sum() == 8  # points on spheres


def test_mis_weights_equal_reference_getSamples(G, D):
    assert G["samples"].shape[1] == D.n_samples == 10 * (1 + 2 + 1) + 60
    w = D.mis_weights(G["x"], G["n"], G["wo_ref"], G["wo"], G["prim"], G["samples"])
    np.testing.assert_allclose(w, G["weights"], rtol=1e-12, atol=0)


def test_point_sources_equal_reference(G):
    D = oracle.DirectScene(sources_scene(G, "p_"))
    v = D.direct_light(host.direct_params(abi.DIRECT_MIS), G["px"], G["pn"], G["pd"], G["pprim"])
    np.testing.assert_allclose(v, G["pts_mis"], rtol=1e-12, atol=1e-15)
    assert (v.max(1) > 1e-3).mean() > 0.5  # lit, not trivially zero
    assert (v.max(1) == 0).any()  # and some point shadowed or facing away


def test_analytic_equals_reference(G, D):
    v = D.direct_light(host.direct_params(abi.DIRECT_ANALYTIC), G["x"], G["n"], G["d"], G["prim"])
    np.testing.assert_allclose(v, G["analytic"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("name", ["mis", "analytic_cv16"])
def test_estimator_means_match_reference(G, D, name):
    est, cv, key = (abi.DIRECT_MIS, 0, "mis") if name == "mis" else (abi.DIRECT_ANALYTIC_CV, 16, "cv")
    m, v = oracle_means(D, est, G["x"], G["n"], G["d"], G["prim"], 2048, cv=cv)
    ok, bad = agree(m, v, G[f"{key}_mean"], G[f"{key}_var"])
    assert ok, (bad, m[tuple(bad[0])] if len(bad) else None)
    assert (G[f"{key}_mean"].max(1) > 1e-3).mean() > 0.5


def _cone(sc, x):
    j = int(sc.sphere_sources[0])
    v = sc.spheres["center"][j] - x
    L = np.linalg.norm(v)
    return v / L, np.sqrt(1 - (sc.spheres["radius"][j] / L) ** 2)


def test_cone_samples(G, D):
    sc = sources_scene(G)
    for i in range(len(G["prim"])):
        w, cm = _cone(sc, G["x"][i])
        assert (G["samples"][i, :10] @ w >= cm - 1e-12).all()  # the reference's
    i = 3
    w, cm = _cone(sc, G["x"][i])
    s = np.concatenate([D.mis_samples(G["x"][i], G["n"][i], G["wo_ref"][i], G["wo"][i],
                                      G["prim"][i], i, e, (5, 6))[:10] for e in range(400)])
    c = s @ w
    assert (c >= cm - 1e-12).all()
    mean_c = (1 + cm) / 2
    sd = (1 - cm) / np.sqrt(12) / np.sqrt(len(c))
    assert abs(c.mean() - mean_c) < 5 * sd
    perp = s - c[:, None] * w
    assert np.abs(perp.mean(0)).max() < 5 * np.sqrt(1 - cm ** 2) / np.sqrt(len(c))


def test_sample_slots_follow_getSamples_order(G, D):
    """10 cone slots, 10 per mesh light, 10 for the triangle light, 60 BSDF."""
    sc = sources_scene(G)
    i = 0
    x = G["x"][i]
    s = D.mis_samples(x, G["n"][i], G["wo_ref"][i], G["wo"][i], G["prim"][i], i, 0, (7, 9))
    assert s.shape == (100, 3)
    for j, (first, cnt) in enumerate(sc.lights.tolist()):
        rays = np.c_[np.repeat(x[None], 10, 0), s[10 + 10 * j:20 + 10 * j]]
        assert (oracle.closest_brute(sc.tris[first:first + cnt].astype(np.float64), rays)["id"] >= 0).all()
    t = int(sc.tri_lights[0])
    rays = np.c_[np.repeat(x[None], 10, 0), s[30:40]]
    assert (oracle.closest_brute(sc.tris[t:t + 1].astype(np.float64), rays)["id"] >= 0).all()


def test_camera_rays_on_the_sphere_source_return_its_maxnorm(G):
    sc = sources_scene(G).with_camera(pos=(0.78, 0.55, 1.4), dir=(0.0, 0.27, -0.78), width=24,
                                      height=24, fovx_deg=12.0, fovy_deg=12.0)
    D = oracle.DirectScene(sc)
    params = dict(spp=1, jitter_mode=abi.JITTER_REFERENCE, seed=12345)
    acc, _ = D.render_shade(abi.SHADE_DIRECT_ONLY, params, host.direct_params(abi.DIRECT_MISCV))
    # the camera rays (reference jitter 0.34, 0.20) against every primitive
    pos, d0, cx, cy = sc.camera.basis()
    W = H = 24
    ys, xs = np.mgrid[0:H, 0:W]
    yy = H - 1 - ys
    dirs = cx * ((xs[..., None] + 0.34) / W - 0.5) + cy * ((yy[..., None] + 0.20) / H - 0.5) + d0
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    rays = np.c_[np.repeat(pos[None], W * H, 0), dirs.reshape(-1, 3)]
    sp = np.c_[sc.spheres["center"], sc.spheres["radius"]]
    sid, ts = oracle.spheres_brute(sp, rays)
    th = oracle.closest_brute(sc.tris.astype(np.float64), rays)["t"]
    src = (sid == sc.sphere_sources[0]) & (ts < th)
    assert 5 < src.sum() < W * H
    rad = sc.mats["radiance"][sc.spheres["mat"][sc.sphere_sources[0]]].astype(np.float64)
    np.testing.assert_allclose(acc.reshape(-1, 3)[src], np.broadcast_to(rad / rad.max(), (src.sum(), 3)),
                               rtol=1e-15)
    assert (acc.reshape(-1, 3)[~src].max(1) != 1.0).all()
