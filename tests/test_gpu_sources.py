"""Spheres and the triangle, sphere and point sources on the GPU
(chk_scene_create_desc, chk_mis_weights, chk_direct_light, chk_render_shade)
against the reference's own MIS::getSamples and estimators
(tests/golden/sources.npz) and against the FP64 oracle on the same Philox
samples.

Bars as for the mesh-light estimators: reference weights 1e-10 relative;
point-source values (no random draw) 1e-9 relative; same samples as the
oracle 1e-8 relative per point except points whose visibility differs
between the FP32 triangle traversal and the FP64 brute force (at most 2%);
means against the reference's 4.5 sigma; renders per pixel against the
oracle on the same keys."""
import numpy as np
import pytest

import oracle
from chakumrender_amd import abi, host, scenes
from direct_cases import ESTIMATORS, agree, means, sources_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G(golden):
    return golden("sources.npz")


@pytest.fixture(scope="module")
def env(ctx, G):
    sc = sources_scene(G)
    ds = abi.DeviceScene(ctx, sc)
    return sc, ds, host.SceneLights(ds), oracle.DirectScene(sc)


def _points(sc, k_tri, k_sph, seed):
    a = scenes.shading_points(sc, k_tri, seed=seed)
    b = scenes.sphere_points(sc, k_sph, seed=seed + 1)
    return tuple(np.concatenate([u, v]) for u, v in zip(a, b))


def test_mis_weights_equal_reference_getSamples(env, G):
    _, _, SL, _ = env
    w = SL.mis_weights(G["x"], G["n"], G["wo_ref"], G["wo"], G["prim"], G["samples"])
    np.testing.assert_allclose(w, G["weights"], rtol=1e-10, atol=0)


def test_analytic_equals_reference(env, G):
    _, _, SL, _ = env
    v = SL.getLightFromMeshSource_Analytic(G["x"], G["n"], G["d"], G["prim"])
    np.testing.assert_allclose(v, G["analytic"], rtol=1e-9, atol=1e-14)


def test_point_sources_equal_reference(ctx, G):
    ds = abi.DeviceScene(ctx, sources_scene(G, "p_"))
    v = host.SceneLights(ds).getLightFromAllSources_MIS(G["px"], G["pn"], G["pd"], G["pprim"])
    np.testing.assert_allclose(v, G["pts_mis"], rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("name", list(ESTIMATORS))
def test_same_samples_as_oracle(env, name):
    sc, _, SL, D = env
    x, n, d, prim = _points(sc, 384, 128, 21)
    kw = ESTIMATORS[name]
    for e in (0, 1):
        g = SL.estimate(kw["estimator"], x, n, d, prim, estimate=e,
                        cv_samples=kw.get("cv_samples", 0), flags=kw.get("flags", 0))
        P = host.direct_params(**kw)
        P.estimate = e
        o = D.direct_light(P, x, n, d, prim)
        assert np.isfinite(g).all()
        rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-6)
        bad = (rel > 1e-8).any(1)
        on_sph = prim >= sc.n_tris
        print(f"{name} e={e}: points off {bad.sum()} / {len(prim)} ({bad[on_sph].sum()} on spheres), "
              f"max rel {rel.max():.2e}")
        assert bad.mean() <= 0.02, np.nonzero(bad)[0][:10]
        assert abs(g.sum() - o.sum()) <= 0.01 * abs(o.sum())
        assert (o[on_sph].max(1) > 0).any()


@pytest.mark.parametrize("name,key,cv", [("mis", "mis", 0), ("analytic_cv16", "cv", 16)])
def test_means_match_reference(env, G, name, key, cv):
    _, _, SL, _ = env
    est = ESTIMATORS[name]["estimator"]
    m, v = means(lambda e: SL.estimate(est, G["x"], G["n"], G["d"], G["prim"], estimate=e,
                                       cv_samples=cv), len(G["prim"]), 1024)
    ok, bad = agree(m, v, G[f"{key}_mean"], G[f"{key}_var"])
    assert ok, bad


SHADERS = [(abi.SHADE_DIRECT_ONLY, "mis"), (abi.SHADE_DIRECT_ONLY, "miscv_vis"),
           (abi.SHADE_EXPLICIT, "mis"), (abi.SHADE_IMPLICIT, "mis")]


@pytest.mark.parametrize("shader,name", SHADERS)
def test_render_matches_oracle(ctx, shader, name):
    """The shaders over a scene with spheres (closest hits and shadows) and
    every source kind, on the same keys as the oracle."""
    sc = scenes.sources_room(width=32, height=32)
    ds = abi.DeviceScene(ctx, sc)
    dp = host.direct_params(**ESTIMATORS[name])
    spp = 8
    P = dict(spp=spp, seed_lo=0x5EED0010, flags=abi.RENDER_SUMSQ)
    a, s2, st = host.render_shade(ds, P, shader, dp)
    oa, os2 = oracle.DirectScene(sc).render_shade(shader, P, dp, sumsq=True)
    gm, om = a / spp, oa / spp
    var = np.maximum(s2 / spp - gm ** 2, 0) / spp + np.maximum(os2 / spp - om ** 2, 0) / spp
    z = np.abs(gm - om) / np.sqrt(var + 1e-18)
    close = (np.abs(gm - om) <= 1e-5 * np.maximum(np.abs(om), 1e-3)).all(2)
    iz = abs(gm.mean() - om.mean()) / np.sqrt(var.sum() / var.size ** 2 + 1e-30)
    print(f"shader {shader} {name}: pixels equal to 1e-5 {close.mean():.3f}, max z {z.max():.2f}, "
          f"image z {iz:.2f}, mean {gm.mean():.5f} vs {om.mean():.5f}")
    assert st["paths"] == 32 * 32 * spp
    assert gm.mean() > 0
    assert close.mean() >= 0.8
    assert (z <= 4.0).mean() >= 0.99
    assert iz < 4.0


def test_reference_scene1_renders_like_oracle(ctx, golden):
    """The reference's scene1.xml (sphere walls, two GGX balls, a sphere
    source) at 128 x 96: the shipped shadeDirectOnly with MISCV is black but
    for the visible source (MISCV returns 0 without a mesh light,
    lightSources_CV.cpp:400), exactly as the oracle; with
    getLightFromAllSources_MIS the pixels follow the oracle's."""
    from chakumrender_amd import scene_xml

    G = golden("scene_xml.npz")
    sc = scene_xml.load_scene_xml(str(G["scene1_xml"])).with_camera(width=128, height=96)
    ds = abi.DeviceScene(ctx, sc)
    D = oracle.DirectScene(sc)
    P = dict(spp=2, seed_lo=0x5EED0011, flags=abi.RENDER_SUMSQ)
    dp = host.direct_params(abi.DIRECT_MISCV)
    a, _, _ = host.render_direct(ds, P, dp)
    oa, _ = D.render_direct(P, dp)
    assert (a == oa).all(2).mean() >= 0.999  # a silhouette pixel may part on rounding
    lit = a.max(2) > 0
    assert 0 < lit.mean() < 0.2  # the source only
    dp = host.direct_params(abi.DIRECT_MIS)
    a, _, _ = host.render_direct(ds, P, dp)
    oa, _ = D.render_direct(P, dp)
    close = (np.abs(a - oa) <= 1e-5 * np.maximum(np.abs(oa), 1e-3)).all(2)
    print(f"scene1 MIS: pixels equal to 1e-5 {close.mean():.3f}, mean {a.mean():.5f} vs {oa.mean():.5f}")
    assert (a.max(2) > 0).mean() > 0.9
    assert close.mean() >= 0.95
    assert abs(a.mean() - oa.mean()) <= 2e-3 * abs(oa.mean())


def test_scene_errors(ctx):
    sc = scenes.sources_room(width=8, height=8)

    def bad(**kw):
        from dataclasses import replace
        return replace(sc, **kw)

    with pytest.raises(abi.ChkError, match="triangle light"):
        abi.DeviceScene(ctx, bad(tri_lights=np.array([0], np.uint32)))  # not an emitter
    with pytest.raises(abi.ChkError, match="triangle light"):
        abi.DeviceScene(ctx, bad(tri_lights=np.array([sc.n_tris], np.uint32)))
    s = sc.spheres.copy()
    s[0]["radius"] = 0.0
    with pytest.raises(abi.ChkError, match="sphere 0"):
        abi.DeviceScene(ctx, bad(spheres=s))
    with pytest.raises(abi.ChkError, match="sphere source"):
        abi.DeviceScene(ctx, bad(sphere_sources=np.array([0], np.uint32)))  # not an emitter
    many = np.zeros(4097, scenes.SPHERE_DTYPE)
    many["radius"] = 1.0
    many["center"][:, 0] = np.arange(4097) * 3.0
    with pytest.raises(abi.ChkError, match="4096 spheres"):
        abi.DeviceScene(ctx, bad(spheres=many, sphere_sources=np.zeros(0, np.uint32)))
    ds = abi.DeviceScene(ctx, sc)
    with pytest.raises(abi.ChkError, match="spheres or triangle, sphere, point or environment sources"):
        host.render(ds, dict(spp=1))
    SL = host.SceneLights(ds)
    x, n, d, prim = _points(sc, 2, 2, 5)
    prim[3] = sc.n_tris + len(sc.spheres)
    with pytest.raises(abi.ChkError, match="prim"):
        SL.getLightFromAllSources_MIS(x, n, d, prim)
    prim[3] = sc.n_tris + int(sc.sphere_sources[0])  # an emitter sphere has no BSDF sampler
    with pytest.raises(abi.ChkError, match="sampler"):
        SL.getLightFromAllSources_MIS(x, n, d, prim)


def test_sphere_closest_replaces_triangle_only_when_nearer(ctx):
    """A sphere touching a triangle from behind on the view axis: the
    triangle is hit (the sphere is not nearer, tS < tT fails,
    shader.cpp:348); moved 0.01 forward, the sphere is hit."""
    cam = scenes.Camera((0.0, 0.0, 3.0), (0, 0, -1), (0, 1, 0), 10.0, 10.0, 8, 8)
    mats = np.zeros(2, scenes.MATERIAL_DTYPE)
    mats[0] = scenes.emitter_material((1.0, 0.5, 0.25), 1.0)  # triangle: maxnorm (1, .5, .25)
    mats[1] = scenes._material(scenes.BRDF_NONE, reflectance=0.0, color=(0, 0, 0),
                               light=scenes.LIGHT_VOLUME, radiance=(0.25, 0.5, 1.0))
    tri = np.array([[-5, -5, 0, 5, -5, 0, 0, 5, 0]], np.float32)
    for z, want in ((-1.0, (1.0, 0.5, 0.25)), (-0.99, (0.25, 0.5, 1.0))):
        sph = np.zeros(1, scenes.SPHERE_DTYPE)
        sph[0]["center"] = (0.0, 0.0, z)
        sph[0]["radius"] = 1.0
        sph[0]["mat"] = 1
        sc = scenes.Scene(tri, np.zeros(1, np.uint32), mats, np.zeros(0, scenes.LIGHT_DTYPE), cam,
                          spheres=sph)
        ds = abi.DeviceScene(ctx, sc)
        a, _, _ = host.render_shade(ds, dict(spp=1, jitter_mode=abi.JITTER_REFERENCE),
                                    abi.SHADE_DIRECT_ONLY, host.direct_params(abi.DIRECT_MIS))
        np.testing.assert_allclose(a[3:5, 3:5].reshape(-1, 3), np.broadcast_to(want, (4, 3)),
                                   rtol=1e-7)


def test_scene_without_triangles(ctx):
    """Spheres, a sphere source and point sources but no triangle at all (an
    empty BVH): the shaders render what the oracle renders."""
    sc0 = scenes.sources_room(width=24, height=24)
    sc = scenes.Scene(np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), sc0.mats,
                      np.zeros(0, scenes.LIGHT_DTYPE), sc0.camera, spheres=sc0.spheres,
                      sphere_sources=sc0.sphere_sources, point_sources=sc0.point_sources)
    ds = abi.DeviceScene(ctx, sc)
    assert ds.info()["n_tris"] == 0
    dp = host.direct_params(abi.DIRECT_MIS)
    P = dict(spp=4, seed_lo=0x5EED0014)
    for shader in (abi.SHADE_DIRECT_ONLY, abi.SHADE_EXPLICIT, abi.SHADE_IMPLICIT):
        a, _, _ = host.render_shade(ds, P, shader, dp)
        oa, _ = oracle.DirectScene(sc).render_shade(shader, P, dp)
        close = (np.abs(a - oa) <= 1e-5 * np.maximum(np.abs(oa), 1e-3)).all(2)
        print(f"no triangles, shader {shader}: pixels equal {close.mean():.3f}, mean {a.mean():.5f}")
        assert close.mean() >= 0.95
    assert a.mean() >= 0 and np.isfinite(a).all()
