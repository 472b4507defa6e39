"""Write the golden fixtures of tests/golden/ from the reference's OWN
compiled code (oracle/_ref/libchkref.so, built by `make -C oracle ref` from
the reference sources; see oracle/ref_shim.cpp for the entry points used).

Run in a checkout that has the reference mounted:
    python tests/golden/make_golden.py
Fixtures (inputs and expected outputs only):
    bvh_closest.npz   BvhAccel::intersectT on a 1,933-triangle Cornell scene
                      and a 9,680-triangle displaced geodesic sphere
    shadow.npz        shadow() semantics (intersectT + 1e-5 margin)
    brdf.npz          BasePrimitive::brdf and MaterialType::pdfEval
    ltc.npz           M_GGX and LightSource::getLightFromMeshSource_Analytic
    spheres.npz       BvhAccel::intersectS over Sphere primitives
    scene_xml.npz     the reference's two scene files (inputs) and what its own
                      front end builds from them: triangle list with
                      materials, mesh-light sizes, spheres, sphere sources,
                      camera
    direct.npz        MIS::getSamples (samples and weights),
                      getLightFromMeshSource_Analytic, and the mean and
                      batch-means variance of 8192 calls of
                      getLightFromAllSources_MIS and
                      getLightFromMeshSource_AnalyticCV (16 samples per light)
                      at 32 shading points of the direct-lighting room
                      (chakumrender_amd.scenes.direct_room; the scene arrays
                      are stored too).  The reference seeds its samplers from
                      clock(), so the two statistical rows differ from run to
                      run within their variance.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from chakumrender_amd import scenes  # noqa: E402

OUT = Path(__file__).resolve().parent


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def rays_for(scene, rng, m):
    """Camera-like rays through the scene plus rays between random points in
    the scene bounds; float32-representable origins and directions."""
    t = scene.tris.astype(np.float64).reshape(-1, 3, 3)
    finite = t[np.all(np.abs(t) < 1e5, axis=(1, 2))]
    lo, hi = finite.reshape(-1, 3).min(0), finite.reshape(-1, 3).max(0)
    pos, d, cx, cy = scene.camera.basis()
    k = m // 2
    u = rng.uniform(-0.5, 0.5, (k, 2))
    cd = unit(cx * u[:, :1] + cy * u[:, 1:] + d)
    co = np.repeat(pos[None], k, 0)
    a = rng.uniform(lo, hi, (m - k, 3))
    b = rng.uniform(lo, hi, (m - k, 3))
    rd = unit(b - a)
    o = np.concatenate([co, a]).astype(np.float32).astype(np.float64)
    dd = unit(np.concatenate([cd, rd])).astype(np.float32).astype(np.float64)
    return np.concatenate([o, dd], 1)


def gen_bvh(R, rng):
    out = {}
    cfg1 = scenes.cornell()
    ico = scenes.icosphere_scene(freq=22, n_lights=8)
    for name, sc in (("cornell", cfg1), ("ico", ico)):
        tris = sc.tris.astype(np.float64)
        rays = rays_for(sc, rng, 8192)
        m = rays.shape[0]
        ids = np.empty(m, np.int32)
        t = np.empty(m)
        N = np.empty((m, 3))
        R.ref_bvh_closest(np.ascontiguousarray(tris), tris.shape[0], np.ascontiguousarray(rays), m,
                          ids, t, N)
        out[f"{name}_tris"] = sc.tris
        out[f"{name}_rays"] = rays
        out[f"{name}_id"] = ids
        out[f"{name}_t"] = t
        out[f"{name}_N"] = N
    np.savez_compressed(OUT / "bvh_closest.npz", **out)
    return cfg1


def gen_shadow(R, rng, sc):
    tris = sc.tris.astype(np.float64)
    rays = rays_for(sc, rng, 4096)
    # distances: up to the first hit (+/- a bit) so both outcomes occur
    ids = np.empty(rays.shape[0], np.int32)
    t = np.empty(rays.shape[0])
    R.ref_bvh_closest(np.ascontiguousarray(tris), tris.shape[0], np.ascontiguousarray(rays),
                      rays.shape[0], ids, t, np.empty((rays.shape[0], 3)))
    tt = np.where(ids >= 0, t, 3.0)
    dist = tt * rng.uniform(0.5, 1.5, tt.shape)
    r7 = np.ascontiguousarray(np.concatenate([rays, dist[:, None]], 1))
    occ = np.empty(r7.shape[0], np.uint8)
    R.ref_bvh_occluded(np.ascontiguousarray(tris), tris.shape[0], r7, r7.shape[0], occ)
    np.savez_compressed(OUT / "shadow.npz", tris=sc.tris, rays7=r7, occluded=occ)


def gen_spheres(R, rng):
    """BvhAccel::intersectS over 64 random spheres (plus the scene front end's
    dummy sphere, prepareScene.cpp:238) for rays from inside and outside."""
    k = 64
    c = rng.uniform(-2, 2, (k, 3))
    r = rng.uniform(0.05, 0.6, k)
    spheres = np.concatenate([np.c_[c, r], [[27.0, 1e5, 47.0, 1.0]]]).astype(np.float32).astype(np.float64)
    m = 8192
    o = rng.uniform(-3, 3, (m, 3))
    o[: m // 4] = c[rng.integers(0, k, m // 4)]  # a quarter start inside a sphere
    d = unit(rng.normal(size=(m, 3)))
    rays = np.concatenate([o, d], 1).astype(np.float32).astype(np.float64)
    rays[:, 3:] = unit(rays[:, 3:])
    ids = np.empty(m, np.int32)
    t = np.empty(m)
    R.ref_bvh_spheres(np.ascontiguousarray(spheres), spheres.shape[0], np.ascontiguousarray(rays),
                      m, ids, t)
    np.savez_compressed(OUT / "spheres.npz", spheres=spheres, rays=rays, id=ids, t=t)


def hemi(rng, n, k):
    v = unit(rng.normal(size=(k, 3)))
    return np.where((v * n).sum(1, keepdims=True) < 0, -v, v)


def gen_brdf(R, rng):
    cases = []
    # (brdf, ks, alpha, phong_exp, int_ior, ext_ior, reflectance)
    for ks in (0.0, 0.3, 1.0):
        cases.append((scenes.CLASSIC_PHONG, ks, 0.0, 20.0, 0.0, 0.0, 0.9))
    for bt in (scenes.GGX, scenes.BECKMANN):
        for alpha in (0.05, 0.3, 0.8):
            for ks in (0.5, 1.0):
                cases.append((bt, ks, alpha, 0.0, 1.5, 1.0, 0.8))
    for alpha in (2.0, 50.0, 798.0):
        for ks in (0.5, 1.0):
            cases.append((scenes.PHONG, ks, alpha, 0.0, 1.5, 1.0, 0.8))
    k = 512
    rows = {k_: [] for k_ in ("case", "n", "wo_ref", "wo", "wi", "f", "pdf")}
    for ci, (bt, ks, alpha, pe, ii, ei, refl) in enumerate(cases):
        n = unit(rng.normal(size=(k, 3)))
        wo = hemi(rng, n, k)
        # bias half the wi toward the mirror direction so specular peaks are sampled
        wo_ref = unit(2 * (n * wo).sum(1, keepdims=True) * n - wo)
        wi_u = hemi(rng, n, k)
        wi_s = unit(wo_ref + rng.normal(size=(k, 3)) * 0.15)
        wi = np.where(np.arange(k)[:, None] % 2 == 0, wi_u, wi_s)
        wi = np.where((wi * n).sum(1, keepdims=True) <= 0.02, wi_u, wi)
        keep = ((wi * n).sum(1) > 0.02) & ((wo * n).sum(1) > 0.02)
        n, wo, wo_ref, wi = n[keep], wo[keep], wo_ref[keep], wi[keep]
        # float32-representable directions, renormalised in FP64: the
        # reference formulas assume unit vectors (materialTypes.cpp:8) and
        # amplify |v| - 1 for small roughness; the GPU receives the float32
        # rounding of these unit vectors
        n, wo, wo_ref, wi = (unit(a.astype(np.float32).astype(np.float64)) for a in (n, wo, wo_ref, wi))
        m = n.shape[0]
        f = np.empty(m)
        pdf = np.empty(m)
        # material parameters as the float32 chk_material record holds them
        ks, alpha, pe, ii, ei, refl = (float(np.float32(v)) for v in (ks, alpha, pe, ii, ei, refl))
        R.ref_brdf(bt, ks, alpha, pe, ii, ei, refl, m, np.ascontiguousarray(n),
                   np.ascontiguousarray(wo_ref), np.ascontiguousarray(wo), np.ascontiguousarray(wi),
                   f, pdf)
        rows["case"].append(np.full(m, ci, np.int32))
        for key, arr in (("n", n), ("wo_ref", wo_ref), ("wo", wo), ("wi", wi), ("f", f), ("pdf", pdf)):
            rows[key].append(arr)
    out = {key: np.concatenate(v) for key, v in rows.items()}
    out["cases"] = np.array(cases, np.float64)
    np.savez_compressed(OUT / "brdf.npz", **out)


def gen_ltc(R, rng):
    # M_GGX on a grid that straddles bin edges plus random points
    th = np.concatenate([rng.uniform(0, np.pi / 2, 1500), rng.uniform(np.pi / 2, 2.0, 100),
                         (np.arange(0, 64) + 0.5) / 64 * np.pi / 2]).astype(np.float32)
    # alphas whose sqrt(alpha)*64 lies within a few ulp of a bin edge decide
    # how M_GGX's sqrt is evaluated (ltc.h:20)
    edge = np.float32(((np.arange(1, 64)) / 64.0) ** 2)
    near = np.concatenate([np.nextafter(edge, np.float32(0)), edge, np.nextafter(edge, np.float32(1))])
    al = np.concatenate([rng.uniform(0.0, 1.0, 1600 - near.size),
                         ((np.arange(0, 64) + 0.5) / 64) ** 2, near]).astype(np.float32)
    th = th[:al.size] if th.size >= al.size else np.concatenate([th, rng.uniform(0, np.pi / 2, al.size - th.size).astype(np.float32)])
    M = np.empty((th.size, 9))
    Mi = np.empty((th.size, 9))
    amp = np.empty(th.size)
    det = np.empty(th.size)
    for i in range(th.size):
        R.ref_M_GGX(float(th[i]), float(al[i]), M[i], Mi[i], amp[i:i + 1], det[i:i + 1])
    # shading points
    P = scenes.ltc_points(3000, seed=0x5EED0103, n_quads=16)
    tris, rad = scenes.quads_to_tris(P["quads"], P["radiance"])
    tris = tris.astype(np.float32).astype(np.float64)
    rad = rad.astype(np.float32).astype(np.float64)
    x = P["x"].astype(np.float64)
    n = P["n"].astype(np.float64)
    v = P["v"].astype(np.float64)
    alpha = P["alpha"].astype(np.float64)
    k = x.shape[0]
    ks = np.array([0.0, 0.5, 1.0])[np.arange(k) % 3]
    rgb = np.empty((k, 3))
    R.ref_ltc_shade(k, np.ascontiguousarray(x), np.ascontiguousarray(n), np.ascontiguousarray(v),
                    np.ascontiguousarray(alpha), np.ascontiguousarray(ks), np.ascontiguousarray(tris),
                    np.ascontiguousarray(rad), tris.shape[0], rgb)
    # an axis-aligned light (A.x == C.x on a triangle) exercising the CA-edge rule
    q = np.array([[[-0.5, 2.0, -0.5], [0.5, 2.0, -0.5], [0.5, 2.0, 0.5], [-0.5, 2.0, 0.5]]])
    t2, r2 = scenes.quads_to_tris(q, [[5.0, 5.0, 5.0]])
    k2 = 1000
    x2 = np.c_[rng.uniform(-2, 2, k2), rng.uniform(1.0, 2.5, k2), rng.uniform(-2, 2, k2)].astype(np.float32).astype(np.float64)
    n2 = unit(rng.normal(size=(k2, 3)) + [0, 1.0, 0]).astype(np.float32).astype(np.float64)
    v2 = hemi(rng, n2, k2).astype(np.float32).astype(np.float64)
    a2 = rng.uniform(0.02, 1.0, k2).astype(np.float32).astype(np.float64)
    ks2 = np.array([0.0, 0.5, 1.0])[np.arange(k2) % 3]
    rgb2 = np.empty((k2, 3))
    R.ref_ltc_shade(k2, np.ascontiguousarray(x2), np.ascontiguousarray(n2), np.ascontiguousarray(v2),
                    np.ascontiguousarray(a2), np.ascontiguousarray(ks2), np.ascontiguousarray(t2),
                    np.ascontiguousarray(r2), t2.shape[0], rgb2)
    np.savez_compressed(OUT / "ltc.npz", theta=th, alpha_m=al, M=M, Minv=Mi, amp=amp, det=det,
                        x=x, n=n, v=v, alpha=alpha, ks=ks, tris=tris, rad=rad, rgb=rgb,
                        x2=x2, n2=n2, v2=v2, alpha2=a2, ks2=ks2, tris2=t2, rad2=r2, rgb2=rgb2)


def gen_direct():
    sc = scenes.direct_room()
    x, n, d, prim = scenes.shading_points(sc, 32)
    wo = -d
    wo_ref = unit(n * 2.0 * (n * wo).sum(1, keepdims=True) + d)  # lightSources.cpp:44-45
    Rd = oracle.RefDirect(sc)
    smp, w = Rd.mis_get_samples(x, n, wo_ref, wo, prim)
    analytic, _ = Rd.stats(0, x, n, d, prim, reps=1, batches=1)
    mis_m, mis_v = Rd.stats(2, x, n, d, prim, reps=8192, batches=64)
    cv_m, cv_v = Rd.stats(1, x, n, d, prim, reps=8192, cv_samples=16, batches=64)
    np.savez_compressed(OUT / "direct.npz", tris=sc.tris, tri_mat=sc.tri_mat, mats=sc.mats,
                        lights=sc.lights, x=x, n=n, d=d, prim=prim, wo_ref=wo_ref, wo=wo,
                        samples=smp, weights=w, analytic=analytic, mis_mean=mis_m, mis_var=mis_v,
                        cv_mean=cv_m, cv_var=cv_v, cv_samples=16, reps=8192)


def _scene_arrays(prefix, sc):
    return {f"{prefix}{k}": getattr(sc, k) for k in
            ("tris", "tri_mat", "mats", "lights", "spheres", "tri_lights", "sphere_sources",
             "point_sources")}


def _points(sc, k_tri, k_sph):
    a = scenes.shading_points(sc, k_tri, seed=3)
    b = scenes.sphere_points(sc, k_sph, seed=4)
    return tuple(np.concatenate([u, v]) for u, v in zip(a, b))


def gen_sources():
    """sources_room: spheres, a sphere source, a triangle light, point
    sources and mesh lights.  The reference's MIS::getSamples (samples and
    weights), its getLightFromMeshSource_Analytic value and its
    getLightFromAllSources_MIS and _AnalyticCV means at points on triangles
    and on spheres; and its getLightFromAllSources_MIS values on a scene whose
    only sources are the point sources (no random draw then matters)."""
    sc = scenes.sources_room()
    x, n, d, prim = _points(sc, 24, 8)
    wo = -d
    wo_ref = unit(n * 2.0 * (n * wo).sum(1, keepdims=True) + d)  # lightSources.cpp:44-45
    Rd = oracle.RefDirect(sc)
    smp, w = Rd.mis_get_samples(x, n, wo_ref, wo, prim)
    analytic, _ = Rd.stats(0, x, n, d, prim, reps=1, batches=1)
    mis_m, mis_v = Rd.stats(2, x, n, d, prim, reps=8192, batches=64)
    cv_m, cv_v = Rd.stats(1, x, n, d, prim, reps=8192, cv_samples=16, batches=64)
    sp = scenes.sources_room(mesh=False, sphere_source=False, tri_light=False)
    px, pn, pd, pp = _points(sp, 24, 8)
    Rp = oracle.RefDirect(sp)
    pts_mis, _ = Rp.stats(2, px, pn, pd, pp, reps=1, batches=1)
    np.savez_compressed(OUT / "sources.npz", **_scene_arrays("", sc), **_scene_arrays("p_", sp),
                        x=x, n=n, d=d, prim=prim, wo_ref=wo_ref, wo=wo, samples=smp, weights=w,
                        analytic=analytic, mis_mean=mis_m, mis_var=mis_v, cv_mean=cv_m,
                        cv_var=cv_v, cv_samples=16, reps=8192, px=px, pn=pn, pd=pd, pprim=pp,
                        pts_mis=pts_mis)


def env_scene(axis=1):
    """sources_room under the sky environment light."""
    sc = scenes.sources_room()
    sc.env = scenes.EnvMap(scenes.sky_image(), axis, (1.0, 0.9, 0.8), 2.0)
    return sc


def gen_env(R):
    """EnvSource (lightSources.h:33-109, lightSources.cpp:662-782) over the
    sky image, polar axis y, in sources_room: the reference's
    normalizationArea, getRadiance of random directions, the (theta, phi)
    grid indices of the first 20,000 of its importance samples, its
    MIS::getSamples (samples and weights) and its getLightFromAllSources_MIS
    means at points on triangles and spheres; normalizationArea with the x and
    z axes; and readImage / writeImage (ppm.cpp:36-152) on P6 files."""
    import tempfile

    sc = env_scene(1)
    x, n, d, prim = _points(sc, 16, 8)
    wo = -d
    wo_ref = unit(n * 2.0 * (n * wo).sum(1, keepdims=True) + d)  # lightSources.cpp:44-45
    Rd = oracle.RefDirect(sc)
    area = Rd.env_area()
    rng = np.random.default_rng(20261205)
    dirs = unit(rng.normal(size=(1024, 3)))
    rad = Rd.env_radiance(dirs)
    s = Rd.env_samples(20000)
    ti = np.rint(np.arccos(np.clip(s[:, 2], -1, 1)) / (np.pi / 3000)).astype(np.uint16)
    pj = (np.rint((np.arctan2(s[:, 1], s[:, 0]) % (2 * np.pi)) / (2 * np.pi / 4000)) % 4000).astype(np.uint16)
    smp, w = Rd.mis_get_samples(x, n, wo_ref, wo, prim)
    mis_m, mis_v = Rd.stats(2, x, n, d, prim, reps=8192, batches=64)
    areas = {}
    for ax in (0, 2):
        areas[ax] = oracle.RefDirect(env_scene(ax)).env_area()
    # P6: a file with a comment line read by readImage; writeImage's bytes
    img = scenes.sky_image(width=7, height=5)
    raw = np.round(img * 255).astype(np.uint8)
    ppm_in = b"P6\n# a comment\n# another\n7 5\n255\n" + raw.tobytes()
    out_img = np.linspace(-0.6, 1.4, 7 * 5 * 3).reshape(5, 7, 3)
    with tempfile.TemporaryDirectory() as tmp:
        Path(tmp, "env1.ppm").write_bytes(ppm_in)
        buf, wh = np.zeros((64, 3)), np.zeros(2, np.uint32)
        assert R.ref_read_ppm(tmp.encode(), buf, 64, wh) == 1
        ppm_in_ref = buf[:int(wh[0] * wh[1])].reshape(int(wh[1]), int(wh[0]), 3)
        assert R.ref_write_ppm(tmp.encode(), np.ascontiguousarray(out_img.reshape(-1, 3)), 7, 5) == 1
        ppm_out_ref = Path(tmp, "env1_cpy.ppm").read_bytes()
    np.savez_compressed(OUT / "env.npz", env_image=sc.env.image, env_axis=1,
                        env_col=np.array(sc.env.col), env_intensity=sc.env.intensity,
                        area=area, area_axis0=areas[0], area_axis2=areas[2], dirs=dirs, rad=rad,
                        sample_theta_idx=ti, sample_phi_idx=pj, x=x, n=n, d=d, prim=prim,
                        wo_ref=wo_ref, wo=wo, samples=smp, weights=w, mis_mean=mis_m,
                        mis_var=mis_v, reps=8192, ppm_in=np.frombuffer(ppm_in, np.uint8),
                        ppm_in_ref=ppm_in_ref, ppm_out_img=out_img,
                        ppm_out_ref=np.frombuffer(ppm_out_ref, np.uint8))


def gen_scene_xml(R):
    """The reference's own front end (initSceneParser, prepareScreenCamera,
    prepareObjects, prepareLightSources) on its two scene files; the files
    themselves are stored as the inputs."""
    import ctypes as C

    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    up = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    R.ref_parse_scene.restype = C.c_int
    R.ref_parse_scene.argtypes = [C.c_char_p, up, dp, dp, dp, up, dp, dp, C.c_uint32]
    out = {}
    cap = 4096
    for name in ("scene0", "scene1"):
        path = oracle.REF_ROOT / f"{name}.xml"
        counts = np.zeros(8, np.uint32)
        cam, tris, tmat = np.zeros(12), np.zeros((cap, 9)), np.zeros((cap, 15))
        mesh, sph, ssrc = np.zeros(cap, np.uint32), np.zeros((cap, 19)), np.zeros((cap, 7))
        with oracle._quiet():
            ok = R.ref_parse_scene(str(path).encode(), counts, cam, tris, tmat, mesh, sph, ssrc, cap)
        assert ok, name
        n, nm, nt, ns, nss = (int(c) for c in counts[:5])
        out.update({f"{name}_xml": np.array(path.read_text()), f"{name}_counts": counts,
                    f"{name}_cam": cam, f"{name}_tris": tris[:n], f"{name}_tmat": tmat[:n],
                    f"{name}_mesh": mesh[:nm + nt], f"{name}_spheres": sph[:ns],
                    f"{name}_sphere_sources": ssrc[:nss]})
    np.savez_compressed(OUT / "scene_xml.npz", **out)


def main():
    """python tests/golden/make_golden.py [fixture ...]: all fixtures, or the
    named ones (bvh_closest, shadow, brdf, ltc, spheres, direct, scene_xml, sources, env)."""
    oracle.build(ref=True)
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libchkref.so missing: the reference checkout is needed")
    want = set(sys.argv[1:])

    def on(name):
        return not want or name in want

    rng = np.random.default_rng(20261203)
    sc = gen_bvh(R, rng)  # the fixtures below share this generator's stream
    if on("shadow") or on("brdf") or on("ltc") or not want:
        gen_shadow(R, rng, sc)
        gen_brdf(R, rng)
        gen_ltc(R, rng)
    if on("spheres"):
        gen_spheres(R, np.random.default_rng(20261204))  # own stream: earlier fixtures unchanged
    if on("direct"):
        gen_direct()
    if on("scene_xml"):
        gen_scene_xml(R)
    if on("sources"):
        gen_sources()
    if on("env"):
        gen_env(R)
    for p in sorted(OUT.glob("*.npz")):
        print(p.name, p.stat().st_size)


if __name__ == "__main__":
    main()
