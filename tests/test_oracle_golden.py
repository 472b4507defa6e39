"""The FP64 oracle against golden vectors written by the reference's own
compiled code (tests/golden/make_golden.py), and live against that code when
oracle/_ref/libchkref.so was built here."""
import numpy as np
import pytest

import oracle
from chakumrender_amd import scenes


def brdf_materials(cases):
    mats = np.zeros(len(cases), scenes.MATERIAL_DTYPE)
    for i, (bt, ks, al, pe, ii, ei, refl) in enumerate(cases):
        mats[i]["brdf"] = bt
        mats[i]["ks"] = ks
        mats[i]["alpha"] = al
        mats[i]["phong_exp"] = pe
        mats[i]["int_ior"] = ii
        mats[i]["ext_ior"] = ei
        mats[i]["reflectance"] = refl
    return mats


@pytest.mark.parametrize("name", ["cornell", "ico"])
def test_closest_hit_ids_match_reference_bvh(golden, name):
    G = golden("bvh_closest.npz")
    tris, rays = G[f"{name}_tris"].astype(np.float64), G[f"{name}_rays"]
    gid, gt, gN = G[f"{name}_id"], G[f"{name}_t"], G[f"{name}_N"]
    b = oracle.closest_brute(tris, rays)
    assert np.array_equal(b["id"], gid)
    hit = gid >= 0
    assert hit.mean() > 0.5
    np.testing.assert_allclose(b["t"][hit], gt[hit], rtol=1e-12)
    np.testing.assert_allclose(b["N"][hit], gN[hit], atol=1e-12)
    # restated reference BVH (midpoint split, std::partition) agrees as well
    bv = oracle.RefBvh(tris).closest(rays)
    assert np.array_equal(bv["id"], gid)
    np.testing.assert_allclose(bv["t"][hit], gt[hit], rtol=1e-12)
    bp = oracle.RefBvh(tris).closest(rays, prune=True)
    assert np.array_equal(bp["id"], gid)
    assert bp["nv"].sum() < bv["nv"].sum()


def test_reference_bvh_shape(golden):
    """2N-1 nodes with one primitive per leaf (bvhAccel.cpp:9,43)."""
    G = golden("bvh_closest.npz")
    tris = G["cornell_tris"].astype(np.float64)
    B = oracle.RefBvh(tris)
    n = tris.shape[0]
    assert B.node_count <= 2 * n - 1
    o = B.ordered()
    assert sorted(o.tolist()) == list(range(n))


def test_shadow_matches_reference(golden):
    S = golden("shadow.npz")
    occ = oracle.occluded_brute(S["tris"].astype(np.float64), S["rays7"])
    assert np.array_equal(occ, S["occluded"])
    assert 0.2 < S["occluded"].mean() < 0.8


def test_brdf_pdf_match_reference(golden):
    B = golden("brdf.npz")
    mats = brdf_materials(B["cases"])
    f, p = oracle.brdf_eval(mats, B["case"].astype(np.uint32), B["n"], B["wo_ref"], B["wo"], B["wi"])
    np.testing.assert_allclose(f, B["f"], rtol=1e-11, atol=0)
    np.testing.assert_allclose(p, B["pdf"], rtol=1e-11, atol=0)
    # every BRDF type of the configs is covered
    assert set(B["cases"][:, 0].astype(int)) == {scenes.CLASSIC_PHONG, scenes.GGX, scenes.BECKMANN,
                                                 scenes.PHONG}


def test_M_GGX_matches_reference(golden):
    L = golden("ltc.npz")
    for i in range(L["theta"].size):
        M, Minv, amp, det = oracle.M_GGX(L["theta"][i], L["alpha_m"][i])
        np.testing.assert_allclose(M, L["M"][i], rtol=1e-14, atol=1e-13)
        np.testing.assert_allclose(Minv, L["Minv"][i], rtol=1e-14, atol=1e-11)
        assert abs(amp - L["amp"][i]) <= 1e-15
        assert abs(det - L["det"][i]) <= 1e-14 * max(1.0, abs(det))


def test_M_GGX_alpha_bin_uses_double_sqrt(golden):
    """ltc.h:20 bins floor(sqrt(alpha)*64): the golden set holds alphas within
    one float ulp of every bin edge, where float and double sqrt disagree;
    the oracle's double evaluation must reproduce the reference there."""
    L = golden("ltc.npz")
    al = L["alpha_m"].astype(np.float32)
    bf = np.floor(np.sqrt(al) * np.float32(64)).astype(int)
    bd = np.floor(np.sqrt(al.astype(np.float64)) * 64).astype(int)
    assert (bf != bd).sum() >= 10


@pytest.mark.parametrize("suffix", ["", "2"])
def test_ltc_analytic_matches_reference(golden, suffix):
    L = golden("ltc.npz")
    o = oracle.ltc_shade(L["x" + suffix], L["n" + suffix], L["v" + suffix], L["alpha" + suffix],
                         L["ks" + suffix], L["tris" + suffix], L["rad" + suffix])
    g = L["rgb" + suffix]
    scale = np.maximum(np.abs(g).max(1), 1e-12)
    err = np.abs(o["rgb"] - g).max(1) / scale
    # the axis-aligned light ("2") puts clip points on A.x == C.x edges where
    # FP64 rounding order decides the reference's x-only test
    assert err.max() < (1e-10 if suffix == "" else 1e-6)
    assert (np.abs(g).max(1) > 0).mean() > 0.5


@pytest.mark.skipif(oracle.ref_lib() is None, reason="reference library not built here")
def test_live_reference_brdf_random_materials():
    R = oracle.ref_lib()
    rng = np.random.default_rng(7)
    for bt in (scenes.GGX, scenes.BECKMANN, scenes.PHONG, scenes.CLASSIC_PHONG):
        ks = float(np.float32(rng.uniform(0, 1)))
        al = float(np.float32(rng.uniform(0.05, 0.8) if bt != scenes.PHONG else rng.uniform(1, 300)))
        pe = float(np.float32(rng.uniform(1, 100)))
        k = 300
        n = rng.normal(size=(k, 3))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        wo = rng.normal(size=(k, 3))
        wo /= np.linalg.norm(wo, axis=1, keepdims=True)
        wo = np.where((wo * n).sum(1, keepdims=True) < 0, -wo, wo)
        wi = rng.normal(size=(k, 3))
        wi /= np.linalg.norm(wi, axis=1, keepdims=True)
        wi = np.where((wi * n).sum(1, keepdims=True) < 0, -wi, wi)
        wr = 2 * (n * wo).sum(1, keepdims=True) * n - wo
        f = np.empty(k)
        p = np.empty(k)
        R.ref_brdf(bt, ks, al, pe, 1.5, 1.0, 1.0, k, *[np.ascontiguousarray(a) for a in (n, wr, wo, wi)], f, p)
        mats = brdf_materials([(bt, ks, al, pe, 1.5, 1.0, 1.0)])
        fo, po = oracle.brdf_eval(mats, np.zeros(k, np.uint32), n, wr, wo, wi)
        ok = np.isfinite(f)
        np.testing.assert_allclose(fo[ok], f[ok], rtol=1e-10)
        np.testing.assert_allclose(po[ok], p[ok], rtol=1e-10)


def test_spheres_match_reference_intersectS(golden):
    """orc_spheres_brute against the compiled BvhAccel::intersectS over Sphere
    primitives (same FP64 inputs): ids equal, t to 1e-12."""
    G = golden("spheres.npz")
    ids, t = oracle.spheres_brute(G["spheres"], G["rays"])
    np.testing.assert_array_equal(ids, G["id"])
    hit = G["id"] >= 0
    assert 0.2 < hit.mean() < 0.95
    np.testing.assert_allclose(t[hit], G["t"][hit], rtol=1e-12)


def test_sphere_closed_form():
    """A ray along an axis through the centre hits at |c - o| - r from
    outside and at r - |c - o| from inside; roots at or below eps = 1e-4 are
    skipped (geometryPrimitives.cpp:26)."""
    s = np.array([1.0, 2.0, 3.0, 0.5])
    assert oracle.sphere_intersect(s, [1.0, 2.0, 0.0], [0, 0, 1.0]) == pytest.approx(2.5, rel=1e-15)
    assert oracle.sphere_intersect(s, [1.0, 2.0, 2.9], [0, 0, 1.0]) == pytest.approx(0.6, rel=1e-14)
    assert oracle.sphere_intersect(s, [1.0, 2.0, 0.0], [0, 0, -1.0]) == 1e20
    # origin on the surface: the root at t = 0 is below eps, the far one counts
    assert oracle.sphere_intersect(s, [1.0, 2.0, 2.5], [0, 0, 1.0]) == pytest.approx(1.0, rel=1e-14)
