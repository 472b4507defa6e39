"""Pins of the FP64 oracle that do not come from the reference's code:
published known-answer vectors, closed forms, invariants and independent
quadratures, chosen so that a dropped term, wrong sign or transposed operand
fails at least one of them."""
import math

import numpy as np
import pytest

import oracle
from chakumrender_amd import abi, scenes
from conftest import unit


# ------------------------------------------------------------------ Philox
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expect):
    """Random123 known-answer vectors for philox4x32-10."""
    assert list(oracle.philox(ctr, key)) == expect


# ------------------------------------------------------------------ geometry
def test_triangle_closed_form():
    A, B, C = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    rng = np.random.default_rng(0)
    for _ in range(2000):
        x, y, h = rng.uniform(-0.5, 1.5), rng.uniform(-0.5, 1.5), rng.uniform(0.1, 3)
        N = np.empty(3)
        t = oracle.lib().orc_tri_intersect(A, B, C, np.array([x, y, h]), np.array([0.0, 0, -1]), N)
        inside = x >= 0 and y >= 0 and x + y <= 1
        if inside:
            assert abs(t - h) < 1e-12 and np.allclose(N, [0, 0, 1])
        else:
            assert t == 1e20
    # edges are inside (closed triangle), behind the origin is a miss, and a
    # ray in the triangle's plane is rejected by the (d.n)^2 <= 1e-10 test
    N = np.empty(3)
    lib = oracle.lib()
    assert lib.orc_tri_intersect(A, B, C, np.array([0.5, 0.5, 1.0]), np.array([0.0, 0, -1]), N) == 1.0
    assert lib.orc_tri_intersect(A, B, C, np.array([0.2, 0.2, -1.0]), np.array([0.0, 0, -1]), N) == 1e20
    d = unit(np.array([1.0, 0.0, 1e-6]))
    assert lib.orc_tri_intersect(A, B, C, np.array([-1.0, 0.2, -1e-6]), d, N) == 1e20


def test_aabb_closed_form():
    lib = oracle.lib()
    lo, hi = np.array([0.0, 0, 0]), np.array([1.0, 2, 3])
    assert lib.orc_aabb_intersect(lo, hi, np.array([-1.0, 1, 1]), np.array([1.0, 0, 0])) == 1.0
    # box behind the origin still reports its (negative) entry t (geometryPrimitives.cpp:107-139)
    assert lib.orc_aabb_intersect(lo, hi, np.array([3.0, 1, 1]), np.array([1.0, 0, 0])) == -3.0
    assert lib.orc_aabb_intersect(lo, hi, np.array([-1.0, 5, 1]), np.array([1.0, 0, 0])) == 1e20


def test_bvh_equals_brute_force_on_random_soup():
    rng = np.random.default_rng(4)
    c = rng.uniform(-1, 1, (3000, 1, 3))
    tris = (c + rng.normal(scale=0.05, size=(3000, 3, 3))).reshape(-1, 9)
    o = rng.uniform(-1.5, 1.5, (4000, 3))
    d = unit(rng.normal(size=(4000, 3)))
    rays = np.concatenate([o, d], 1)
    b = oracle.closest_brute(tris, rays)
    B = oracle.RefBvh(tris)
    for prune in (False, True):
        r = B.closest(rays, prune)
        ok = (b["id"] < 0) | (np.abs(b["t2"] - b["t"]) > 1e-9)
        assert np.array_equal(r["id"][ok], b["id"][ok])


# ------------------------------------------------------------------ materials
def material(bt, ks, alpha=0.3, pe=20.0, refl=1.0, iior=1.5, eior=1.0):
    m = np.zeros(1, scenes.MATERIAL_DTYPE)
    m["brdf"], m["ks"], m["alpha"], m["phong_exp"] = bt, ks, alpha, pe
    m["int_ior"], m["ext_ior"], m["reflectance"] = iior, eior, refl
    return m


def hemi(rng, n, k):
    v = unit(rng.normal(size=(k, 3)))
    return np.where((v * n).sum(1, keepdims=True) < 0, -v, v)


@pytest.mark.parametrize("bt,alpha", [(scenes.GGX, 0.3), (scenes.BECKMANN, 0.4), (scenes.PHONG, 30.0),
                                      (scenes.CLASSIC_PHONG, 0.0)])
def test_brdf_reciprocity(bt, alpha):
    rng = np.random.default_rng(1)
    k = 2000
    n = unit(rng.normal(size=(k, 3)))
    wo, wi = hemi(rng, n, k), hemi(rng, n, k)
    refl = lambda w: unit(2 * (n * w).sum(1, keepdims=True) * n - w)  # noqa: E731
    m = material(bt, 0.7, alpha)
    idx = np.zeros(k, np.uint32)
    f1, _ = oracle.brdf_eval(m, idx, n, refl(wo), wo, wi)
    f2, _ = oracle.brdf_eval(m, idx, n, refl(wi), wi, wo)
    ok = np.isfinite(f1) & np.isfinite(f2)
    np.testing.assert_allclose(f1[ok], f2[ok], rtol=1e-9)


def test_lambert_white_furnace():
    """ks = 0: f = reflectance/pi, pdf = cos/pi; E[f cos/pdf] = reflectance."""
    rng = np.random.default_rng(2)
    k = 4000
    n = unit(rng.normal(size=(k, 3)))
    wo = hemi(rng, n, k)
    wr = unit(2 * (n * wo).sum(1, keepdims=True) * n - wo)
    m = material(scenes.CLASSIC_PHONG, 0.0, refl=0.7)
    u = rng.uniform(size=(k, 3))
    wi, pdf = oracle.brdf_sample(m, np.zeros(k, np.uint32), n, wr, wo, u)
    f, p2 = oracle.brdf_eval(m, np.zeros(k, np.uint32), n, wr, wo, wi)
    cos = (wi * n).sum(1)
    r = float(np.float32(0.7))  # the material record holds float32
    np.testing.assert_allclose(f, r / 3.14159265359, rtol=1e-12)
    np.testing.assert_allclose(pdf, cos / 3.14159265359, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(f * cos / pdf, r, rtol=1e-9)


@pytest.mark.parametrize("bt,alpha,ks", [(scenes.GGX, 0.35, 1.0), (scenes.BECKMANN, 0.3, 1.0),
                                         (scenes.PHONG, 25.0, 1.0), (scenes.CLASSIC_PHONG, 0.0, 1.0),
                                         (scenes.GGX, 0.5, 0.5)])
def test_sampler_histogram_matches_pdf(bt, alpha, ks):
    """Chi-square of 200k sampled directions against pdfEval integrated over
    each (cos theta, phi) bin of the upper hemisphere.  A wrong Jacobian
    (e.g. dropping 1/(4 h.wo)) or a wrong tan(theta) law fails it."""
    rng = np.random.default_rng(11)
    n0 = np.array([0.0, 0.0, 1.0])
    wo0 = unit(np.array([0.3, 0.1, 1.0]))
    wr0 = unit(2 * wo0.dot(n0) * n0 - wo0)
    m = material(bt, ks, alpha, pe=15.0)
    N = 200_000
    rep = lambda v: np.repeat(v[None], N, 0)  # noqa: E731
    u = rng.uniform(size=(N, 3))
    wi, _ = oracle.brdf_sample(m, np.zeros(N, np.uint32), rep(n0), rep(wr0), rep(wo0), u)
    up = wi[:, 2] > 0
    nb_c, nb_p = 12, 12
    ci = np.minimum((wi[up, 2] * nb_c).astype(int), nb_c - 1)
    phi = np.arctan2(wi[up, 1], wi[up, 0]) % (2 * np.pi)
    pi_ = np.minimum((phi / (2 * np.pi) * nb_p).astype(int), nb_p - 1)
    counts = np.bincount(ci * nb_p + pi_, minlength=nb_c * nb_p).astype(float)
    # expected: midpoint quadrature of pdfEval in (cos theta, phi); d omega = dcos dphi
    s = 10
    cc = (np.arange(nb_c * s) + 0.5) / (nb_c * s)
    pp = (np.arange(nb_p * s) + 0.5) / (nb_p * s) * 2 * np.pi
    C, P = np.meshgrid(cc, pp, indexing="ij")
    S = np.sqrt(1 - C * C)
    W = np.stack([S * np.cos(P), S * np.sin(P), C], -1).reshape(-1, 3)
    K = W.shape[0]
    _, pdf = oracle.brdf_eval(m, np.zeros(K, np.uint32), np.repeat(n0[None], K, 0),
                              np.repeat(wr0[None], K, 0), np.repeat(wo0[None], K, 0), W)
    cell = (1.0 / (nb_c * s)) * (2 * np.pi / (nb_p * s))
    exp_cells = (pdf * cell).reshape(nb_c, s, nb_p, s).sum((1, 3)).reshape(-1) * N
    ok = exp_cells > 20
    chi2 = ((counts[ok] - exp_cells[ok]) ** 2 / exp_cells[ok]).sum()
    dof = ok.sum()
    # loose bound: mean dof, sd sqrt(2 dof); quadrature error adds a little
    assert chi2 < dof + 6 * math.sqrt(2 * dof) + 0.002 * N ** 0.5 * dof ** 0.5, (chi2, dof)
    assert abs(counts.sum() - exp_cells.sum()) / N < 0.02


# ------------------------------------------------------------------ polygon integrals
def corner_ff(a, b, h):
    """Form factor from a differential area to a parallel rectangle with one
    corner above it (Howell's catalogue): odd in a and b."""
    A, B = a / h, b / h
    sa, sb = math.sqrt(1 + A * A), math.sqrt(1 + B * B)
    return (A / sa * math.atan(B / sa) + B / sb * math.atan(A / sb)) / (2 * math.pi)


def rect_ff(x, z, h, x1, x2, z1, z2):
    return (corner_ff(x2 - x, z2 - z, h) - corner_ff(x1 - x, z2 - z, h)
            - corner_ff(x2 - x, z1 - z, h) + corner_ff(x1 - x, z1 - z, h))


def test_diffuse_polygon_integral_closed_form():
    """ks = 0: Analytic = radiance * form factor of the light polygon."""
    rng = np.random.default_rng(6)
    q = np.array([[[-0.3, 1.0, -0.2], [0.4, 1.0, -0.2], [0.4, 1.0, 0.5], [-0.3, 1.0, 0.5]]])
    tris, rad = scenes.quads_to_tris(q, [[2.0, 2.0, 2.0]])
    k = 300
    x = np.c_[rng.uniform(-1, 1, k), np.zeros(k), rng.uniform(-1, 1, k)]
    n = np.tile([0.0, 1.0, 0.0], (k, 1))
    v = unit(rng.normal(size=(k, 3)) + [0, 2, 0])
    o = oracle.ltc_shade(x, n, v, np.full(k, 0.3), np.zeros(k), tris, rad)
    ff = np.array([rect_ff(a, c, 1.0, -0.3, 0.4, -0.2, 0.5) for a, c in zip(x[:, 0], x[:, 2])])
    np.testing.assert_allclose(o["rgb"][:, 0], 2.0 * ff, rtol=2e-7, atol=1e-12)


def test_ltc_closed_box_integrates_to_one():
    """A closed, rotated box around the point: the clipped cosine-domain
    integral is exactly the hemisphere, 1, for the LTC and diffuse parts."""
    rng = np.random.default_rng(9)
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    c = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float) @ R.T * 3.0
    faces = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]
    tris, rad = scenes.quads_to_tris(np.array([[c[i] for i in f] for f in faces]), np.ones((6, 3)))
    k = 256
    x = rng.uniform(-0.5, 0.5, (k, 3))
    n = unit(rng.normal(size=(k, 3)))
    v = hemi(rng, n, k)
    o = oracle.ltc_shade(x, n, v, rng.uniform(0.05, 1, k), np.full(k, 0.5), tris, rad)
    np.testing.assert_allclose(o["ltc"][:, 0], 1.0, rtol=1e-10)
    np.testing.assert_allclose(o["diffuse"][:, 0], 1.0, rtol=1e-6)


def test_ltc_integral_matches_quadrature_of_ltc_distribution():
    """Away from interpolation (theta on a bin edge), the analytic LTC
    integral equals a Monte-Carlo integral of approxLTC_BRDF's cosine-lobe
    density over the light's solid angle (independent formula: M, det M)."""
    rng = np.random.default_rng(12)
    q = np.array([[[-0.6, 1.0, -0.5], [0.5, 1.2, -0.5], [0.5, 1.2, 0.6], [-0.6, 1.0, 0.6]]])
    tris, rad = scenes.quads_to_tris(q, [[1.0, 1.0, 1.0]])
    x = np.array([[0.1, 0.0, 0.05]])
    nrm = np.array([[0.0, 1.0, 0.0]])
    th = np.float32(20 / 64 * np.pi / 2)  # exactly a theta bin edge -> no lerp
    v = np.array([[math.sin(th), math.cos(th), 0.0]])
    alpha = np.array([0.35])
    o = oracle.ltc_shade(x, nrm, v, alpha, np.ones(1), tris, rad)
    # MC over the light area: integral of D(w) dw = sum D(w) cos_l / r^2 * area
    N = 400_000
    P = np.empty((N, 3))
    T = tris.reshape(-1, 3, 3)
    areas = np.array([0.5 * np.linalg.norm(np.cross(t[2] - t[0], t[1] - t[0])) for t in T])
    pick = rng.choice(2, N, p=areas / areas.sum())
    r1, r2 = np.sqrt(rng.uniform(size=N)), rng.uniform(size=N)
    P = T[pick, 0] * (1 - r1)[:, None] + T[pick, 1] * (r1 * (1 - r2))[:, None] + T[pick, 2] * (r1 * r2)[:, None]
    dvec = P - x
    r = np.linalg.norm(dvec, axis=1)
    w = dvec / r[:, None]
    nl = unit(np.cross(T[pick, 2] - T[pick, 0], T[pick, 1] - T[pick, 0]))
    cl = np.abs((w * nl).sum(1))
    D = oracle.ltc_brdf_cos(np.repeat(nrm, N, 0), np.repeat(v, N, 0), w, np.repeat(alpha, N))
    est = (D * cl / r ** 2).mean() * areas.sum()
    amp = o["amp"][0]
    np.testing.assert_allclose(o["ltc"][0, 0] * amp, est, rtol=0.02)


# ------------------------------------------------------------------ integrator
def plane_and_light(width=24, height=16):
    """A Lambert ground (y = 0, albedo 0.5) under one square light (y = 1,
    radiance 10), seen by a camera looking down at it from the side."""
    mats = [scenes.lambert_material((0.5, 0.5, 0.5))]
    ground = scenes.quad((0, 0, 0), (3, 0, 0), (0, 0, 3))
    light = (scenes.quad((0.1, 1.0, 0.0), (0.3, 0, 0), (0, 0, 0.3)), (1, 1, 1), 10.0)
    cam = scenes.Camera((0.0, 0.7, 2.5), (0.0, -0.45, -1.0), (0, 1, 0), 40.0, 40.0, width, height)
    return scenes._assemble(ground, np.zeros(2, np.uint32), mats, [light], cam, "plane", {})


def plane_expectation(sc):
    """Per-pixel closed form for max_depth 1 with the fixed reference
    jitter: every sample of a pixel hits the same ground point x, whose
    radiance is albedo * L * F(x) (F: form factor to the light square)."""
    pos, d, cx, cy = sc.camera.basis()
    W, H = sc.camera.width, sc.camera.height
    exp = np.full((H, W), np.nan)
    for row in range(H):
        y = H - 1 - row
        for x in range(W):
            dirv = cx * ((x + 0.34) / W - 0.5) + cy * ((y + 0.20) / H - 0.5) + d
            dirv = dirv / np.linalg.norm(dirv)
            if dirv[1] >= 0:
                continue
            t = -pos[1] / dirv[1]
            p = pos + dirv * t
            if abs(p[0]) > 2.9 or abs(p[2]) > 2.9:
                continue
            # skip rays that pass near the light quad (it occludes/emits)
            tl = (1.0 - pos[1]) / dirv[1]
            if tl > 0:
                pl = pos + dirv * tl
                if abs(pl[0] - 0.1) < 0.35 and abs(pl[2]) < 0.35:
                    continue
            exp[row, x] = 0.5 * 10.0 * rect_ff(p[0], p[2], 1.0, -0.2, 0.4, -0.3, 0.3)
    return exp


def test_integrator_direct_light_closed_form():
    sc = plane_and_light()
    spp = 512
    P = dict(spp=spp, max_depth=1, seed=21, jitter_mode=abi.JITTER_REFERENCE)
    acc, sq, paths, _ = oracle.render(sc, P, sumsq=True)
    mean = acc[..., 0] / spp
    var = np.maximum(sq[..., 0] / spp - mean ** 2, 0) / spp
    exp = plane_expectation(sc)
    ok = np.isfinite(exp)
    assert ok.sum() > 100
    z = (mean[ok] - exp[ok]) / np.sqrt(var[ok] + 1e-30)
    assert np.mean(np.abs(z) < 4) > 0.99, np.abs(z).max()
    assert abs(mean[ok].sum() / exp[ok].sum() - 1) < 0.01
