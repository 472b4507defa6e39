"""Closest-hit and any-hit traversal on the GPU (chk_trace_closest /
chk_trace_any) against the reference's BvhAccel::intersectT and shadow().

Bar: triangle ids bit-exact (original indices) outside the tie/grazing band
of tests/parity.py, t within 1e-6 relative; occlusion bit-exact outside the
band around tmax."""
import ctypes as C

import numpy as np
import pytest

import oracle
from chakumrender_amd import abi, host, scenes
from parity import classify_closest, oracle_closest_bvh, random_rays, trace_scene

pytestmark = pytest.mark.gpu


def gpu_closest(ctx, scene, rays):
    acc = host.BvhAccel(ctx, scene)
    h = acc.closest_raw(host.make_rays(rays[:, :3], rays[:, 3:6]))
    info = acc.ds.info()
    return h[:, 1].view(np.int32).copy(), h[:, 0].astype(np.float64), h[:, 2:], info


def check_closest(tris, rays, ref_id, ref_t, gid, gt, ref_t2=None, max_band=0.005):
    band, mism = classify_closest(tris, rays, ref_id, ref_t, gid, gt, ref_t2)
    print(f"rays {rays.shape[0]}, hits {np.mean(ref_id >= 0):.3f}, excluded {band.sum()} "
          f"({band.mean():.2e}), mismatches {mism.sum()}")
    assert mism.sum() == 0, np.nonzero(mism)[0][:10]
    assert band.mean() < max_band
    same = (~band) & (ref_id >= 0)
    rel = np.abs(gt[same] - ref_t[same]) / ref_t[same]
    assert rel.max() < 1e-6, rel.max()
    miss = (~band) & (ref_id < 0)
    assert np.all(gt[miss] >= 1e29)


@pytest.mark.parametrize("name", ["cornell", "ico"])
def test_closest_matches_reference_golden(ctx, golden, name):
    G = golden("bvh_closest.npz")
    tris, rays = G[f"{name}_tris"], G[f"{name}_rays"]
    gid, gt, _, info = gpu_closest(ctx, trace_scene(tris), rays)
    # the reference's own t / id; the band needs the second-nearest t (FP64 oracle)
    b = oracle.closest_brute(tris.astype(np.float64), rays)
    assert np.array_equal(b["id"], G[f"{name}_id"])
    check_closest(tris, rays, G[f"{name}_id"], G[f"{name}_t"], gid, gt, b["t2"])
    assert info["n_tris_bvh"] == tris.shape[0] - 1  # the degenerate dummy triangle is dropped


def test_closest_100k_icosphere_vs_reference_bvh(ctx):
    sc = scenes.icosphere_scene(freq=71, n_lights=8, width=64, height=64)
    assert sc.n_tris == 100_820 + 2 + 1 + 16
    rng = np.random.default_rng(11)
    rays = random_rays(sc, rng, 1 << 16)
    # camera rays too
    pos, d, cx, cy = sc.camera.basis()
    u = rng.uniform(-0.5, 0.5, (1 << 15, 2))
    cd = cx * u[:, :1] + cy * u[:, 1:] + d
    cd /= np.linalg.norm(cd, axis=1, keepdims=True)
    cam = np.concatenate([np.repeat(pos[None], u.shape[0], 0), cd], 1).astype(np.float32).astype(np.float64)
    rays = np.concatenate([rays, cam])
    gid, gt, _, info = gpu_closest(ctx, sc, rays)
    rid, rt = oracle_closest_bvh(sc.tris, rays)
    check_closest(sc.tris, rays, rid, rt, gid, gt)
    assert info["stack_bound"] <= info["stack_size"]


def test_skip_and_tmax(ctx, golden):
    """tmax bounds the reported hit; skip removes one original triangle."""
    G = golden("bvh_closest.npz")
    tris, rays = G["cornell_tris"], G["cornell_rays"][:2048]
    acc = host.BvhAccel(ctx, trace_scene(tris))
    ref = oracle.closest_brute(tris.astype(np.float64), rays)
    hit = ref["id"] >= 0
    skip = np.where(hit, ref["id"], -1)
    h = acc.closest_raw(host.make_rays(rays[:, :3], rays[:, 3:6], skip=skip))
    gid = h[:, 1].view(np.int32)
    assert not np.any(gid[hit] == skip[hit])
    # tmax just below the first hit: nothing is reported
    h2 = acc.closest_raw(host.make_rays(rays[:, :3], rays[:, 3:6],
                                        tmax=np.where(hit, ref["t"] * 0.999, 1e30)))
    g2 = h2[:, 1].view(np.int32)
    assert np.all(g2[hit] == -1)


def test_any_hit_matches_reference_shadow(ctx, golden):
    S = golden("shadow.npz")
    tris, r7 = S["tris"], S["rays7"]
    acc = host.BvhAccel(ctx, trace_scene(tris))
    vis = acc.shadow(r7[:, :3], r7[:, 3:6], r7[:, 6])
    occ = (vis < 0.5).astype(np.uint8)
    # band: the closest FP64 hit within 1e-5 of the shadow limit, or grazing
    b = oracle.closest_brute(tris.astype(np.float64), r7[:, :6])
    lim = r7[:, 6] - 1e-5
    band = (b["id"] >= 0) & ((np.abs(b["t"] - lim) <= 1e-5 * np.maximum(1, lim)) | (b["margin"] < 1e-5)
                             | (b["cos2"] < 1.001e-10))
    mism = (~band) & (occ != S["occluded"])
    print(f"shadow rays {r7.shape[0]}, occluded {S['occluded'].mean():.3f}, excluded {band.sum()}")
    assert mism.sum() == 0, np.nonzero(mism)[0][:10]


def test_any_hit_consistent_with_closest_100k(ctx):
    """Occlusion with tmax just above / below the closest hit is 1 / 0."""
    sc = scenes.icosphere_scene(freq=71, n_lights=8, width=64, height=64)
    rays = random_rays(sc, np.random.default_rng(5), 1 << 15)
    gid, gt, _, _ = gpu_closest(ctx, sc, rays)
    acc = host.BvhAccel(ctx, sc)
    hit = gid >= 0
    above = acc.shadow(rays[:, :3], rays[:, 3:6], np.where(hit, gt * 1.0001 + 1e-5, 1e6))
    below = acc.shadow(rays[:, :3], rays[:, 3:6], np.where(hit, gt * 0.9999 + 1e-5, 1e6))
    assert np.all(above[hit] == 0.0)
    assert np.all(below[hit] == 1.0)
    assert np.all(above[~hit] == 1.0)


def test_empty_and_degenerate_inputs(ctx):
    # zero rays is a no-op
    acc = host.BvhAccel(ctx, trace_scene(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)))
    abi.check(acc.L.chk_trace_closest(acc.ds.h, None, 0, None, None), "zero rays")
    # a scene of only degenerate triangles: every ray misses
    deg = np.tile(np.array([[1e6, 1e6, 1e6] * 3], np.float32), (3, 1))
    acc2 = host.BvhAccel(ctx, trace_scene(deg))
    hit, tt, ids = acc2.intersectT(np.zeros((4, 3)), np.tile([[0, 0, 1.0]], (4, 1)))
    assert not hit.any() and np.all(ids == -1)
    assert acc2.ds.info()["n_tris_bvh"] == 0
    # a ray along a triangle's plane (d.n = 0) never hits it (geometryPrimitives.cpp:42)
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    acc3 = host.BvhAccel(ctx, trace_scene(tri))
    hit, _, _ = acc3.intersectT([[-1, 0.2, 0.0]], [[1, 0, 0]])
    assert not hit.any()
    hit, tt, ids = acc3.intersectT([[0.2, 0.2, 1.0]], [[0, 0, -1]])
    assert hit.all() and ids[0] == 0 and abs(tt[0] - 1.0) < 1e-6


def test_spheres_match_reference_intersectS(golden):
    """chk_trace_spheres against the compiled BvhAccel::intersectS: the GPU gets
    the FP32 rounding of the unit directions, so ids must agree with the
    reference except where the nearest two roots (or a root and eps) are within
    1e-5 relative.  That rounding moves t of near-tangent rays by more than
    1e-6 (sqrt(det) amplifies it), so t is held to 1e-6 relative against the
    oracle restatement (pinned to the same golden vectors on the CPU) run on
    exactly the FP32-rounded rays the GPU sees, where ids must also agree."""
    G = golden("spheres.npz")
    sp, rays = G["spheres"], G["rays"]
    r = host.make_rays(rays[:, :3], rays[:, 3:], tmax=np.full(rays.shape[0], np.inf, np.float32))
    h = host.intersect_spheres(sp, r)
    gid = h[:, 1].view(np.uint32).astype(np.int64)
    gid[gid == 0xFFFFFFFF] = -1
    # tie band from the oracle: roots of every sphere per ray
    o, d = rays[:, :3], rays[:, 3:]
    roots = np.full((rays.shape[0], sp.shape[0]), np.inf)
    for k in range(sp.shape[0]):
        op = sp[k, :3] - o
        b = (op * d).sum(1)
        det = b * b - (op * op).sum(1) + sp[k, 3] ** 2
        sq = np.sqrt(np.maximum(det, 0))
        t1, t2 = b - sq, b + sq
        t = np.where(t1 > 1e-4, t1, np.where(t2 > 1e-4, t2, np.inf))
        roots[:, k] = np.where(det >= 0, t, np.inf)
    srt = np.sort(roots, 1)
    near_tie = (np.abs(srt[:, 1] - srt[:, 0]) <= 1e-5 * np.abs(srt[:, 0])) | \
               (np.abs(srt[:, 0] - 1e-4) <= 1e-5)
    ok = ~near_tie
    mism = (gid != G["id"]) & ok
    print(f"rays {rays.shape[0]}, hits {(G['id'] >= 0).mean():.3f}, excluded {(~ok).sum()}, "
          f"mismatches {mism.sum()}")
    assert mism.sum() == 0
    oid, ot = oracle.spheres_brute(sp, rays.astype(np.float32).astype(np.float64))
    assert np.array_equal(gid[ok], oid[ok])
    hit = (oid >= 0) & ok
    np.testing.assert_allclose(h[hit, 0], ot[hit], rtol=1e-6)
    assert np.all(h[oid < 0, 0] >= 1e29)
    # against the reference's t on its exact unit directions, the input rounding
    # alone bounds the difference away from tangency
    far = (G["id"] >= 0) & ok & (np.abs(h[:, 0] - G["t"]) > 1e-6 * G["t"])
    print(f"t beyond 1e-6 of the reference from direction rounding: {far.sum()}")
    assert far.mean() < 0.01


@pytest.mark.parametrize("name", ["cornell", "ico"])
def test_dummy_accel_equals_bvh_and_reference(ctx, golden, name):
    """CHK_ACCEL_DUMMY (the reference's DummyAccel: sqrt(n) runs of triangles
    in original order, each with its box) returns the BVH's hits: ids bit
    for bit and the same FP64-refined t, against the reference's golden
    closest hits as the BVH test does; occlusion equal to the BVH's."""
    G = golden("bvh_closest.npz")
    tris, rays = G[f"{name}_tris"], G[f"{name}_rays"]
    sc = trace_scene(tris)
    acc = host.BvhAccel(ctx, sc)
    r = host.make_rays(rays[:, :3], rays[:, 3:6])
    hb = acc.closest_raw(r)
    acc.ds.set_accel(abi.ACCEL_DUMMY)
    print(f"{name}: dummy accel over {tris.shape[0]} triangles, {rays.shape[0]} rays", flush=True)
    hd = acc.closest_raw(r)
    gid = hd[:, 1].view(np.int32).copy()
    b = oracle.closest_brute(tris.astype(np.float64), rays)
    check_closest(tris, rays, G[f"{name}_id"], G[f"{name}_t"], gid, hd[:, 0].astype(np.float64), b["t2"])
    # equal to the BVH's hits outside the tie band (of two equally near
    # triangles DummyAccel keeps the first in the original order)
    band, _ = classify_closest(tris, rays, G[f"{name}_id"], G[f"{name}_t"], gid,
                               hd[:, 0].astype(np.float64), b["t2"])
    np.testing.assert_array_equal(gid[~band], hb[~band, 1].view(np.int32))
    np.testing.assert_array_equal(hd[~band, 0], hb[~band, 0])
    # occlusion: segments to the closest hit (shortened) and beyond it
    t = np.where(hb[:, 0] < 1e29, hb[:, 0], 10.0).astype(np.float64)
    for scale in (0.5, 2.0):
        sh = host.make_rays(rays[:, :3], rays[:, 3:6], tmax=t * scale)
        od = acc.occluded_raw(sh)
        acc.ds.set_accel(abi.ACCEL_BVH)
        ob = acc.occluded_raw(sh)
        acc.ds.set_accel(abi.ACCEL_DUMMY)
        np.testing.assert_array_equal(od, ob)
    with pytest.raises(abi.ChkError, match="kind"):
        acc.ds.set_accel(7)
