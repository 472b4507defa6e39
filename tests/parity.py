"""Helpers for the GPU parity tests: FP64 barycentric margins that define the
tie / grazing exclusion band for bit-exact closest-hit ids.

The reference test (geometryPrimitives.cpp:29-71) is FP64 and accepts
gamma, beta in [0, 1] with gamma + beta <= 1 and (d.n)^2 > 1e-10; the GPU
test is FP32 watertight.  Both agree on every ray whose winning hit is not
within 1e-5 (barycentric) of an edge, not within 1e-5 relative distance of
another triangle's hit, and not near the parallel-ray threshold.  Rays
inside that band are counted and reported, not compared.
"""
import numpy as np

import oracle
from chakumrender_amd import scenes

BAND = 1e-5


def bary_fp64(tris, ids, rays):
    """Reference-formula barycentrics (gamma, beta), t and (d.n)^2 of rays
    against the given triangles (FP64).  ids < 0 -> NaN."""
    ok = ids >= 0
    T = tris[np.where(ok, ids, 0)].reshape(-1, 3, 3).astype(np.float64)
    A, B, C = T[:, 0], T[:, 1], T[:, 2]
    o, d = rays[:, :3], rays[:, 3:6]
    n = np.cross(C - A, B - A)
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    dn = (d * n).sum(1)
    with np.errstate(all="ignore"):
        t = ((A - o) * n).sum(1) / dn
        P = o + d * t[:, None]
        v0, v1, v2 = P - A, B - A, C - A
        k0, k1 = (v0 * v1).sum(1), (v0 * v2).sum(1)
        b0, b1 = (v1 * v1).sum(1), (v1 * v2).sum(1)
        c0, c1 = b1, (v2 * v2).sum(1)
        gamma = (k0 * b1 - k1 * b0) / (c0 * b1 - c1 * b0)
        beta = (k0 - gamma * c0) / b0
    margin = np.minimum(np.minimum(gamma, beta), 1 - gamma - beta)
    margin[~ok] = np.nan
    return t, margin, dn * dn


def classify_closest(tris, rays, ref_id, ref_t, gpu_id, gpu_t, ref_t2=None):
    """Returns (excluded mask, mismatch mask among non-excluded)."""
    tris = np.asarray(tris, np.float64).reshape(-1, 9)
    _, m_ref, c_ref = bary_fp64(tris, ref_id, rays)
    tg, m_gpu, c_gpu = bary_fp64(tris, gpu_id, rays)
    band = np.zeros(rays.shape[0], bool)
    hit = ref_id >= 0
    band |= hit & (m_ref < BAND)
    band |= hit & (c_ref < 1e-10 * (1 + 1e-3))
    if ref_t2 is not None:
        band |= hit & (np.abs(ref_t2 - ref_t) <= BAND * np.maximum(ref_t, 1e-6))
    ghit = gpu_id >= 0
    # the GPU reports a triangle the reference rejects by less than the band
    band |= ghit & (gpu_id != ref_id) & (m_gpu > -BAND)
    band |= ghit & (gpu_id != ref_id) & hit & (np.abs(tg - ref_t) <= BAND * np.maximum(ref_t, 1e-6))
    band |= ghit & (c_gpu < 1e-10 * (1 + 1e-3))
    mism = (~band) & (ref_id != gpu_id)
    return band, mism


def trace_scene(tris):
    """A Scene holding only geometry (one Lambert material, no lights)."""
    tris = np.asarray(tris, np.float32).reshape(-1, 9)
    mats = np.zeros(1, scenes.MATERIAL_DTYPE)
    mats[0] = scenes.lambert_material((0.5, 0.5, 0.5))
    cam = scenes.Camera((0, 0, 5), (0, 0, -1), (0, 1, 0), 40, 40, 16, 16)
    return scenes.Scene(tris, np.zeros(tris.shape[0], np.uint32), mats,
                        np.zeros(0, scenes.LIGHT_DTYPE), cam, "geometry")


def random_rays(scene, rng, m):
    t = scene.tris.astype(np.float64).reshape(-1, 3, 3)
    finite = t[np.all(np.abs(t) < 1e5, axis=(1, 2))]
    lo, hi = finite.reshape(-1, 3).min(0), finite.reshape(-1, 3).max(0)
    a = rng.uniform(lo, hi, (m, 3))
    b = rng.uniform(lo, hi, (m, 3))
    d = b - a
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([a.astype(np.float32), d.astype(np.float32)], 1).astype(np.float64)


def oracle_closest_bvh(tris, rays):
    b = oracle.RefBvh(np.asarray(tris, np.float64)).closest(rays, prune=True)
    return b["id"], b["t"]
