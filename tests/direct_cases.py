"""Shared pieces of the direct-lighting tests: the fixture's scene, estimate
means, and the statistical comparison used throughout."""
from __future__ import annotations

import numpy as np

from chakumrender_amd import abi, host, scenes


def fixture_scene(G):
    """The scene stored in tests/golden/direct.npz."""
    cam = scenes.direct_room().camera
    return scenes.Scene(np.asarray(G["tris"], np.float32), np.asarray(G["tri_mat"], np.uint32),
                        np.asarray(G["mats"]), np.asarray(G["lights"]), cam, "direct_room")


def sources_scene(G, prefix=""):
    """A scene stored in tests/golden/sources.npz (prefix "" or "p_")."""
    cam = scenes.sources_room().camera
    g = lambda k: np.asarray(G[prefix + k])  # noqa: E731
    return scenes.Scene(g("tris").astype(np.float32), g("tri_mat").astype(np.uint32), g("mats"),
                        g("lights"), cam, "sources_room", spheres=g("spheres"),
                        tri_lights=g("tri_lights").astype(np.uint32),
                        sphere_sources=g("sphere_sources").astype(np.uint32),
                        point_sources=g("point_sources"))


def means(fn, k, E):
    """Mean and variance of the mean over E estimates fn(e) -> (k, 3)."""
    acc, acc2 = np.zeros((k, 3)), np.zeros((k, 3))
    for e in range(E):
        v = fn(e)
        acc += v
        acc2 += v * v
    m = acc / E
    return m, np.maximum(acc2 / E - m * m, 0.0) / E


def oracle_means(D, estimator, x, n, d, prim, E, cv=0, flags=0):
    def one(e):
        P = host.direct_params(estimator, cv_samples=cv, flags=flags)
        P.estimate = e
        return D.direct_light(P, x, n, d, prim)

    return means(one, len(prim), E)


def agree(m1, v1, m2, v2, z=4.5, rel_floor=1e-9):
    """|m1 - m2| within z standard deviations (plus a rounding floor for the
    deterministic components)."""
    tol = z * np.sqrt(v1 + v2) + rel_floor * np.maximum(np.abs(m1), np.abs(m2)) + 1e-15
    bad = np.abs(m1 - m2) > tol
    return not bad.any(), np.argwhere(bad)[:8]


ESTIMATORS = {
    "analytic": dict(estimator=abi.DIRECT_ANALYTIC),
    "analytic_cv16": dict(estimator=abi.DIRECT_ANALYTIC_CV, cv_samples=16),
    "mis": dict(estimator=abi.DIRECT_MIS),
    "miscv": dict(estimator=abi.DIRECT_MISCV),
    "miscv_vis": dict(estimator=abi.DIRECT_MISCV, flags=abi.DIRECT_VISIBILITY),
}
