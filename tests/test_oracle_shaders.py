"""The FP64 restatement of the reference's path shaders (oracle/cpu_ref.cpp:
shade_implicit, shade_explicit; shader.cpp:111-207,276-337, which do not
compile in the reference: shade() calls an estimator its class never
declares).  Pinned by identities that fix their means and their structure:

  * shadeImplicit (cosine sampling, up to 4 vertices, an emitter's
    radiance.maxnorm()) has the mean of the configs' path tracer at max
    depth 3 when every emitter's radiance has maximum component 1 (then
    maxnorm is the identity): both estimate the same 0..3-bounce transport.
  * shadeExplicit stopped by Russian roulette at its first vertex returns
    exactly shadeDirectOnly's sample (same estimator and counter); with the
    MIS estimator its mean is the full light transport of a deep path tracer
    (its continuation's 2 pi weight ignores the excluded light directions,
    a bias of 2 pi / (2 pi - solid angle of the lights), kept below 1e-3 by
    small lights)."""
import numpy as np
import pytest

import oracle
from chakumrender_amd import abi, host, scenes

SEED = 0x5EED000A


def _image_z(m1, v1, m2, v2):
    return abs(m1.mean() - m2.mean()) / np.sqrt(v1.sum() / v1.size ** 2 + v2.sum() / v2.size ** 2)


def _stats(acc, sq, spp):
    m = acc / spp
    return m, np.maximum(sq / spp - m * m, 0) / spp


def test_implicit_mean_equals_depth3_path_tracer():
    sc = scenes.direct_room(width=12, height=12, light_intensity=1.0, light_scale=1.6, fan=False)
    D = oracle.DirectScene(sc)
    spp = 4096
    a, s2 = D.render_shade(abi.SHADE_IMPLICIT, dict(spp=spp, seed_lo=SEED, flags=abi.RENDER_SUMSQ),
                           host.direct_params(abi.DIRECT_MIS), sumsq=True)
    b, t2, _, _ = oracle.render(sc, dict(spp=1024, max_depth=3, seed_lo=SEED + 1,
                                         flags=abi.RENDER_SUMSQ), sumsq=True)
    m1, v1 = _stats(a, s2, spp)
    m2, v2 = _stats(b, t2, 1024)
    z = np.abs(m1 - m2) / np.sqrt(v1 + v2 + 1e-30)
    print(f"per-pixel max z {z.max():.2f}, image z {_image_z(m1, v1, m2, v2):.2f}, "
          f"means {m1.mean():.5f} {m2.mean():.5f}")
    assert (z < 4.5).mean() >= 0.995
    assert _image_z(m1, v1, m2, v2) < 4.0
    assert m1.mean() > 1e-3


def test_explicit_roulette_stop_is_direct_only():
    sc = scenes.direct_room(width=16, height=16)
    D = oracle.DirectScene(sc)
    for est in (abi.DIRECT_MISCV, abi.DIRECT_MIS):
        dp = host.direct_params(est)
        P = dict(spp=1, seed_lo=SEED)
        ex, _ = D.render_shade(abi.SHADE_EXPLICIT, P, dp)
        do, _ = D.render_shade(abi.SHADE_DIRECT_ONLY, P, dp)
        W, H = 16, 16
        stop = np.zeros((H, W), bool)
        for idx in range(W * H):
            r = oracle.philox([idx, 0, 0xFF, abi.SHADE_STREAM], [SEED, 0])
            stop[idx // W, idx % W] = (r[0] + 0.5) / 2 ** 32 < 0.3
        assert 0.15 < stop.mean() < 0.45
        np.testing.assert_array_equal(ex[stop], do[stop])
        if est == abi.DIRECT_MIS:  # non-negative terms: the continuation only adds
            assert (ex[~stop] >= do[~stop]).all()
            assert (ex[~stop] > do[~stop]).mean() > 0.5


def test_explicit_mis_mean_equals_deep_path_tracer():
    sc = scenes.direct_room(width=10, height=10, light_scale=0.15, fan=False)
    D = oracle.DirectScene(sc)
    spp = 192
    a, s2 = D.render_shade(abi.SHADE_EXPLICIT, dict(spp=spp, seed_lo=SEED, flags=abi.RENDER_SUMSQ),
                           host.direct_params(abi.DIRECT_MIS), sumsq=True)
    b, t2, _, _ = oracle.render(sc, dict(spp=4096, max_depth=16, seed_lo=SEED + 1,
                                         flags=abi.RENDER_SUMSQ), sumsq=True)
    m1, v1 = _stats(a, s2, spp)
    m2, v2 = _stats(b, t2, 4096)
    iz = _image_z(m1, v1, m2, v2)
    rel = abs(m1.mean() - m2.mean()) / m2.mean()
    print(f"image z {iz:.2f}, rel {rel:.2e}, means {m1.mean():.6f} {m2.mean():.6f}")
    assert iz < 4.0 or rel < 1e-3
