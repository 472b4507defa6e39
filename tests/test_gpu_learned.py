"""chk_render_learned: the learned light-selection proposal inside the
wavefront integrator.

* An untrained proposal (q' = 1/L, c = 1/2) is the plain estimator: with the
  same random numbers the image equals chk_render's except where the float
  CDF of q' moves a light-choice boundary by an ulp.
* Online training keeps the estimator unbiased: per-pixel means of a
  trained render agree with an independent plain render within 3 sigma.
* Training lowers the per-pixel variance on a scene whose 16 lights differ
  about 100x in power (uniform light choice wastes most light samples).
"""
import numpy as np
import pytest

from chakumrender_amd import abi, host, scenes
from test_gpu_render import zscores

pytestmark = pytest.mark.gpu


def multi_light_scene(w=64, h=64):
    return scenes.cornell(seed=0x5EED0007, sphere_tris=960, n_lights=16, width=w, height=h,
                          light_rng=True)


def test_untrained_proposal_is_plain_mis(ctx):
    sc = multi_light_scene()
    ds = abi.DeviceScene(ctx, sc)
    nl = ds.info()["n_lights"]
    prop = host.Proposal(ctx, nl, seed=1)
    spp = 64
    P = dict(spp=spp, max_depth=4, seed=0x5EED, flags=abi.RENDER_SUMSQ)
    a_p, s_p, _ = host.render(ds, P)
    a_l, s_l, st = host.render_learned(ds, P, prop)
    assert st["train_steps"] == 0 and st["paths"] == 64 * 64 * spp
    z = zscores(a_l, s_l, spp, a_p, s_p, spp)
    frac = np.mean(np.abs(z) <= 3)
    close = np.abs(a_l - a_p) / spp <= 1e-4 * (1 + np.abs(a_p) / spp)
    print(f"within 3 sigma {frac:.4f}; pixels equal to 1e-4: {close.mean():.4f}")
    assert frac >= 0.99
    assert close.mean() > 0.95
    prop.close()


def test_training_is_unbiased_and_lowers_variance(ctx):
    sc = multi_light_scene()
    ds = abi.DeviceScene(ctx, sc)
    nl = ds.info()["n_lights"]
    prop = host.Proposal(ctx, nl, seed=2)
    p0 = prop.get_params()
    train = dict(train_spp=32, batch=4096, steps_per_bounce=32, lr=1e-3)
    # training pass: samples 0..31 of every pixel
    _, _, st = host.render_learned(ds, dict(spp=32, max_depth=4, seed=0x5EED), prop, train=train)
    p1 = prop.get_params()
    print(f"train steps {st['train_steps']}, |dW4| {np.abs(p1 - p0)[-(64 * 33 + 33):].max():.3g}")
    assert st["train_steps"] > 0
    assert np.isfinite(p1).all() and np.abs(p1 - p0).max() > 0
    # inference on later samples vs an independent plain render
    spp = 256
    P = dict(spp=spp, sample_offset=32, max_depth=4, seed=0x5EED, flags=abi.RENDER_SUMSQ)
    a_l, s_l, st_l = host.render_learned(ds, P, prop, train=train)
    assert st_l["train_steps"] == 0  # sample_offset >= train_spp
    a_p, s_p, _ = host.render(ds, dict(P, seed=0xBEEF))
    z = zscores(a_l, s_l, spp, a_p, s_p, spp)
    frac = np.mean(np.abs(z) <= 3)
    ml, mp = a_l / spp, a_p / spp
    vl = np.maximum(s_l / spp - ml * ml, 0).sum()
    vp = np.maximum(s_p / spp - mp * mp, 0).sum()
    print(f"within 3 sigma {frac:.4f}; image mean learned {ml.mean():.5f} plain {mp.mean():.5f}; "
          f"variance ratio learned/plain {vl / vp:.3f}")
    assert frac >= 0.99
    assert abs(ml.mean() - mp.mean()) / mp.mean() < 0.01
    assert vl / vp < 0.7
    prop.close()


def test_learned_render_argument_errors(ctx):
    sc = multi_light_scene(16, 16)
    ds = abi.DeviceScene(ctx, sc)
    nl = ds.info()["n_lights"]
    wrong = host.Proposal(ctx, max(1, nl - 1), seed=1)
    with pytest.raises(abi.ChkError, match="lights"):
        host.render_learned(ds, dict(spp=4), wrong)
    prop = host.Proposal(ctx, nl, seed=1)
    with pytest.raises(abi.ChkError, match="CHK_MIS_LEARNED"):
        host.render_learned(ds, dict(spp=4, mis_mode=abi.MIS_BALANCE), prop)
    with pytest.raises(abi.ChkError, match="batch"):
        host.render_learned(ds, dict(spp=4), prop, train=dict(batch=8))
    with pytest.raises(abi.ChkError, match="chk_render_learned"):
        host.render(ds, dict(spp=4, mis_mode=abi.MIS_LEARNED))
    wrong.close()
    prop.close()
