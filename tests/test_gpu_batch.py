"""chk_render_batch: many requests (different scenes, sizes, spp, depths) in
one wavefront.  Every request's image must equal its own chk_render image
bit for bit (same Philox keys, same per-path arithmetic, same per-pixel
summation order), in eager mode, with captured hipGraphs replayed several
times, and with the autotuner's choice."""
import numpy as np
import pytest

from chakumrender_amd import abi, host, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene_set(ctx):
    a = abi.DeviceScene(ctx, scenes.cornell(width=40, height=30))
    b = abi.DeviceScene(ctx, scenes.icosphere_scene(freq=12, n_lights=4, width=24, height=16))
    c = abi.DeviceScene(ctx, scenes.cornell(seed=0x5EED0007, sphere_tris=2000, n_lights=8,
                                            width=32, height=32, light_rng=True))
    return a, b, c


def batch_items(scene_set):
    a, b, c = scene_set
    return [
        dict(ds=a, params=dict(spp=16, max_depth=4, seed=11, flags=abi.RENDER_SUMSQ)),
        dict(ds=b, params=dict(spp=4, max_depth=6, seed=12)),
        dict(ds=c, params=dict(spp=64, max_depth=3, seed=13)),
        dict(ds=a, params=dict(spp=1, max_depth=2, seed=14, jitter_mode=abi.JITTER_REFERENCE)),
        dict(ds=b, params=dict(spp=32, sample_offset=32, max_depth=5, seed=15)),
        dict(ds=c, params=dict(spp=8, max_depth=4, seed=16, flags=abi.RENDER_SUMSQ)),
    ]


def single(items):
    out = []
    for it in items:
        acc, sq, _ = host.render(it["ds"], it["params"])
        out.append((acc, sq))
    return out


@pytest.mark.parametrize("tuning", [None, dict(chunk_paths=1 << 16, trace_block=64, shade_block=512),
                                    dict(chunk_paths=1 << 17, trace_block=256, shade_block=128)])
def test_batch_equals_single_renders(ctx, scene_set, tuning):
    items = batch_items(scene_set)
    ref = single(items)
    got, lat, st = host.render_batch(ctx, items, tuning)
    print(f"tuning {tuning}: chunks {st['chunks']}, segments {st['segments']}, "
          f"{st['batch_ms']:.2f} ms, latencies {np.round(lat, 3)}")
    assert st["paths"] == sum(it["ds"].scene.camera.width * it["ds"].scene.camera.height *
                              it["params"]["spp"] for it in items)
    for k, ((a, s), (ra, rs)) in enumerate(zip(got, ref)):
        np.testing.assert_array_equal(a, ra, err_msg=f"request {k}")
        if rs is not None:
            np.testing.assert_array_equal(s, rs, err_msg=f"request {k} sumsq")
    # every request's latency lies within the batch (consecutive chunks run
    # on two streams and overlap, so a later chunk can finish first)
    assert min(lat) > 0 and max(lat) <= st["batch_ms"] + 1e-3


def test_graph_replays_match_eager(ctx, scene_set):
    items = batch_items(scene_set)
    tun = dict(chunk_paths=1 << 16, trace_block=128, shade_block=256)
    eager = host.Batch(ctx, items)
    graph = host.Batch(ctx, items)
    for rep in range(3):
        le, se = eager.run(tun)
        lg, sg = graph.run(dict(tun, use_graph=1))
        assert sg["graph_cached"] == (1 if rep > 0 else 0)
        assert sg["chunks"] > 2  # the one graph holds both streams' chunks
        # latencies come from device clock stamps inside the graph
        for lat, s in ((le, se), (lg, sg)):
            assert min(lat) > 0 and max(lat) <= s["batch_ms"] + 1e-3, (lat, s["batch_ms"])
        # the program's own eager work between replays
        host.render(scene_set[1], dict(spp=2, max_depth=3, seed=99))
        for k, ((a, s), (b, t)) in enumerate(zip(eager.get(), graph.get())):
            np.testing.assert_array_equal(a, b, err_msg=f"replay {rep}, request {k}")
            if s is not None:
                np.testing.assert_array_equal(s, t, err_msg=f"replay {rep}, request {k} sumsq")
    print(f"capture {sg['capture_ms']:.2f} ms (last call cached), eager {se['batch_ms']:.2f} ms, "
          f"graph {sg['batch_ms']:.2f} ms")


def test_autotune_choice_renders_identically(ctx, scene_set):
    items = batch_items(scene_set)
    best, ms = host.batch_autotune(ctx, items)
    print(f"autotuned {best}, {ms:.3f} ms")
    assert best["chunk_paths"] in (1 << 22, 1 << 23, 1 << 24, 1 << 25)
    assert best["trace_block"] in (64, 128) and best["shade_block"] == 256
    ref = single(items)
    got, _, _ = host.render_batch(ctx, items, best)
    for (a, _), (ra, _) in zip(got, ref):
        np.testing.assert_array_equal(a, ra)


def test_batch_errors(ctx, scene_set):
    items = batch_items(scene_set)
    with pytest.raises(abi.ChkError, match="tuning"):
        host.render_batch(ctx, items, dict(trace_block=96))
    bad = [dict(items[0], params=dict(spp=4, mis_mode=abi.MIS_LEARNED))]
    with pytest.raises(abi.ChkError, match="CHK_MIS_BALANCE"):
        host.render_batch(ctx, bad)
    _, _, st = host.render_batch(ctx, [])
    assert st["paths"] == 0
