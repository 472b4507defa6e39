"""chk_render_noalloc: chk_render without device allocation, its path
queues in the library's fixed per-device scratch array.  Its images must
equal chk_render's bit for bit (same keys, same per-path arithmetic, whole
pixels per chunk so the same per-pixel summation order), on consecutive
calls whose second is smaller than the first (the array keeps the first
call's contents), on a render of many chunks and on a deep tree that needs
the overflow retry; sizes beyond the array are refused."""
import ctypes as C

import numpy as np
import pytest

from chakumrender_amd import abi, host, scenes

pytestmark = pytest.mark.gpu

MAX_CHUNK = 356_352  # paths per chunk that CHK_NOALLOC_SCRATCH_BYTES holds


def render_noalloc(ds, params, camera=None):
    cam = abi.camera_struct(camera or ds.scene.camera)
    P = abi.params_struct(params)
    img = host.Image(cam.width, cam.height, sumsq=bool(P.flags & abi.RENDER_SUMSQ))
    st = abi.chk_render_stats()
    abi.check(ds.L.chk_render_noalloc(ds.h, C.byref(cam), C.byref(P), img.acc.ptr,
                                      img.sq.ptr if img.sq is not None else None, C.byref(st),
                                      None), "chk_render_noalloc")
    a, s = img.get()
    return a, s, {k: getattr(st, k) for k, _ in abi.chk_render_stats._fields_}


def test_noalloc_equals_render_on_two_calls_second_smaller(ctx):
    sc = scenes.cornell(width=192, height=128, sphere_tris=4000, n_lights=4, light_rng=True)
    ds = abi.DeviceScene(ctx, sc)
    big = dict(spp=32, max_depth=4, seed=0x5EED0031, flags=abi.RENDER_SUMSQ)
    small_cam = scenes.Camera((0.5, 0.5, 2.5), (0, 0, -1), (0, 1, 0), 36.0, 36.0, 40, 24)
    small = dict(spp=8, max_depth=3, seed=0x5EED0032)
    print("noalloc: 192x128x32 (3 chunks) then 40x24x8", flush=True)
    a1, s1, st1 = render_noalloc(ds, big)
    r1, q1, rt1 = host.render(ds, big)
    assert st1["chunks"] == -(-192 * 128 // (MAX_CHUNK // 32)) == 3
    assert st1["paths"] == rt1["paths"] and st1["rays_extend"] == rt1["rays_extend"]
    np.testing.assert_array_equal(a1, r1)
    np.testing.assert_array_equal(s1, q1)
    a2, _, st2 = render_noalloc(ds, small, camera=small_cam)
    r2, _, rt2 = host.render(ds, small, camera=small_cam)
    assert st2["chunks"] == 1 and st2["rays_shadow"] == rt2["rays_shadow"]
    np.testing.assert_array_equal(a2, r2)
    assert a2.mean() > 0


def test_noalloc_deep_tree_retries_equal_render(ctx):
    """A tree whose worst-case stack exceeds 32 entries: the overflow list
    lives in the scratch array too."""
    sc = scenes.cornell(sphere_tris=100_000, width=48, height=32)
    ds = abi.DeviceScene(ctx, sc)
    assert ds.info()["stack_bound"] > 32
    P = dict(spp=4, max_depth=4, seed=0x5EED0033)
    print("noalloc: deep tree 48x32x4", flush=True)
    a, _, st = render_noalloc(ds, P)
    r, _, _ = host.render(ds, P)
    np.testing.assert_array_equal(a, r)


def test_noalloc_rejects_what_it_cannot_hold(ctx):
    ds = abi.DeviceScene(ctx, scenes.cornell(width=8, height=8))
    with pytest.raises(abi.ChkError, match="capacity"):
        render_noalloc(ds, dict(spp=MAX_CHUNK + 1, max_depth=1, seed=1))
    with pytest.raises(abi.ChkError, match="PROFILE"):
        render_noalloc(ds, dict(spp=1, max_depth=1, seed=1, flags=abi.RENDER_PROFILE))
    # the largest spp that fits still renders
    a, _, st = render_noalloc(ds, dict(spp=MAX_CHUNK // 64 * 64, max_depth=0, seed=1),
                              camera=scenes.Camera((0.5, 0.5, 2.5), (0, 0, -1), (0, 1, 0), 36.0,
                                                   36.0, 1, 1))
    assert st["paths"] == MAX_CHUNK // 64 * 64
