"""ORACLE -- test infrastructure only.

ctypes bindings to the FP64 CPU restatement (oracle/liboracle.so, built from
cpu_ref.cpp) and, where it was built, to the reference's own compiled code
(oracle/_ref/libchkref.so, built from the reference sources by
oracle/Makefile).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or the timed CPU
baseline.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libchkref.so"
REF_ROOT = Path(os.environ.get("CHK_REFERENCE", "/root/reference"))
LTC_TABLE = HERE.parent / "chakumrender_amd" / "data" / "ltc_ggx_64.bin"

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def build(ref: bool | None = None):
    """Compile liboracle.so (and the reference library when the reference
    checkout is present, or when ref=True)."""
    if ref is None:
        ref = REF_ROOT.is_dir()
    target = ["ref"] if ref else ["all"]
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8"] + target, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build(ref=False)
        L = C.CDLL(str(LIB_PATH))
        L.orc_tri_intersect.restype = C.c_double
        L.orc_tri_intersect.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_aabb_intersect.restype = C.c_double
        L.orc_aabb_intersect.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_closest_brute.argtypes = [_dp, C.c_uint32, _dp, C.c_uint32, _ip, _dp, _dp, _dp, _dp, _dp]
        L.orc_occluded_brute.argtypes = [_dp, C.c_uint32, _dp, C.c_uint32, _bp]
        L.orc_sphere_intersect.restype = C.c_double
        L.orc_sphere_intersect.argtypes = [_dp, _dp, _dp]
        L.orc_spheres_brute.argtypes = [_dp, C.c_uint32, _dp, C.c_uint32, _ip, _dp]
        L.orc_bvh_build.restype = _vp
        L.orc_bvh_build.argtypes = [_dp, C.c_uint32]
        L.orc_bvh_free.argtypes = [_vp]
        L.orc_bvh_node_count.restype = C.c_uint32
        L.orc_bvh_node_count.argtypes = [_vp]
        L.orc_bvh_ordered.argtypes = [_vp, _up]
        L.orc_bvh_closest.argtypes = [_vp, _dp, C.c_uint32, C.c_int, _ip, _ip, _dp, _dp, _up, _up]
        L.orc_brdf_eval.argtypes = [_vp, _up, C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_brdf_sample.argtypes = [_vp, _up, C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_onb.argtypes = [_dp, _dp, _dp]
        L.orc_ltc_set.argtypes = [_dp]
        L.orc_M_GGX.argtypes = [C.c_float, C.c_float, _dp, _dp, _dp, _dp]
        L.orc_ltc_shade.argtypes = [C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_uint32,
                                    _dp, _dp, _dp, _dp]
        L.orc_ltc_brdf_cos.argtypes = [C.c_uint32, _dp, _dp, _dp, _dp, _dp]
        L.orc_philox4x32_10.argtypes = [_up, _up, _up]
        L.orc_render.restype = C.c_uint64
        L.orc_render.argtypes = [_fp, C.c_uint32, _up, _vp, C.c_uint32, _vp, C.c_uint32, _vp, _vp,
                                 C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _dp, _vp,
                                 C.POINTER(C.c_uint64)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_direct_scene.restype = _vp
        L.orc_direct_scene.argtypes = [_fp, C.c_uint32, _up, _vp, C.c_uint32, _vp, C.c_uint32]
        L.orc_direct_scene_free.argtypes = [_vp]
        L.orc_direct_scene_desc.restype = _vp
        L.orc_direct_scene_desc.argtypes = [_vp]
        L.orc_mis_count.restype = C.c_uint32
        L.orc_mis_count.argtypes = [_vp]
        L.orc_env_tables.restype = C.c_double
        L.orc_env_tables.argtypes = [_vp, _up, _up]
        L.orc_env_radiance.argtypes = [_vp, C.c_uint32, _dp, _dp]
        L.orc_env_sample.argtypes = [_vp, C.c_uint32, _dp, _dp]
        L.orc_mis_weights.argtypes = [_vp, C.c_uint32, _dp, _dp, _dp, _dp, _up, _dp, C.c_uint32,
                                      _dp]
        L.orc_mis_samples.argtypes = [_vp, _dp, _dp, _dp, _dp, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32, _dp]
        L.orc_direct_light.argtypes = [_vp, _vp, C.c_uint32, _dp, _dp, _dp, _up, C.c_int, _dp]
        L.orc_render_shade.argtypes = [_vp, _vp, _vp, C.c_uint32, _vp, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_int, _dp, _vp]
        table = np.fromfile(LTC_TABLE, "<f8")
        L.orc_ltc_set(np.ascontiguousarray(table))
        _lib = L
    return _lib


def ref_lib():
    """The compiled reference (None when oracle/_ref/libchkref.so is absent)."""
    global _ref
    if _ref is None and REF_PATH.exists():
        R = C.CDLL(str(REF_PATH))
        R.ref_bvh_closest.argtypes = [_dp, C.c_uint32, _dp, C.c_uint32, _ip, _dp, _dp]
        R.ref_bvh_occluded.argtypes = [_dp, C.c_uint32, _dp, C.c_uint32, _bp]
        R.ref_bvh_spheres.argtypes = [_dp, C.c_uint32, _dp, C.c_uint32, _ip, _dp]
        R.ref_brdf.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp]
        R.ref_M_GGX.argtypes = [C.c_float, C.c_float, _dp, _dp, _dp, _dp]
        R.ref_ltc_shade.argtypes = [C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_uint32, _dp]
        R.ref_direct_setup.argtypes = [_dp, C.c_uint32, _up, _dp, C.c_uint32, _up, C.c_uint32]
        R.ref_direct_setup_ex.argtypes = [_dp, C.c_uint32, _up, _dp, C.c_uint32, _up, C.c_uint32,
                                          _vp, C.c_uint32, _vp, _vp, C.c_uint32, _vp, C.c_uint32,
                                          _vp, C.c_uint32]
        R.ref_env_setup.argtypes = [_dp, C.c_uint32, C.c_uint32, C.c_uint32, _dp, C.c_double]
        R.ref_env_area.restype = C.c_double
        R.ref_env_radiance.argtypes = [C.c_uint32, _dp, _dp]
        R.ref_env_samples.argtypes = [C.c_uint32, _dp]
        R.ref_read_ppm.argtypes = [C.c_char_p, _dp, C.c_uint32, _up]
        R.ref_write_ppm.argtypes = [C.c_char_p, _dp, C.c_uint32, C.c_uint32]
        R.ref_mis_get_samples.restype = C.c_uint32
        R.ref_mis_get_samples.argtypes = [C.c_uint32, _dp, _dp, _dp, _dp, _up, _dp, _dp]
        R.ref_direct_stats.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       _dp, _dp, _dp, _up, _dp, _dp]
        _ref = R
    return _ref


def _d(a, shape=None):
    a = np.ascontiguousarray(a, np.float64)
    return a if shape is None else a.reshape(shape)


# ---------------------------------------------------------------- geometry
def closest_brute(tris, rays):
    tris, rays = _d(tris, (-1, 9)), _d(rays, (-1, 6))
    m = rays.shape[0]
    out = dict(id=np.empty(m, np.int32), t=np.empty(m), N=np.empty((m, 3)), t2=np.empty(m),
               margin=np.empty(m), cos2=np.empty(m))
    lib().orc_closest_brute(tris, tris.shape[0], rays, m, out["id"], out["t"], out["N"],
                            out["t2"], out["margin"], out["cos2"])
    return out


def spheres_brute(spheres, rays):
    """Sphere::intersect closest over a sphere list: (id, t)."""
    spheres, rays = _d(spheres, (-1, 4)), _d(rays, (-1, 6))
    m = rays.shape[0]
    ids, t = np.empty(m, np.int32), np.empty(m)
    lib().orc_spheres_brute(spheres, spheres.shape[0], rays, m, ids, t)
    return ids, t


def sphere_intersect(sphere, o, d):
    return lib().orc_sphere_intersect(_d(sphere), _d(o), _d(d))


def occluded_brute(tris, rays7):
    tris, rays7 = _d(tris, (-1, 9)), _d(rays7, (-1, 7))
    occ = np.empty(rays7.shape[0], np.uint8)
    lib().orc_occluded_brute(tris, tris.shape[0], rays7, rays7.shape[0], occ)
    return occ


class RefBvh:
    """The reference's BVH (midpoint split, 1 primitive per leaf) restated."""

    def __init__(self, tris):
        self.tris = _d(tris, (-1, 9))
        self.h = lib().orc_bvh_build(self.tris, self.tris.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_free(self.h)
            self.h = None

    @property
    def node_count(self):
        return lib().orc_bvh_node_count(self.h)

    def ordered(self):
        o = np.empty(self.tris.shape[0], np.uint32)
        lib().orc_bvh_ordered(self.h, o)
        return o

    def closest(self, rays, prune=False):
        rays = _d(rays, (-1, 6))
        m = rays.shape[0]
        out = dict(id=np.empty(m, np.int32), rid=np.empty(m, np.int32), t=np.empty(m),
                   N=np.empty((m, 3)), nv=np.empty(m, np.uint32), nt=np.empty(m, np.uint32))
        lib().orc_bvh_closest(self.h, rays, m, int(prune), out["id"], out["rid"], out["t"],
                              out["N"], out["nv"], out["nt"])
        return out


# ---------------------------------------------------------------- materials
def brdf_eval(mats, mat_idx, n, wo_ref, wo, wi):
    mats = np.ascontiguousarray(mats)
    mat_idx = np.ascontiguousarray(mat_idx, np.uint32)
    k = mat_idx.shape[0]
    f = np.empty(k)
    pdf = np.empty(k)
    lib().orc_brdf_eval(mats.ctypes.data, mat_idx, k, _d(n, (-1, 3)), _d(wo_ref, (-1, 3)),
                        _d(wo, (-1, 3)), _d(wi, (-1, 3)), f, pdf)
    return f, pdf


def brdf_sample(mats, mat_idx, n, wo_ref, wo, u):
    mats = np.ascontiguousarray(mats)
    mat_idx = np.ascontiguousarray(mat_idx, np.uint32)
    k = mat_idx.shape[0]
    wi = np.empty((k, 3))
    pdf = np.empty(k)
    lib().orc_brdf_sample(mats.ctypes.data, mat_idx, k, _d(n, (-1, 3)), _d(wo_ref, (-1, 3)),
                          _d(wo, (-1, 3)), _d(u, (-1, 3)), wi, pdf)
    return wi, pdf


def onb(n):
    t = np.empty(3)
    b = np.empty(3)
    lib().orc_onb(_d(n), t, b)
    return t, b


# ---------------------------------------------------------------- LTC
def M_GGX(theta, alpha):
    M = np.empty(9)
    Minv = np.empty(9)
    amp = np.empty(1)
    det = np.empty(1)
    lib().orc_M_GGX(float(theta), float(alpha), M, Minv, amp, det)
    return M, Minv, float(amp[0]), float(det[0])


def ltc_shade(x, n, v, alpha, ks, tris, radiance):
    x, n, v = _d(x, (-1, 3)), _d(n, (-1, 3)), _d(v, (-1, 3))
    k = x.shape[0]
    alpha = _d(alpha).reshape(-1)
    ks = _d(np.broadcast_to(ks, (k,)).copy())
    tris, radiance = _d(tris, (-1, 9)), _d(radiance, (-1, 3))
    rgb, ltc, dif, amp = np.empty((k, 3)), np.empty((k, 3)), np.empty((k, 3)), np.empty(k)
    lib().orc_ltc_shade(k, x, n, v, alpha, ks, tris, radiance, tris.shape[0], rgb, ltc, dif, amp)
    return dict(rgb=rgb, ltc=ltc, diffuse=dif, amp=amp)


def ltc_brdf_cos(n, wo, wi, alpha):
    n, wo, wi = _d(n, (-1, 3)), _d(wo, (-1, 3)), _d(wi, (-1, 3))
    k = n.shape[0]
    out = np.empty(k)
    lib().orc_ltc_brdf_cos(k, n, wo, wi, _d(alpha).reshape(-1), out)
    return out


# ---------------------------------------------------------------- RNG
def philox(ctr, key):
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


# ---------------------------------------------------------------- render
def render(scene, params, rows=None, threads=0, sumsq=False, row_step=1):
    """Render rows r0, r0 + row_step, ... < r1 of `scene` with the FP64 estimator.  params: a
    chakumrender_amd.RenderParams or dict with spp, max_depth, seed ... .
    Returns (accum (H, W, 3) float64 sums over samples, sumsq or None,
    paths, rays)."""
    from chakumrender_amd import abi

    cam = abi.camera_struct(scene.camera)
    P = abi.params_struct(params)
    W, H = scene.camera.width, scene.camera.height
    r0, r1 = (0, H) if rows is None else rows
    acc = np.zeros((H, W, 3))
    sq = np.zeros((H, W, 3)) if sumsq else None
    nr = C.c_uint64(0)
    mats = np.ascontiguousarray(scene.mats)
    lights = np.ascontiguousarray(scene.lights)
    paths = lib().orc_render(np.ascontiguousarray(scene.tris, np.float32), scene.n_tris,
                             np.ascontiguousarray(scene.tri_mat, np.uint32), mats.ctypes.data,
                             mats.shape[0], lights.ctypes.data, lights.shape[0], C.addressof(cam),
                             C.addressof(P), r0, r1, int(row_step), int(threads), acc,
                             sq.ctypes.data if sq is not None else None, C.byref(nr))
    return acc, sq, int(paths), int(nr.value)


# ---------------------------------------------------------------- direct lighting
class DirectScene:
    """orc_direct_scene: the shadeDI estimators' FP64 restatement over a
    scene (triangles, materials, mesh lights)."""

    def __init__(self, scene):
        from chakumrender_amd import abi

        self.scene = scene
        self.mats = np.ascontiguousarray(scene.mats)
        self.lights = np.ascontiguousarray(scene.lights)
        d, keep = abi.scene_desc(scene)
        self.h = lib().orc_direct_scene_desc(C.addressof(d))
        del keep

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_direct_scene_free(self.h)
            self.h = None

    @property
    def n_samples(self):
        """MIS::getSamples' sample count: 10 per sphere source, mesh light and
        triangle light, 60 BSDF."""
        return int(lib().orc_mis_count(self.h))

    def env_tables(self):
        """(normalizationArea, inverse marginal CDF (2000), inverse
        conditional CDFs (3000, 2000)) of the scene's environment light."""
        im, ic = np.zeros(2000, np.uint32), np.zeros((3000, 2000), np.uint32)
        area = lib().orc_env_tables(self.h, im, ic)
        return area, im, ic

    def env_radiance(self, dirs):
        d = _d(dirs, (-1, 3))
        out = np.empty_like(d)
        lib().orc_env_radiance(self.h, d.shape[0], d, out)
        return out

    def env_sample(self, u):
        u = _d(u, (-1, 2))
        out = np.empty((u.shape[0], 3))
        lib().orc_env_sample(self.h, u.shape[0], u, out)
        return out

    def mis_weights(self, x, n, wo_ref, wo, prim, samples):
        smp = _d(samples)
        k, S = smp.shape[0], smp.shape[1]
        w = np.empty((k, S))
        lib().orc_mis_weights(self.h, k, _d(x, (-1, 3)), _d(n, (-1, 3)), _d(wo_ref, (-1, 3)),
                              _d(wo, (-1, 3)), np.ascontiguousarray(prim, np.uint32),
                              smp.reshape(-1, 3), S, w)
        return w

    def mis_samples(self, x, n, wo_ref, wo, prim, point, estimate, seed):
        out = np.empty((self.n_samples, 3))
        lib().orc_mis_samples(self.h, _d(x), _d(n), _d(wo_ref), _d(wo), int(prim), int(point),
                              int(estimate), int(seed[0]), int(seed[1]), out)
        return out

    def direct_light(self, params, x, n, d, prim, threads=0):
        """params: chakumrender_amd.abi.chk_direct_params."""
        k = int(np.asarray(prim).size)
        out = np.empty((k, 3))
        lib().orc_direct_light(self.h, C.addressof(params), k, _d(x, (-1, 3)), _d(n, (-1, 3)),
                               _d(d, (-1, 3)), np.ascontiguousarray(prim, np.uint32),
                               int(threads), out)
        return out

    def render_shade(self, shader, params, direct, rows=None, threads=0, sumsq=False,
                     row_step=1):
        """orc_render_shade: shader 0 shadeDirectOnly, 1 shadeExplicit, 2 shadeImplicit."""
        from chakumrender_amd import abi

        cam = abi.camera_struct(self.scene.camera)
        P = abi.params_struct(params)
        W, H = cam.width, cam.height
        r0, r1 = (0, H) if rows is None else rows
        acc = np.zeros((H, W, 3))
        sq = np.zeros((H, W, 3)) if sumsq else None
        lib().orc_render_shade(self.h, C.addressof(cam), C.addressof(P), int(shader),
                               C.addressof(direct), r0, r1, int(row_step), int(threads), acc,
                               sq.ctypes.data if sq is not None else None)
        return acc, sq

    def render_direct(self, params, direct, **kw):
        return self.render_shade(0, params, direct, **kw)


class _quiet:
    """Silence the reference's own prints (it reports unset LTC parameters
    and negative control-variate sums on stdout/stderr on every call)."""

    def __enter__(self):
        import sys

        sys.stdout.flush()
        sys.stderr.flush()
        self.null = os.open(os.devnull, os.O_WRONLY)
        self.saved = (os.dup(1), os.dup(2))
        os.dup2(self.null, 1)
        os.dup2(self.null, 2)

    def __exit__(self, *a):
        import ctypes.util

        libc = C.CDLL(ctypes.util.find_library("c"))
        libc.fflush(None)
        os.dup2(self.saved[0], 1)
        os.dup2(self.saved[1], 2)
        for f in (self.null,) + self.saved:
            os.close(f)


class RefDirect:
    """The reference's own LightSource / MIS objects over a scene, through
    the shim in oracle/_ref (None-safe: check ref_lib() first)."""

    def __init__(self, scene):
        R = ref_lib()
        mats = np.ascontiguousarray(scene.mats)
        # per-material doubles: brdf, ks, alpha, pe, iior, eior, refl, c[3], rad[3], light
        md = np.zeros((mats.shape[0], 14))
        md[:, 0] = mats["brdf"]
        md[:, 1] = mats["ks"]
        md[:, 2] = mats["alpha"]
        md[:, 3] = mats["phong_exp"]
        md[:, 4] = mats["int_ior"]
        md[:, 5] = mats["ext_ior"]
        md[:, 6] = mats["reflectance"]
        md[:, 7:10] = mats["color"]
        md[:, 10:13] = mats["radiance"]
        md[:, 13] = mats["light"]
        lights = np.ascontiguousarray(scene.lights)
        lt = np.ascontiguousarray(np.stack([lights["first_tri"], lights["n_tris"]], 1), np.uint32)
        tris = _d(np.asarray(scene.tris, np.float32).astype(np.float64), (-1, 9))
        tri_mat = np.ascontiguousarray(scene.tri_mat, np.uint32)
        sph = np.asarray(getattr(scene, "spheres", np.zeros(0)))
        tl = np.ascontiguousarray(getattr(scene, "tri_lights", np.zeros(0)), np.uint32)
        ss = np.ascontiguousarray(getattr(scene, "sphere_sources", np.zeros(0)), np.uint32)
        pts = np.asarray(getattr(scene, "point_sources", np.zeros(0)))
        env = getattr(scene, "env", None)
        with _quiet():
            if sph.shape[0] or tl.shape[0] or ss.shape[0] or pts.shape[0] or env is not None:
                # the scene's spheres as given (a scene file's list holds the
                # dummy sphere of prepareScene.cpp:238 already)
                sd = np.ascontiguousarray(np.concatenate(
                    [sph["center"], sph["radius"][:, None]], 1) if sph.shape[0] else np.zeros((0, 4)))
                sm = np.ascontiguousarray(sph["mat"] if sph.shape[0] else np.zeros(0), np.uint32)
                pd = np.ascontiguousarray(np.concatenate(
                    [pts["pos"], pts["power"]], 1) if pts.shape[0] else np.zeros((0, 6)))
                ptr = lambda a: a.ctypes.data if a.shape[0] else None  # noqa: E731
                R.ref_direct_setup_ex(tris, scene.n_tris, tri_mat, np.ascontiguousarray(md),
                                      md.shape[0], lt, lt.shape[0], ptr(tl), tl.shape[0], ptr(sd),
                                      ptr(sm), sd.shape[0], ptr(ss), ss.shape[0], ptr(pd),
                                      pd.shape[0])
            else:
                R.ref_direct_setup(tris, scene.n_tris, tri_mat, np.ascontiguousarray(md),
                                   md.shape[0], lt, lt.shape[0])
            if env is not None:
                img = _d(env.image)
                R.ref_env_setup(img.reshape(-1, 3), img.shape[1], img.shape[0], int(env.axis),
                                _d(env.col), float(env.intensity))
        self.R = R
        self.n_lights = lt.shape[0]

    def env_area(self):
        return self.R.ref_env_area()

    def env_radiance(self, dirs):
        d = _d(dirs, (-1, 3))
        out = np.empty_like(d)
        self.R.ref_env_radiance(d.shape[0], d, out)
        return out

    def env_samples(self, n):
        out = np.zeros((n, 3))
        self.R.ref_env_samples(n, out)
        return out

    def mis_get_samples(self, x, n, wo_ref, wo, prim):
        """MIS::getSamples per point: (samples (k, S, 3), weights (k, S))."""
        k = int(np.asarray(prim).size)
        smp = np.zeros((k, 200, 3))
        w = np.zeros((k, 200))
        with _quiet():
            S = self.R.ref_mis_get_samples(k, _d(x, (-1, 3)), _d(n, (-1, 3)), _d(wo_ref, (-1, 3)),
                                           _d(wo, (-1, 3)), np.ascontiguousarray(prim, np.uint32),
                                           smp, w)
        return smp[:, :S], w[:, :S]

    def stats(self, estimator, x, n, d, prim, reps, cv_samples=0, batches=32):
        """reps calls of the reference's estimator per point: mean (k, 3) and
        the variance of that mean from `batches` batch means."""
        k = int(np.asarray(prim).size)
        mean, var = np.zeros((k, 3)), np.zeros((k, 3))
        with _quiet():
            self.R.ref_direct_stats(int(estimator), int(cv_samples), int(reps), int(batches), k,
                                    _d(x, (-1, 3)), _d(n, (-1, 3)), _d(d, (-1, 3)),
                                    np.ascontiguousarray(prim, np.uint32), mean, var)
        return mean, var
