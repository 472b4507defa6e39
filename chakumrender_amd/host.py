"""Host-side mirror of the reference's per-sample-path interfaces over the
C ABI, batched.  Names and argument meanings follow the reference:

  BvhAccel.intersectT   bvhAccel.cpp:200-249   closest hit (t, N, id)
  BvhAccel.shadow       shader.cpp:76-97        1.0 if unoccluded else 0.0
  MaterialType.brdf_pdf / sample
                        materialTypes.cpp:9-153 (brdf, pdfEval,
                        getBrdfDirectionSampleRandomized)
  LightSource.getLightFromMeshSource_Analytic
                        lightSources_CV.cpp:323-333
  SceneLights           the shadeDI estimators over a scene's mesh lights
                        (lightSources.cpp:32-109, lightSources_CV.cpp:323-450)
                        and MIS::getSamples' weights (mis.cpp:18-223)
  render_direct         shadeDirectOnly (shader.cpp:339-363), the reference's
                        shipped shade()
  render_shade          any of its shaders: shadeDirectOnly, shadeExplicit
                        (shader.cpp:111-207), shadeImplicit (:276-337)
  render                main.cpp:62-94 around shade() (shader.cpp:362)
  write_ppm             main.cpp:27,98-117

Differences a caller sees: rays are batched numpy arrays; ids are ORIGINAL
triangle indices (the reference returns indices into its reordered array);
random numbers for sampling are explicit inputs.  Device memory is
allocated through libchakum itself (abi.DevBuf); every computation runs in
libchakum.so.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

from . import abi

DevBuf = abi.DevBuf


def make_rays(o, d, tmax=None, skip=None):
    """chk_ray records (n, 8) float32: (o, tmax), (d, skip bits)."""
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = o.shape[0]
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 1e30 if tmax is None else np.asarray(tmax, np.float32)
    r[:, 4:7] = d
    sk = np.full(n, -1, np.int32) if skip is None else np.asarray(skip, np.int32)
    r[:, 7] = sk.view(np.float32)
    return r


class BvhAccel:
    """BvhAccel over a scene on the GPU (bvhAccel.h:49-76)."""

    def __init__(self, ctx: abi.Context, scene):
        self.ds = abi.DeviceScene(ctx, scene)
        self.L = ctx.L

    def trace_closest(self, rays_dev, n, hits_dev, stream=None):
        abi.check(self.L.chk_trace_closest(self.ds.h, rays_dev, n, hits_dev, stream),
                  "chk_trace_closest")

    def trace_any(self, rays_dev, n, occ_dev, stream=None):
        abi.check(self.L.chk_trace_any(self.ds.h, rays_dev, n, occ_dev, stream), "chk_trace_any")

    def closest_raw(self, rays):
        """rays: (n, 8) chk_ray records -> (n, 4) hit records (t, tri bits, u, v)."""
        n = rays.shape[0]
        rd = DevBuf.from_array(np.asarray(rays, np.float32))
        hd = DevBuf(16 * n)
        self.trace_closest(rd.ptr, n, hd.ptr)
        return hd.download((n, 4), np.float32)

    def intersectT(self, o, d):
        """Closest hit for rays (o, d): (hit bool, t, id) arrays; t = 1e30 and
        id = -1 on a miss (the reference: INF, 0xFFFFFFFF)."""
        h = self.closest_raw(make_rays(o, d))
        tri = h[:, 1].view(np.int32).copy()
        return tri >= 0, h[:, 0].astype(np.float64), tri

    def occluded_raw(self, rays):
        n = rays.shape[0]
        rd = DevBuf.from_array(np.asarray(rays, np.float32))
        od = DevBuf(max(n, 16))
        self.trace_any(rd.ptr, n, od.ptr)
        return od.download((n,), np.uint8)

    def shadow(self, o, d, dist):
        """shader.cpp:76-97: 0.0 if some triangle is hit with t + 1e-5 < dist."""
        occ = self.occluded_raw(make_rays(o, d, tmax=np.asarray(dist, np.float32) - np.float32(1e-5)))
        return 1.0 - occ.astype(np.float64)


def intersect_spheres(spheres, rays):
    """BvhAccel::intersectS over a sphere list (n x 4: centre, radius) for
    chk_ray records: (n, 4) hit records (t, sphere id bits, 0, 0)."""
    L = abi.load()
    sp = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = r.shape[0]
    sd, rd, hd = DevBuf.from_array(sp), DevBuf.from_array(r), DevBuf(16 * max(n, 1))
    abi.check(L.chk_trace_spheres(sd.ptr, sp.shape[0], rd.ptr, n, hd.ptr, None), "chk_trace_spheres")
    return hd.download((n, 4), np.float32)


class MaterialType:
    """Batched MaterialType (materialTypes.h:20-120) over a material table.
    FP64 directions in and FP64 results out, like the reference's Vec
    interface; fp32=True evaluates with the render kernels' FP32 arithmetic."""

    FP32 = 1  # CHK_BRDF_FP32

    def __init__(self, mats):
        self.mats = np.ascontiguousarray(mats)
        self.L = abi.load()

    def _bufs(self, mat_idx, *vecs):
        md = DevBuf.from_array(self.mats.view(np.uint8))
        idx = DevBuf.from_array(np.asarray(mat_idx, np.uint32))
        vs = [DevBuf.from_array(np.asarray(v, np.float64).reshape(-1, 3)) for v in vecs]
        return md, idx, vs

    def brdf_pdf(self, mat_idx, n, wo_ref, wo, wi, fp32=False):
        """(BasePrimitive::brdf, MaterialType::pdfEval) per row."""
        k = len(mat_idx)
        md, idx, (nd, rd, od, wd) = self._bufs(mat_idx, n, wo_ref, wo, wi)
        f, p = DevBuf(8 * k), DevBuf(8 * k)
        abi.check(self.L.chk_brdf_eval(md.ptr, self.mats.shape[0], idx.ptr, nd.ptr, rd.ptr, od.ptr,
                                       wd.ptr, k, self.FP32 if fp32 else 0, f.ptr, p.ptr, None),
                  "chk_brdf_eval")
        return f.download((k,), np.float64), p.download((k,), np.float64)

    def sample(self, mat_idx, n, wo_ref, wo, u, fp32=False):
        """getBrdfDirectionSampleRandomized with explicit uniforms u (k, 3)."""
        k = len(mat_idx)
        md, idx, (nd, rd, od, ud) = self._bufs(mat_idx, n, wo_ref, wo, u)
        wi, p = DevBuf(24 * k), DevBuf(8 * k)
        abi.check(self.L.chk_brdf_sample(md.ptr, self.mats.shape[0], idx.ptr, nd.ptr, rd.ptr, od.ptr,
                                         ud.ptr, k, self.FP32 if fp32 else 0, wi.ptr, p.ptr, None),
                  "chk_brdf_sample")
        return wi.download((k, 3), np.float64), p.download((k,), np.float64)


class LightSource:
    """Mesh-light analytic shading (lightSources_CV.cpp:221-333) on the GPU."""

    def __init__(self, ctx: abi.Context, light_tris, radiance):
        self.ctx = ctx
        self.L = ctx.L
        self.n = int(np.asarray(light_tris).reshape(-1, 9).shape[0])
        self.tris = DevBuf.from_array(np.asarray(light_tris, np.float32).reshape(-1, 9))
        self.rad = DevBuf.from_array(np.asarray(radiance, np.float32).reshape(-1, 3))

    def eval_dev(self, bufs, k, rgb, stream=None):
        """bufs: DevBufs (x, n, v, alpha, ks); rgb: DevBuf (k, 3)."""
        abi.check(self.L.chk_ltc_eval(self.ctx.h, *[b.ptr for b in bufs], k, self.tris.ptr,
                                      self.rad.ptr, self.n, rgb.ptr, stream), "chk_ltc_eval")

    def getLightFromMeshSource_Analytic(self, x, n, v, alpha, ks):
        arrs = [np.asarray(a, np.float32) for a in (x, n, v, alpha, ks)]
        k = arrs[0].reshape(-1, 3).shape[0]
        bufs = [DevBuf.from_array(a) for a in arrs]
        rgb = DevBuf(12 * k)
        self.eval_dev(bufs, k, rgb)
        return rgb.download((k, 3), np.float32)


class SceneLights:
    """The reference's LightSource / MIS estimators over a scene's mesh
    lights (chk_direct_light, chk_mis_weights).  Shading points are the
    reference's (r.d, n, x, primitive): FP64 x, n (facing the viewer), d
    (the ray direction), and prim, the original index of the triangle x lies
    on.  Every call returns one estimate per point (FP64, k x 3)."""

    def __init__(self, ds: abi.DeviceScene, seed=(0x5EED0006, 0)):
        self.ds = ds
        self.L = ds.L
        self.seed = seed

    def _pts(self, *arrs):
        return [DevBuf.from_array(np.ascontiguousarray(a, np.float64).reshape(-1, 3)) for a in arrs]

    def estimate(self, estimator, x, n, d, prim, estimate=0, cv_samples=0, flags=0, seed=None):
        k = int(np.asarray(prim).size)
        xd, nd, dd = self._pts(x, n, d)
        pd = DevBuf.from_array(np.ascontiguousarray(prim, np.uint32))
        P = abi.chk_direct_params()
        P.estimator, P.cv_samples, P.flags, P.estimate = estimator, cv_samples, flags, estimate
        P.seed_lo, P.seed_hi = seed or self.seed
        out = DevBuf(24 * max(k, 1))
        abi.check(self.L.chk_direct_light(self.ds.h, xd.ptr, nd.ptr, dd.ptr, pd.ptr, k, C.byref(P),
                                          out.ptr, None), "chk_direct_light")
        return out.download((k, 3), np.float64)

    def getLightFromMeshSource_Analytic(self, x, n, d, prim):
        """lightSources_CV.cpp:323-333 (deterministic)."""
        return self.estimate(abi.DIRECT_ANALYTIC, x, n, d, prim)

    def getLightFromMeshSource_AnalyticCV(self, x, n, d, prim, nSamples, estimate=0):
        """lightSources_CV.cpp:335-397."""
        return self.estimate(abi.DIRECT_ANALYTIC_CV, x, n, d, prim, estimate, cv_samples=nSamples)

    def getLightFromAllSources_MIS(self, x, n, d, prim, estimate=0):
        """lightSources.cpp:32-109."""
        return self.estimate(abi.DIRECT_MIS, x, n, d, prim, estimate)

    def getLightFromMeshSource_Analytic_MISCV(self, x, n, d, prim, estimate=0, visibility=False):
        """lightSources_CV.cpp:399-450; visibility=True applies the shadow
        test the reference computes and ignores."""
        return self.estimate(abi.DIRECT_MISCV, x, n, d, prim, estimate,
                             flags=abi.DIRECT_VISIBILITY if visibility else 0)

    def mis_weights(self, x, n, wo_ref, wo, prim, samples):
        """The weights MIS::getSamples returns for these samples
        (mis.cpp:209-220); samples (k, S, 3) -> weights (k, S)."""
        smp = np.ascontiguousarray(samples, np.float64)
        k, S = smp.shape[0], smp.shape[1]
        xd, nd, rd, od = self._pts(x, n, wo_ref, wo)
        pd = DevBuf.from_array(np.ascontiguousarray(prim, np.uint32))
        sd = DevBuf.from_array(smp.reshape(-1, 3))
        w = DevBuf(8 * max(k * S, 1))
        abi.check(self.L.chk_mis_weights(self.ds.h, xd.ptr, nd.ptr, rd.ptr, od.ptr, pd.ptr, k, sd.ptr,
                                         S, w.ptr, None), "chk_mis_weights")
        return w.download((k, S), np.float64)


def direct_params(estimator, cv_samples=0, flags=0, seed=(0x5EED0006, 0)):
    P = abi.chk_direct_params()
    P.estimator, P.cv_samples, P.flags = estimator, cv_samples, flags
    P.seed_lo, P.seed_hi = seed
    return P


def render_shade(ds: abi.DeviceScene, params, shader, direct=None, camera=None,
                 image: "Image | None" = None, stats=True, stream=None, download=True):
    """chk_render_shade: one of the reference's shaders (abi.SHADE_DIRECT_ONLY,
    SHADE_EXPLICIT, SHADE_IMPLICIT; shader.cpp:111-378) with the direct-lighting
    estimator of `direct` (a chk_direct_params or a dict of its fields; unused
    by shadeImplicit)."""
    cam = abi.camera_struct(camera or ds.scene.camera)
    P = abi.params_struct(params)
    if isinstance(direct, dict):
        direct = direct_params(**direct)
    if image is None:
        image = Image(cam.width, cam.height, sumsq=bool(P.flags & abi.RENDER_SUMSQ))
    st = abi.chk_render_stats()
    abi.check(ds.L.chk_render_shade(ds.h, C.byref(cam), C.byref(P), int(shader),
                                    C.byref(direct) if direct is not None else None,
                                    image.acc.ptr, image.sq.ptr if image.sq is not None else None,
                                    C.byref(st) if stats else None, stream),
              "chk_render_shade")
    sd = {k: getattr(st, k) for k, _ in abi.chk_render_stats._fields_} if stats else {}
    if not download:
        return image, sd
    a, s2 = image.get()
    return a, s2, sd


def render_direct(ds: abi.DeviceScene, params, direct, camera=None, image: "Image | None" = None,
                  stats=True, stream=None, download=True):
    """chk_render_direct: the reference's shipped shade() (shadeDirectOnly,
    shader.cpp:339-363) with the estimator of `direct` (a chk_direct_params
    or a dict of its fields)."""
    cam = abi.camera_struct(camera or ds.scene.camera)
    P = abi.params_struct(params)
    if isinstance(direct, dict):
        direct = direct_params(**direct)
    if image is None:
        image = Image(cam.width, cam.height, sumsq=bool(P.flags & abi.RENDER_SUMSQ))
    st = abi.chk_render_stats()
    abi.check(ds.L.chk_render_direct(ds.h, C.byref(cam), C.byref(P), C.byref(direct),
                                     image.acc.ptr, image.sq.ptr if image.sq is not None else None,
                                     C.byref(st) if stats else None, stream),
              "chk_render_direct")
    sd = {k: getattr(st, k) for k, _ in abi.chk_render_stats._fields_} if stats else {}
    if not download:
        return image, sd
    a, s2 = image.get()
    return a, s2, sd


class Proposal:
    """Learned light-selection proposal (chk_proposal_t)."""

    def __init__(self, ctx: abi.Context, n_lights: int, seed: int = 1):
        self.ctx = ctx
        self.L = ctx.L
        self.n_lights = int(n_lights)
        h = C.c_void_p()
        abi.check(self.L.chk_proposal_create(ctx.h, self.n_lights, int(seed), C.byref(h)),
                  "chk_proposal_create")
        self.h = h

    def get_params(self):
        p = np.empty(abi.N_PROPOSAL_PARAMS, np.float32)
        abi.check(self.L.chk_proposal_get_params(self.h, p.ctypes.data, p.size), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        abi.check(self.L.chk_proposal_set_params(self.h, p.ctypes.data, p.size), "set_params")

    def eval(self, inputs):
        s = np.ascontiguousarray(inputs, np.float32).reshape(-1, 7)
        k = s.shape[0]
        sd = DevBuf.from_array(s)
        pr, mx, lg = DevBuf(128 * k), DevBuf(4 * k), DevBuf(132 * k)
        abi.check(self.L.chk_proposal_eval(self.h, sd.ptr, k, pr.ptr, mx.ptr, lg.ptr, None),
                  "chk_proposal_eval")
        return (pr.download((k, 32), np.float32), mx.download((k,), np.float32),
                lg.download((k, 33), np.float32))

    def grad(self, inputs, dlogits):
        s = np.ascontiguousarray(inputs, np.float32).reshape(-1, 7)
        k = s.shape[0]
        sd = DevBuf.from_array(s)
        dd = DevBuf.from_array(np.ascontiguousarray(dlogits, np.float32).reshape(k, 33))
        g = DevBuf(4 * abi.N_PROPOSAL_PARAMS)
        abi.check(self.L.chk_proposal_grad(self.h, sd.ptr, dd.ptr, k, g.ptr, None), "chk_proposal_grad")
        return g.download((abi.N_PROPOSAL_PARAMS,), np.float32), g

    def loss_grad(self, records, inputs, probs, occluded, flags=0):
        """chk_proposal_loss_grad: the online-training dlogits (k, 33) of
        `records` (abi.TRAIN_RECORD_DTYPE) drawn with pmfs `probs` (k, 32)
        at network inputs `inputs` (k, 7); occluded: the occlusion flags the
        records' sh_slot index."""
        rec = np.ascontiguousarray(records, abi.TRAIN_RECORD_DTYPE)
        k = rec.shape[0]
        bufs = [DevBuf.from_array(rec),
                DevBuf.from_array(np.ascontiguousarray(inputs, np.float32).reshape(k, 7)),
                DevBuf.from_array(np.ascontiguousarray(probs, np.float32).reshape(k, 32)),
                DevBuf.from_array(np.ascontiguousarray(occluded, np.uint8).reshape(-1))]
        out = DevBuf(132 * k)
        abi.check(self.L.chk_proposal_loss_grad(self.h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                                bufs[3].ptr, k, int(flags), out.ptr, None),
                  "chk_proposal_loss_grad")
        return out.download((k, 33), np.float32)

    def adam(self, grad_buf, lr=1e-3):
        abi.check(self.L.chk_proposal_adam(self.h, grad_buf.ptr, float(lr), None), "chk_proposal_adam")

    def close(self):
        if getattr(self, "h", None):
            self.L.chk_proposal_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Image:
    """FP64 accumulator (H, W, 3) (+ optional sum of squares) in HBM."""

    def __init__(self, width, height, sumsq=False):
        self.W, self.H = int(width), int(height)
        nb = 8 * 3 * self.W * self.H
        self.acc = DevBuf.zeros(nb)
        self.sq = DevBuf.zeros(nb) if sumsq else None

    def zero(self):
        self.acc.zero()
        if self.sq is not None:
            self.sq.zero()

    def get(self):
        a = self.acc.download((self.H, self.W, 3), np.float64)
        s = self.sq.download((self.H, self.W, 3), np.float64) if self.sq is not None else None
        return a, s


def render(ds: abi.DeviceScene, params, camera=None, image: Image | None = None, stats=True,
           count=False, stream=None, download=True):
    """Render into an FP64 image accumulator; returns (accum, sumsq or None,
    stats dict) as numpy arrays (or the Image itself when download=False)."""
    cam = abi.camera_struct(camera or ds.scene.camera)
    P = abi.params_struct(params)
    if image is None:
        image = Image(cam.width, cam.height, sumsq=bool(P.flags & abi.RENDER_SUMSQ))
    st = abi.chk_render_stats()
    if count:
        abi.check(ds.L.chk_render_count(ds.h, C.byref(cam), C.byref(P), image.acc.ptr, C.byref(st),
                                        stream), "chk_render_count")
    else:
        abi.check(ds.L.chk_render(ds.h, C.byref(cam), C.byref(P), image.acc.ptr,
                                  image.sq.ptr if image.sq is not None else None,
                                  C.byref(st) if stats else None, stream), "chk_render")
    sd = {k: getattr(st, k) for k, _ in abi.chk_render_stats._fields_} if (stats or count) else {}
    if not download:
        return image, sd
    a, s = image.get()
    return a, s, sd


def render_learned(ds: abi.DeviceScene, params, proposal: "Proposal", train=None, camera=None,
                   image: Image | None = None, stats=True, stream=None, download=True):
    """chk_render_learned: params with mis_mode MIS_LEARNED; train a dict
    (train_spp, batch, steps_per_bounce, lr) or None for inference only."""
    cam = abi.camera_struct(camera or ds.scene.camera)
    d = dict(params)
    d.setdefault("mis_mode", abi.MIS_LEARNED)
    P = abi.params_struct(d)
    if image is None:
        image = Image(cam.width, cam.height, sumsq=bool(P.flags & abi.RENDER_SUMSQ))
    T = None
    if train is not None:
        T = abi.chk_train_params()
        T.train_spp = int(train.get("train_spp", 32))
        T.batch = int(train.get("batch", 16384))
        T.steps_per_bounce = int(train.get("steps_per_bounce", 16))
        T.lr = float(train.get("lr", 1e-3))
        T.flags = int(train.get("flags", 0))
    st = abi.chk_render_stats()
    abi.check(ds.L.chk_render_learned(ds.h, C.byref(cam), C.byref(P), proposal.h,
                                      C.byref(T) if T is not None else None, image.acc.ptr,
                                      image.sq.ptr if image.sq is not None else None,
                                      C.byref(st) if stats else None, stream),
              "chk_render_learned")
    sd = {k: getattr(st, k) for k, _ in abi.chk_render_stats._fields_} if stats else {}
    if not download:
        return image, sd
    a, s2 = image.get()
    return a, s2, sd


def _requests(items):
    """items: dicts (ds, params, camera=None) -> (chk_request array, Images)."""
    arr = (abi.chk_request * len(items))()
    images = []
    for k, it in enumerate(items):
        ds = it["ds"]
        cam = abi.camera_struct(it.get("camera") or ds.scene.camera)
        P = abi.params_struct(it["params"])
        img = Image(cam.width, cam.height, sumsq=bool(P.flags & abi.RENDER_SUMSQ))
        arr[k].scene = ds.h.value if hasattr(ds.h, "value") else ds.h
        arr[k].camera = cam
        arr[k].params = P
        arr[k].accum = img.acc.ptr
        arr[k].sumsq = img.sq.ptr if img.sq is not None else None
        images.append(img)
    return arr, images


def _tuning(t):
    if t is None:
        return None
    T = abi.chk_batch_tuning()
    for k, _ in abi.chk_batch_tuning._fields_:
        setattr(T, k, int(t.get(k, 0)) if isinstance(t, dict) else int(getattr(t, k)))
    return T


class Batch:
    """A fixed set of requests and their accumulators; run() renders them all
    with chk_render_batch (repeated runs add to the same images, and with
    use_graph reuse the captured graphs)."""

    def __init__(self, ctx: abi.Context, items):
        self.ctx = ctx
        self.n = len(items)
        self.arr, self.images = _requests(items)

    def run(self, tuning=None, stream=None):
        T = _tuning(tuning)
        lat = (C.c_double * self.n)()
        st = abi.chk_batch_stats()
        abi.check(self.ctx.L.chk_render_batch(self.ctx.h, self.arr, self.n,
                                              C.byref(T) if T is not None else None, lat,
                                              C.byref(st), stream), "chk_render_batch")
        return list(lat), {k: getattr(st, k) for k, _ in abi.chk_batch_stats._fields_}

    def get(self):
        return [im.get() for im in self.images]


def render_batch(ctx: abi.Context, items, tuning=None, stream=None):
    """chk_render_batch over items (dicts with ds, params, optional camera).
    Returns ([(accum, sumsq or None)], latency_ms list, stats dict)."""
    b = Batch(ctx, items)
    lat, sd = b.run(tuning, stream)
    return b.get(), lat, sd


def batch_autotune(ctx: abi.Context, items, stream=None):
    """chk_batch_autotune: the fastest tuning (dict) and its batch time."""
    arr, images = _requests(items)
    T = abi.chk_batch_tuning()
    ms = C.c_double(0)
    abi.check(ctx.L.chk_batch_autotune(ctx.h, arr, len(items), C.byref(T), C.byref(ms), stream),
              "chk_batch_autotune")
    del images
    return {k: getattr(T, k) for k, _ in abi.chk_batch_tuning._fields_}, ms.value


def to_display(x):
    """toDisplayValue (main.cpp:27): int(pow(clamp(x,0,1), 1/2.2)*255 + .5)."""
    return (np.power(np.clip(x, 0.0, 1.0), 1.0 / 2.2) * 255 + 0.5).astype(np.int64)


def write_ppm(path, image):
    """P3 PPM exactly as main.cpp:110-117: header "P3\\n%d %d\\n%d\\n", then
    "%d %d %d " per pixel, row 0 first.  image: (H, W, 3) linear radiance
    (already divided by the sample count)."""
    img = np.asarray(image, np.float64)
    H, W, _ = img.shape
    v = to_display(img).reshape(-1, 3)
    with open(path, "w") as f:
        f.write("P3\n%d %d\n%d\n" % (W, H, 255))
        f.write("".join("%d %d %d " % (a, b, c) for a, b, c in v))


def read_ppm6(path):
    """readImage (ppm.cpp:36-115) of a binary P6 file: the first line (at
    most 15 characters, fgets) must start with "P6"; lines starting with '#'
    right after it are skipped; then width and height, then the maximum
    component, which must be 255; the rest of that line is skipped and width
    * height * 3 bytes follow.  Returns (H, W, 3) float64 of byte / 255.0,
    row 0 first.  Raises ValueError where readImage returns 0."""
    data = Path(path).read_bytes()
    nl = data.find(b"\n")
    first = data[:min(15, nl + 1 if nl >= 0 else len(data))]
    if len(first) < 2 or first[:2] != b"P6":
        raise ValueError("Invalid image format (must be 'P6')")
    pos = len(first)
    while pos < len(data) and data[pos:pos + 1] == b"#":
        e = data.find(b"\n", pos)
        pos = len(data) if e < 0 else e + 1

    def scan_int(pos):  # fscanf "%d": whitespace, an optional sign, digits
        while pos < len(data) and data[pos:pos + 1].isspace():
            pos += 1
        m = re.match(rb"[+-]?\d+", data[pos:pos + 32])
        if not m:
            raise ValueError("Invalid image size or rgb component")
        return int(m.group()), pos + m.end()

    w, pos = scan_int(pos)
    h, pos = scan_int(pos)
    maxval, pos = scan_int(pos)
    if maxval != 255:
        raise ValueError("does not have 8-bits components")
    e = data.find(b"\n", pos)
    if e < 0:
        raise ValueError("no pixel data")
    raw = data[e + 1:e + 1 + 3 * w * h]
    if w <= 0 or h <= 0 or len(raw) != 3 * w * h:
        raise ValueError("Error loading image")
    return np.frombuffer(raw, np.uint8).reshape(h, w, 3).astype(np.float64) / 255.0


def write_ppm6(path, image, comment=None):
    """writeImage (ppm.cpp:117-152): "P6\\n", an optional "# comment\\n",
    "%d %d\\n255\\n", then per component (min(v, 1) * 255) converted to a byte
    as the reference's x86 build converts it (truncated toward zero, the low
    8 bits kept).  image: (H, W, 3), row 0 first."""
    img = np.asarray(image, np.float64)
    H, W, _ = img.shape
    v = np.trunc(np.where(img > 1, 1.0, img) * 255).astype(np.int64) & 0xFF
    with open(path, "wb") as f:
        f.write(b"P6\n")
        if comment is not None:
            f.write(b"# " + comment.encode() + b"\n")
        f.write(b"%d %d\n%d\n" % (W, H, 255))
        f.write(v.astype(np.uint8).tobytes())
