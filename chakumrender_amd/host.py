"""Host-side mirror of the reference's per-sample-path interfaces over the
C ABI, batched.  Names and argument meanings follow the reference:

  BvhAccel.intersectT   bvhAccel.cpp:200-249   closest hit (t, N, id)
  shadow                shader.cpp:76-97        1.0 if unoccluded else 0.0
  MaterialType.brdf / pdfEval / getBrdfDirectionSampleRandomized
                        materialTypes.cpp:9-153
  LightSource.getLightFromMeshSource_Analytic
                        lightSources_CV.cpp:323-333
  render                main.cpp:62-94 around shade() (shader.cpp:362)
  write_ppm             main.cpp:27,98-117

Differences a caller sees: rays are batched numpy arrays; ids are ORIGINAL
triangle indices (the reference returns indices into its reordered array);
random numbers for sampling are explicit inputs.  Device memory comes from
torch (allocation only); every computation runs in libchakum.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t

        if not t.cuda.is_available():
            raise abi.ChkError("no GPU visible: libchakum needs a HIP device")
        _torch = t
    return _torch


def to_dev(a, dtype=None):
    t = torch()
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return t.from_numpy(a).cuda()


def empty_dev(shape, dtype):
    t = torch()
    return t.empty(shape, dtype=dtype, device="cuda")


def stream_ptr():
    return C.c_void_p(torch().cuda.current_stream().cuda_stream)


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
        abi.check(self.L.chk_trace_closest(self.ds.h, rays_dev, n, hits_dev,
                                           stream if stream is not None else stream_ptr()),
                  "chk_trace_closest")

    def trace_any(self, rays_dev, n, occ_dev, stream=None):
        abi.check(self.L.chk_trace_any(self.ds.h, rays_dev, n, occ_dev,
                                       stream if stream is not None else stream_ptr()),
                  "chk_trace_any")

    def intersectT(self, o, d):
        """Closest hit for rays (o, d): returns (hit bool, t, id) numpy arrays;
        t = 1e30 and id = -1 on a miss (the reference: INF, 0xFFFFFFFF)."""
        t = torch()
        rays = make_rays(o, d)
        n = rays.shape[0]
        rd = to_dev(rays)
        hd = empty_dev((n, 4), t.float32)
        self.trace_closest(C.c_void_p(rd.data_ptr()), n, C.c_void_p(hd.data_ptr()))
        h = hd.cpu().numpy()
        tri = h[:, 1].view(np.int32)
        return tri >= 0, h[:, 0], tri

    def shadow(self, o, d, dist):
        """shader.cpp:76-97: 0.0 if some triangle is hit with t + 1e-5 < dist."""
        t = torch()
        rays = make_rays(o, d, tmax=np.asarray(dist, np.float32) - np.float32(1e-5))
        n = rays.shape[0]
        rd = to_dev(rays)
        od = empty_dev((n,), t.uint8)
        self.trace_any(C.c_void_p(rd.data_ptr()), n, C.c_void_p(od.data_ptr()))
        return 1.0 - od.cpu().numpy().astype(np.float64)


class MaterialType:
    """Batched MaterialType (materialTypes.h:20-120) over a material table."""

    def __init__(self, mats):
        self.mats = np.ascontiguousarray(mats)
        self.L = abi.load()

    def _args(self, mat_idx, *vecs):
        t = torch()
        md = to_dev(self.mats.view(np.uint8))
        idx = to_dev(np.asarray(mat_idx, np.uint32).view(np.int32))
        vs = [to_dev(np.asarray(v, np.float32).reshape(-1, 3)) for v in vecs]
        return t, md, idx, vs

    def brdf_pdf(self, mat_idx, n, wo_ref, wo, wi):
        """(BasePrimitive::brdf, MaterialType::pdfEval) per row."""
        t, md, idx, (nd, rd, od, wd) = self._args(mat_idx, n, wo_ref, wo, wi)
        k = idx.shape[0]
        f = empty_dev((k,), t.float32)
        p = empty_dev((k,), t.float32)
        abi.check(self.L.chk_brdf_eval(md.data_ptr(), self.mats.shape[0], idx.data_ptr(),
                                       nd.data_ptr(), rd.data_ptr(), od.data_ptr(), wd.data_ptr(),
                                       k, f.data_ptr(), p.data_ptr(), stream_ptr()),
                  "chk_brdf_eval")
        return f.cpu().numpy(), p.cpu().numpy()

    def sample(self, mat_idx, n, wo_ref, wo, u):
        """getBrdfDirectionSampleRandomized with explicit uniforms u (k, 3)."""
        t, md, idx, (nd, rd, od, ud) = self._args(mat_idx, n, wo_ref, wo, u)
        k = idx.shape[0]
        wi = empty_dev((k, 3), t.float32)
        p = empty_dev((k,), t.float32)
        abi.check(self.L.chk_brdf_sample(md.data_ptr(), self.mats.shape[0], idx.data_ptr(),
                                         nd.data_ptr(), rd.data_ptr(), od.data_ptr(), ud.data_ptr(),
                                         k, wi.data_ptr(), p.data_ptr(), stream_ptr()),
                  "chk_brdf_sample")
        return wi.cpu().numpy(), p.cpu().numpy()


class LightSource:
    """Mesh-light analytic shading (lightSources_CV.cpp:221-333) on the GPU."""

    def __init__(self, ctx: abi.Context, light_tris, radiance):
        self.ctx = ctx
        self.L = ctx.L
        self.tris = to_dev(np.asarray(light_tris, np.float32).reshape(-1, 9))
        self.rad = to_dev(np.asarray(radiance, np.float32).reshape(-1, 3))

    def getLightFromMeshSource_Analytic(self, x, n, v, alpha, ks, out=None):
        t = torch()
        xs = [a if hasattr(a, "data_ptr") else to_dev(np.asarray(a, np.float32))
              for a in (x, n, v, alpha, ks)]
        k = xs[0].shape[0]
        rgb = out if out is not None else empty_dev((k, 3), t.float32)
        abi.check(self.L.chk_ltc_eval(self.ctx.h, *[a.data_ptr() for a in xs], k,
                                      self.tris.data_ptr(), self.rad.data_ptr(),
                                      self.tris.shape[0], rgb.data_ptr(), stream_ptr()),
                  "chk_ltc_eval")
        return rgb


def render(ds: abi.DeviceScene, params, camera=None, accum=None, sumsq=None, stats=True,
           count=False):
    """Render into an FP64 accumulator (H, W, 3) on the GPU; returns
    (accum tensor, sumsq tensor or None, stats dict)."""
    t = torch()
    cam = abi.camera_struct(camera or ds.scene.camera)
    P = abi.params_struct(params)
    H, W = cam.height, cam.width
    if accum is None:
        accum = t.zeros((H, W, 3), dtype=t.float64, device="cuda")
    if (P.flags & abi.RENDER_SUMSQ) and sumsq is None:
        sumsq = t.zeros((H, W, 3), dtype=t.float64, device="cuda")
    st = abi.chk_render_stats()
    if count:
        abi.check(ds.L.chk_render_count(ds.h, C.byref(cam), C.byref(P), accum.data_ptr(),
                                        C.byref(st), stream_ptr()), "chk_render_count")
    else:
        abi.check(ds.L.chk_render(ds.h, C.byref(cam), C.byref(P), accum.data_ptr(),
                                  sumsq.data_ptr() if sumsq is not None else None,
                                  C.byref(st) if stats else None, stream_ptr()), "chk_render")
    sd = {k: getattr(st, k) for k, _ in abi.chk_render_stats._fields_} if (stats or count) else {}
    return accum, sumsq, sd


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
