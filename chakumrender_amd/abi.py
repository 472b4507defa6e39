"""ctypes binding of libchakum.so (include/chakum_c.h).

The library is the product: every call here goes straight to the HIP
implementation.  Loading fails loudly when libchakum.so is missing; there is
no CPU fallback.  Device buffers are passed as raw pointers (torch tensors
on the GPU are used only as allocations: `tensor.data_ptr()`).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libchakum.so"
LTC_TABLE = PKG / "data" / "ltc_ggx_64.bin"


class ChkError(RuntimeError):
    pass


class chk_material(C.Structure):
    _fields_ = [("brdf", C.c_uint32), ("light", C.c_uint32), ("ks", C.c_float), ("alpha", C.c_float),
                ("phong_exp", C.c_float), ("int_ior", C.c_float), ("ext_ior", C.c_float),
                ("reflectance", C.c_float), ("color", C.c_float * 3), ("radiance", C.c_float * 3)]


class chk_mesh_light(C.Structure):
    _fields_ = [("first_tri", C.c_uint32), ("n_tris", C.c_uint32)]


class chk_sphere(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("radius", C.c_double), ("mat", C.c_uint32),
                ("pad", C.c_uint32)]


class chk_point_source(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("power", C.c_double * 3)]


class chk_env_source(C.Structure):
    _fields_ = [("image", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("axis", C.c_uint32), ("pad", C.c_uint32), ("radiance", C.c_double * 3)]


class chk_scene_desc(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("tri_mat", C.c_void_p), ("mats", C.c_void_p),
                ("mesh_lights", C.c_void_p), ("tri_lights", C.c_void_p), ("spheres", C.c_void_p),
                ("sphere_sources", C.c_void_p), ("point_sources", C.c_void_p),
                ("env", C.POINTER(chk_env_source)),
                ("n_tris", C.c_uint32), ("n_mats", C.c_uint32), ("n_mesh_lights", C.c_uint32),
                ("n_tri_lights", C.c_uint32), ("n_spheres", C.c_uint32),
                ("n_sphere_sources", C.c_uint32), ("n_point_sources", C.c_uint32),
                ("reserved", C.c_uint32)]


def scene_desc(scene):
    """chk_scene_desc of a scenes.Scene and the arrays it points into (keep
    them alive for the call)."""
    arrs = {
        "tris": np.ascontiguousarray(scene.tris, np.float32).reshape(-1, 9),
        "tri_mat": np.ascontiguousarray(scene.tri_mat, np.uint32),
        "mats": np.ascontiguousarray(scene.mats),
        "mesh_lights": np.ascontiguousarray(scene.lights),
        "tri_lights": np.ascontiguousarray(getattr(scene, "tri_lights", np.zeros(0)), np.uint32),
        "spheres": np.ascontiguousarray(getattr(scene, "spheres", np.zeros(0, _SPHERE_NP))),
        "sphere_sources": np.ascontiguousarray(getattr(scene, "sphere_sources", np.zeros(0)),
                                               np.uint32),
        "point_sources": np.ascontiguousarray(getattr(scene, "point_sources",
                                                      np.zeros(0, _POINT_NP))),
    }
    assert arrs["spheres"].dtype.itemsize == C.sizeof(chk_sphere)
    assert arrs["point_sources"].dtype.itemsize == C.sizeof(chk_point_source)
    d = chk_scene_desc()
    for k, a in arrs.items():
        setattr(d, k, a.ctypes.data if a.shape[0] else None)
    env = getattr(scene, "env", None)
    if env is not None:
        img = np.ascontiguousarray(env.image, np.float64)
        assert img.ndim == 3 and img.shape[2] == 3
        e = chk_env_source()
        e.image = img.ctypes.data
        e.height, e.width = img.shape[0], img.shape[1]
        e.axis = int(env.axis)
        e.radiance[:] = [float(v) for v in env.radiance]
        arrs["env_image"], arrs["env"] = img, e
        d.env = C.pointer(e)
    d.n_tris = arrs["tris"].shape[0]
    d.n_mats = arrs["mats"].shape[0]
    d.n_mesh_lights = arrs["mesh_lights"].shape[0]
    d.n_tri_lights = arrs["tri_lights"].shape[0]
    d.n_spheres = arrs["spheres"].shape[0]
    d.n_sphere_sources = arrs["sphere_sources"].shape[0]
    d.n_point_sources = arrs["point_sources"].shape[0]
    return d, arrs


_SPHERE_NP = np.dtype([("center", "<f8", (3,)), ("radius", "<f8"), ("mat", "<u4"), ("pad", "<u4")])
_POINT_NP = np.dtype([("pos", "<f8", (3,)), ("power", "<f8", (3,))])


class chk_camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("up", C.c_float * 3),
                ("fovx_deg", C.c_float), ("fovy_deg", C.c_float), ("width", C.c_uint32),
                ("height", C.c_uint32)]


class chk_render_params(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("sample_offset", C.c_uint32), ("max_depth", C.c_uint32),
                ("mis_mode", C.c_uint32), ("jitter_mode", C.c_uint32), ("seed_lo", C.c_uint32),
                ("seed_hi", C.c_uint32), ("flags", C.c_uint32), ("window", C.c_uint32 * 4)]


class chk_scene_info(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("n_tris_bvh", C.c_uint32), ("n_nodes", C.c_uint32),
                ("n_leaves", C.c_uint32), ("max_depth", C.c_uint32), ("stack_bound", C.c_uint32),
                ("stack_size", C.c_uint32), ("n_lights", C.c_uint32), ("device_bytes", C.c_uint64),
                ("sah_cost", C.c_double), ("build_ms", C.c_double)]


class chk_request(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", chk_camera), ("params", chk_render_params),
                ("accum", C.c_void_p), ("sumsq", C.c_void_p)]


class chk_batch_tuning(C.Structure):
    _fields_ = [("chunk_paths", C.c_uint32), ("trace_block", C.c_uint32),
                ("shade_block", C.c_uint32), ("use_graph", C.c_uint32)]


class chk_batch_stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays_extend", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("chunks", C.c_uint32), ("segments", C.c_uint32), ("batch_ms", C.c_double),
                ("capture_ms", C.c_double), ("graph_cached", C.c_uint32), ("reserved", C.c_uint32)]


class chk_train_params(C.Structure):
    _fields_ = [("train_spp", C.c_uint32), ("batch", C.c_uint32), ("steps_per_bounce", C.c_uint32),
                ("lr", C.c_float), ("flags", C.c_uint32)]


class chk_train_record(C.Structure):
    _fields_ = [("light", C.c_uint32), ("gl", C.c_float), ("pl", C.c_float), ("pbl", C.c_float),
                ("fb", C.c_float), ("le", C.c_float), ("plb", C.c_float), ("pbb", C.c_float),
                ("sh_slot", C.c_uint32), ("light_b", C.c_uint32)]


# the same record as a numpy dtype (arrays of records for chk_proposal_loss_grad)
TRAIN_RECORD_DTYPE = np.dtype([(n, "<u4" if t is C.c_uint32 else "<f4")
                               for n, t in chk_train_record._fields_])


class chk_render_stats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays_extend", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("chunks", C.c_uint32), ("extend_launches", C.c_uint32), ("extend_ms", C.c_double),
                ("shadow_ms", C.c_double), ("shade_ms", C.c_double), ("other_ms", C.c_double),
                ("proposal_ms", C.c_double), ("train_steps", C.c_uint32), ("stack_retries", C.c_uint32)]


class chk_direct_params(C.Structure):
    _fields_ = [("estimator", C.c_uint32), ("cv_samples", C.c_uint32), ("flags", C.c_uint32),
                ("estimate", C.c_uint32), ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32)]


assert C.sizeof(chk_material) == 56
assert C.sizeof(chk_render_params) == 48

# flags (chakum_types.h, chakum_c.h)
RENDER_PHYSICAL_EMITTER = 0x1
RENDER_SUMSQ = 0x2
RENDER_PROFILE = 0x100
RENDER_ONE_STREAM = 0x200
JITTER_BOX, JITTER_REFERENCE = 0, 1
MIS_BALANCE, MIS_LEARNED = 0, 1
TRAIN_FREEZE_MIX, TRAIN_FREEZE_LIGHTS, TRAIN_LIGHTS_CHI2 = 1, 2, 4
DIRECT_ANALYTIC, DIRECT_ANALYTIC_CV, DIRECT_MIS, DIRECT_MISCV = 0, 1, 2, 3
DIRECT_VISIBILITY = 0x1
MIS_LIGHT_SAMPLES, MIS_BSDF_SAMPLES = 10, 60
SHADE_DIRECT_ONLY, SHADE_EXPLICIT, SHADE_IMPLICIT = 0, 1, 2
SHADE_MAX_TRIES = 64
ACCEL_BVH, ACCEL_DUMMY = 0, 1
SHADE_STREAM = 0x53480000  # CHK_STREAM_SHADE

# Every exported symbol of include/chakum_c.h: name -> (restype, argtypes)
_vp = C.c_void_p
_u32 = C.c_uint32
SIGNATURES = {
    "chk_status_string": (C.c_char_p, [C.c_int]),
    "chk_last_error": (C.c_char_p, []),
    "chk_version": (C.c_char_p, []),
    "chk_device_alloc": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(_vp)]),
    "chk_device_free": (C.c_int, [_vp]),
    "chk_memcpy": (C.c_int, [_vp, _vp, C.c_size_t, C.c_int, _vp]),
    "chk_memzero": (C.c_int, [_vp, C.c_size_t, _vp]),
    "chk_stream_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "chk_stream_destroy": (C.c_int, [_vp]),
    "chk_stream_sync": (C.c_int, [_vp]),
    "chk_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "chk_destroy": (C.c_int, [_vp]),
    "chk_set_ltc_table": (C.c_int, [_vp, _vp, C.c_size_t]),
    "chk_scene_create": (C.c_int, [_vp, _vp, _u32, _vp, _vp, _u32, _vp, _u32, C.POINTER(_vp)]),
    "chk_scene_create_desc": (C.c_int, [_vp, C.POINTER(chk_scene_desc), C.POINTER(_vp)]),
    "chk_scene_destroy": (C.c_int, [_vp]),
    "chk_scene_set_trace_stack": (C.c_int, [_vp, _u32]),
    "chk_scene_get_info": (C.c_int, [_vp, C.POINTER(chk_scene_info)]),
    "chk_scene_set_accel": (C.c_int, [_vp, _u32]),
    "chk_trace_closest": (C.c_int, [_vp, _vp, _u32, _vp, _vp]),
    "chk_trace_any": (C.c_int, [_vp, _vp, _u32, _vp, _vp]),
    "chk_trace_spheres": (C.c_int, [_vp, _u32, _vp, _u32, _vp, _vp]),
    "chk_brdf_eval": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp]),
    "chk_brdf_sample": (C.c_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp]),
    "chk_ltc_eval": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _vp]),
    "chk_render": (C.c_int, [_vp, C.POINTER(chk_camera), C.POINTER(chk_render_params), _vp, _vp,
                             C.POINTER(chk_render_stats), _vp]),
    "chk_render_noalloc": (C.c_int, [_vp, C.POINTER(chk_camera), C.POINTER(chk_render_params), _vp,
                                     _vp, C.POINTER(chk_render_stats), _vp]),
    "chk_render_count": (C.c_int, [_vp, C.POINTER(chk_camera), C.POINTER(chk_render_params), _vp,
                                   C.POINTER(chk_render_stats), _vp]),
    "chk_render_batch": (C.c_int, [_vp, C.POINTER(chk_request), _u32, C.POINTER(chk_batch_tuning),
                                   C.POINTER(C.c_double), C.POINTER(chk_batch_stats), _vp]),
    "chk_batch_autotune": (C.c_int, [_vp, C.POINTER(chk_request), _u32, C.POINTER(chk_batch_tuning),
                                     C.POINTER(C.c_double), _vp]),
    "chk_render_learned": (C.c_int, [_vp, C.POINTER(chk_camera), C.POINTER(chk_render_params), _vp,
                                     C.POINTER(chk_train_params), _vp, _vp,
                                     C.POINTER(chk_render_stats), _vp]),
    "chk_mis_weights": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp]),
    "chk_direct_light": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _u32, C.POINTER(chk_direct_params),
                                   _vp, _vp]),
    "chk_render_direct": (C.c_int, [_vp, C.POINTER(chk_camera), C.POINTER(chk_render_params),
                                    C.POINTER(chk_direct_params), _vp, _vp,
                                    C.POINTER(chk_render_stats), _vp]),
    "chk_render_shade": (C.c_int, [_vp, C.POINTER(chk_camera), C.POINTER(chk_render_params), _u32,
                                   C.POINTER(chk_direct_params), _vp, _vp,
                                   C.POINTER(chk_render_stats), _vp]),
    "chk_proposal_create": (C.c_int, [_vp, _u32, C.c_uint64, C.POINTER(_vp)]),
    "chk_proposal_destroy": (C.c_int, [_vp]),
    "chk_proposal_set_params": (C.c_int, [_vp, _vp, C.c_size_t]),
    "chk_proposal_get_params": (C.c_int, [_vp, _vp, C.c_size_t]),
    "chk_proposal_eval": (C.c_int, [_vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "chk_proposal_grad": (C.c_int, [_vp, _vp, _vp, _u32, _vp, _vp]),
    "chk_proposal_adam": (C.c_int, [_vp, _vp, C.c_float, _vp]),
    "chk_proposal_loss_grad": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp]),
}
N_PROPOSAL_PARAMS = 14625

_lib = None


def load(path: str | os.PathLike | None = None):
    """Load libchakum.so (raises ChkError when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ChkError(f"{p} is not built: run `make -C chakumrender_amd/csrc` "
                       "(or __graft_entry__.build())")
    L = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def check(status: int, what: str = ""):
    if status != 0:
        L = load()
        raise ChkError(f"{what}: {L.chk_status_string(status).decode()}: "
                       f"{L.chk_last_error().decode()}")


def camera_struct(cam) -> chk_camera:
    c = chk_camera()
    c.pos[:] = [float(v) for v in cam.pos]
    c.dir[:] = [float(v) for v in cam.dir]
    c.up[:] = [float(v) for v in cam.up]
    c.fovx_deg = cam.fovx_deg
    c.fovy_deg = cam.fovy_deg
    c.width = cam.width
    c.height = cam.height
    return c


def params_struct(p) -> chk_render_params:
    if isinstance(p, chk_render_params):
        return p
    d = dict(p)
    r = chk_render_params()
    r.spp = int(d.get("spp", 1))
    r.sample_offset = int(d.get("sample_offset", 0))
    r.max_depth = int(d.get("max_depth", 4))
    r.mis_mode = int(d.get("mis_mode", MIS_BALANCE))
    r.jitter_mode = int(d.get("jitter_mode", JITTER_BOX))
    if "seed_lo" in d or "seed_hi" in d:
        r.seed_lo = int(d.get("seed_lo", 0)) & 0xFFFFFFFF
        r.seed_hi = int(d.get("seed_hi", 0)) & 0xFFFFFFFF
    else:
        seed = int(d.get("seed", 0x5EED))
        r.seed_lo = seed & 0xFFFFFFFF
        r.seed_hi = (seed >> 32) & 0xFFFFFFFF
    r.flags = int(d.get("flags", 0))
    for i, v in enumerate(d.get("window", (0, 0, 0, 0))):
        r.window[i] = int(v)
    return r


class DevBuf:
    """Device allocation made by libchakum's own HIP runtime."""

    def __init__(self, nbytes: int, device: int = 0):
        self.L = load()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(self.L.chk_device_alloc(device, self.nbytes, C.byref(p)), "chk_device_alloc")
        self.p = p

    @property
    def ptr(self):
        return self.p

    @classmethod
    def from_array(cls, a, device: int = 0):
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 16), device)
        b.upload(a)
        return b

    @classmethod
    def zeros(cls, nbytes: int, device: int = 0):
        b = cls(max(int(nbytes), 16), device)
        b.zero()
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(self.L.chk_memcpy(self.p, a.ctypes.data, a.nbytes, 0, stream), "chk_memcpy h2d")
        check(self.L.chk_stream_sync(stream), "chk_stream_sync")

    def download(self, shape, dtype, stream=None):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(self.L.chk_memcpy(out.ctypes.data, self.p, out.nbytes, 1, stream), "chk_memcpy d2h")
        return out

    def zero(self, stream=None):
        check(self.L.chk_memzero(self.p, self.nbytes, stream), "chk_memzero")

    def free(self):
        if getattr(self, "p", None):
            self.L.chk_device_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync(stream=None):
    check(load().chk_stream_sync(stream), "chk_stream_sync")


class Context:
    """chk_ctx_t on one device (with the LTC table installed)."""

    def __init__(self, device: int = 0, ltc_table: bool = True):
        self.L = load()
        h = C.c_void_p()
        check(self.L.chk_create(device, C.byref(h)), "chk_create")
        self.h = h
        if ltc_table:
            tab = np.ascontiguousarray(np.fromfile(LTC_TABLE, "<f8"))
            check(self.L.chk_set_ltc_table(self.h, tab.ctypes.data, tab.size), "chk_set_ltc_table")

    def close(self):
        if getattr(self, "h", None):
            self.L.chk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """chk_scene_t: triangles, materials, lights and the BVH resident in HBM."""

    def __init__(self, ctx: Context, scene):
        self.ctx = ctx
        self.L = ctx.L
        self.scene = scene
        d, keep = scene_desc(scene)
        h = C.c_void_p()
        check(self.L.chk_scene_create_desc(ctx.h, C.byref(d), C.byref(h)), "chk_scene_create_desc")
        del keep
        self.h = h

    def set_accel(self, kind: int):
        """chk_scene_set_accel: ACCEL_BVH or ACCEL_DUMMY for the trace entry points."""
        check(self.L.chk_scene_set_accel(self.h, int(kind)), "chk_scene_set_accel")

    def set_trace_stack(self, entries: int):
        """chk_scene_set_trace_stack: 8, 16 or 32 LDS stack entries per lane."""
        check(self.L.chk_scene_set_trace_stack(self.h, int(entries)), "chk_scene_set_trace_stack")

    def info(self) -> dict:
        i = chk_scene_info()
        check(self.L.chk_scene_get_info(self.h, C.byref(i)), "chk_scene_get_info")
        return {k: getattr(i, k) for k, _ in chk_scene_info._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.L.chk_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
