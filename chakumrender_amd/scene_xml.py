"""Scene files of the reference's XML front end -> Scene for libchakum.

Mirrors SceneParser (sceneParser.cpp) and the scene preparation of
prepareScene.cpp in main.cpp's order (main.cpp:33-45): screen and camera,
then objects (MESHOBJECT triangles, SPHEREOBJECT spheres, the dummy sphere
and dummy triangle, prepareScene.cpp:180-241), then light sources
(MESHSOURCE mesh lights, TRISOURCE triangle lights, POINTSOURCE,
SPHERESOURCE, prepareScene.cpp:54-157).  Mesh and triangle lights append
their triangles to the scene after the objects, as
LightSource::addMSource / addTSource do (lightSources.h:192-203).

Geometry is computed in FP64 with the reference's operations and order
(quaternion rotations of transformations.cpp:26-42 applied as
R_1 ... R_k S v + t, tempToColor of mathPrimitives.cpp:32-60) and then
rounded to the FP32 the C ABI takes; the FP64 vertices stay in
scene.meta["tris64"].  Sphere objects, the dummy sphere and the sphere
sources' visible spheres fill scene.spheres in vSphereList's order (each
with a material of its own after the triangles'), triangle lights
scene.tri_lights, sphere sources scene.sphere_sources (their spheres) and
point sources scene.point_sources, as LightSource keeps them.
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

from . import scenes

PI = 3.14159265359  # mathPrimitives.h:4
BRDF_NAMES = {"GGX": scenes.GGX, "BECKMANN": scenes.BECKMANN, "PHONG": scenes.PHONG,
              "GGXAPPROX": scenes.GGXAPPROX, "CLASSIC_PHONG": scenes.CLASSIC_PHONG}


# ------------------------------------------------------------------ text fields
def _text(el, name):
    """FirstChildElement(name)->FirstChild() as text (sceneParser.cpp:6-45)."""
    c = el.find(name)
    if c is None or c.text is None:
        return None
    return c.text


def _scan(text, n):
    """sscanf(text, "%lf %lf ...") for n numbers: None unless all n read."""
    out = []
    for tok in text.split()[:n]:
        m = re.match(r"[+-]?((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?|inf(inity)?|nan)", tok, re.I)
        if not m:
            break
        out.append(float(m.group(0)))
        if m.end() != len(tok):
            break
    return out if len(out) == n else None


def _vec3(el, name):
    t = _text(el, name)
    v = None if t is None else _scan(t, 3)
    return None if v is None else tuple(v)


def _double(el, name):
    t = _text(el, name)
    v = None if t is None else _scan(t, 1)
    return None if v is None else v[0]


def _uint(el, name):
    t = _text(el, name)
    if t is None:
        return None
    m = re.match(r"\s*(\d+)", t)
    return int(m.group(1)) if m else None


def temp_to_color(temperature):
    """tempToColor (mathPrimitives.cpp:32-60): black-body colour of a
    temperature in kelvin (Helland's fit), channels in [0, 1]."""
    t = min(max(temperature, 1000.0), 40000.0) / 100.0
    if t <= 66:
        red = 255.0
        green = 99.4708025861 * math.log(t) - 161.1195681661
        blue = 0.0 if t <= 19 else 138.5177312231 * math.log(t - 10.0) - 305.0447927307
    else:
        red = 329.698727446 * math.pow(t - 60.0, -0.1332047592)
        green = 288.1221695283 * math.pow(t - 60, -0.0755148492)
        blue = 255.0

    def c(x):
        return min(max(x, 0.0), 255.0) / 255.0

    return (c(red), c(green), c(blue))


def _color(el):
    """parseColor (sceneParser.cpp:57-72): COLOR, else TEMPERATURE."""
    col = _vec3(el, "COLOR")
    if col is not None:
        return col
    t = _double(el, "TEMPERATURE")
    return None if t is None else temp_to_color(t)


# ------------------------------------------------------------------ matrices (mat3)
def _mat9(*m):
    """mat3's 9-argument constructor takes the matrix column by column
    (mathPrimitives.h:149-163); returned row-major."""
    return [[m[0], m[3], m[6]], [m[1], m[4], m[7]], [m[2], m[5], m[8]]]


def _mul(p, q):
    """mat3::mul(mat3) (mathPrimitives.h:49-63)."""
    return [[p[i][0] * q[0][j] + p[i][1] * q[1][j] + p[i][2] * q[2][j] for j in range(3)]
            for i in range(3)]


def _mulv(m, v):
    """mat3::mul(Vec) (mathPrimitives.h:82-86)."""
    return tuple(m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2] for i in range(3))


def _rotate(axis, theta):
    """rotate(axis, theta) (transformations.cpp:26-42): unit quaternion to matrix."""
    ax, ay, az = axis
    inv = 1.0 / math.sqrt(ax * ax + ay * ay + az * az)  # Vec::norm
    qx, qy, qz = ax * inv, ay * inv, az * inv
    theta /= 2.0
    q0 = math.cos(theta)
    s = math.sin(theta)
    qx, qy, qz = qx * s, qy * s, qz * s
    return _mat9(1.0 - 2.0 * qy * qy - 2.0 * qz * qz, 2.0 * qx * qy + 2.0 * qz * q0,
                 2.0 * qx * qz - 2.0 * qy * q0,
                 2.0 * qx * qy - 2.0 * qz * q0, 1.0 - 2.0 * qx * qx - 2.0 * qz * qz,
                 2.0 * qy * qz + 2.0 * qx * q0,
                 2.0 * qx * qz + 2.0 * qy * q0, 2.0 * qy * qz - 2.0 * qx * q0,
                 1.0 - 2.0 * qx * qx - 2 * qy * qy)


def _mesh_matrix(el):
    """SPVecToMat(scale) then RotSeqToMat(rotations).mul (prepareScene.cpp:40-52)."""
    sc = _vec3(el, "SCALE") if el.find("SCALE") is not None else (1.0, 1.0, 1.0)
    if sc is None:
        raise ValueError("No Mesh Scaling Information.")
    m = _mat9(sc[0], 0, 0, 0, sc[1], 0, 0, 0, sc[2])
    rot = _mat9(1, 0, 0, 0, 1, 0, 0, 0, 1)
    rs = el.find("ROTSEQ")
    if rs is not None:
        rr = rs.findall("R")
        if not rr:
            raise ValueError("No rotaions specified.")
        for r in rr:
            axis, deg = _vec3(r, "AXIS"), _double(r, "ROT")
            if axis is None or deg is None:
                raise ValueError("bad rotation in ROTSEQ")
            rot = _mul(rot, _rotate(axis, deg * PI / 180.0))
    return _mul(rot, m)


def _translate(el):
    if el.find("TRANSLATE") is None:
        return (0.0, 0.0, 0.0)
    t = _vec3(el, "TRANSLATE")
    if t is None:
        raise ValueError("No Mesh Translation Information.")
    return t


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _material_fields(el):
    """parseMaterialProperties (sceneParser.cpp:122-148)."""
    alpha, ks = _double(el, "ALPHA"), _double(el, "SPECCOEF")
    refl, brdf = _double(el, "REFLECTANCE"), _text(el, "BRDF")
    if alpha is None or ks is None or refl is None or brdf is None:
        raise ValueError("material needs ALPHA, SPECCOEF, REFLECTANCE and BRDF")
    iior = _double(el, "INTIOR") if el.find("INTIOR") is not None else 10000.0
    eior = _double(el, "EXTIOR") if el.find("EXTIOR") is not None else 1.0
    return alpha, ks, refl, brdf, iior, eior


def _make_material(col, alpha, ks, refl, brdf_name, iior, eior):
    """stringToBRDF (prepareScene.cpp:159-178) and the MaterialType
    constructors (materialTypes.h:67-98): |ks| clamped to 1; a classic-Phong
    material takes |alpha| as its exponent and no IORs."""
    b = BRDF_NAMES.get(brdf_name, scenes.BRDF_NONE)
    ks = min(abs(ks), 1.0)
    if b == scenes.CLASSIC_PHONG:
        return scenes._material(b, ks=ks, phong_exp=abs(alpha), int_ior=0.0, ext_ior=0.0,
                                reflectance=refl, color=col)
    return scenes._material(b, ks=ks, alpha=alpha, int_ior=iior, ext_ior=eior, reflectance=refl,
                            color=col)


def _emitter(col, intensity):
    """TriLight material: radiance = col.maxnorm() * intensity (lightSources.h:12-16,
    130-138), PLANAR, colour 0, reflectance 0."""
    m = max(col[0], col[1], col[2])
    rad = (col[0] / m * intensity, col[1] / m * intensity, col[2] / m * intensity)
    return scenes._material(scenes.BRDF_NONE, int_ior=0.0, ext_ior=0.0, reflectance=0.0,
                            color=(0, 0, 0), light=scenes.LIGHT_PLANAR, radiance=rad)


def _triangles(el, with_material):
    """parseMesh (sceneParser.cpp:410-455) + the transform of prepareScene."""
    trs = el.findall("TR")
    if not trs:
        raise ValueError("No Triangle fins found in mesh.")
    trans, m = _translate(el), _mesh_matrix(el)
    out = []
    for k, t in enumerate(trs):
        v = [_vec3(t, n) for n in ("V1", "V2", "V3")]
        col = _color(t)
        if any(x is None for x in v) or col is None:
            raise ValueError(f"Triangle: {k} couldn't be parsed.")
        mat = _material_fields(t) if with_material else None
        rad = _double(t, "RADIANCE") if t.find("RADIANCE") is not None else 0.0
        verts = [_add(_mulv(m, p), trans) for p in v]
        out.append((verts, col, mat, rad))
    return out


def load_scene_xml(path_or_text):
    """A reference scene file as a Scene (FP32 triangles with one material
    each, spheres, mesh, triangle, sphere and point sources, camera).
    path_or_text: a path or the XML text."""
    src = str(path_or_text)
    text = Path(src).read_text() if not src.lstrip().startswith("<") else src
    text = re.sub(r"<\?xml[^>]*\?>", "", text)
    doc = ET.fromstring("<SCENEFILE>" + text + "</SCENEFILE>")  # several top-level elements
    # screen and camera (prepareScreenCamera, prepareScene.cpp:12-36)
    screen, cam = doc.find("SCREEN"), doc.find("CAMERA")
    if screen is None or cam is None:
        raise ValueError("No Screen or Camera Information.")
    W, H = _uint(screen, "WIDTH"), _uint(screen, "HEIGHT")
    pos, look, up = _vec3(cam, "POSITION"), _vec3(cam, "LOOKAT"), _vec3(cam, "UP")
    fovx, fovy = _double(cam, "FOVX"), _double(cam, "FOVY")
    if None in (W, H, pos, look, up, fovx, fovy):
        raise ValueError("incomplete SCREEN or CAMERA")
    tris, mats = [], []
    # objects (prepareObjects, prepareScene.cpp:180-241)
    for obj in doc.findall("MESHOBJECT"):
        for verts, col, (alpha, ks, refl, brdf, iior, eior), _ in _triangles(obj, True):
            tris.append(verts)
            mats.append(_make_material(col, alpha, ks, refl, brdf, iior, eior))
    spheres = []
    for obj in doc.findall("SPHEREOBJECT"):
        sps = obj.findall("SP")
        if not sps:
            raise ValueError("No Sphere objects found in mesh.")
        trans = _translate(obj)
        for s in sps:
            r, c, col = _double(s, "RADIUS"), _vec3(s, "POSITION"), _color(s)
            if r is None or c is None or col is None:
                raise ValueError("Sphere couldn't be parsed.")
            spheres.append(dict(center=_add(c, trans), radius=r, color=col,
                                material=_make_material(col, *_material_fields(s))))
    spheres.append(dict(center=(27.0, 1e5, 47.0), radius=1.0, color=(.999, .999, .999),
                        material=_make_material((.999, .999, .999), 200.2, 1.0, 1.0,
                                                "CLASSIC_PHONG", 0, 0)))
    # prepareScene.cpp:241: Triangle(1e6, 1e6, 1e6, ...) converts each 1e6
    # through Vec(double x) into (1e6, 0, 0)
    tris.append([(1e6, 0.0, 0.0)] * 3)
    mats.append(_make_material((0.0, 0, 0.9), 1.0, 0.5, 0.9, "CLASSIC_PHONG", 0, 0))
    # light sources (prepareLightSources, prepareScene.cpp:54-157)
    lights = []
    for src_el in doc.findall("MESHSOURCE"):
        first = len(tris)
        for verts, col, _, rad in _triangles(src_el, False):
            tris.append(verts)
            mats.append(_emitter(col, rad))
        lights.append((first, len(tris) - first))
    tri_lights = []
    for src_el in doc.findall("TRISOURCE"):
        verts, col, _, rad = _triangles(src_el, False)[0]
        tri_lights.append(len(tris))
        tris.append(verts)
        mats.append(_emitter(col, rad))
    points = []
    for s in doc.findall("POINTSOURCE"):
        # parsePointSource (sceneParser.cpp:76-90); PointSource(pos, col,
        # power): BaseSource radiance = col.maxnorm() * power
        c, col, pw = _vec3(s, "POSITION"), _color(s), _double(s, "POWER")
        if c is None or col is None or pw is None:
            raise ValueError("Point Source couldn't be parsed.")
        m = max(col)
        points.append((c, tuple(ch / m * pw for ch in col)))
    sphere_sources = []
    for s in doc.findall("SPHERESOURCE"):
        r, c, col = _double(s, "RADIUS"), _vec3(s, "POSITION"), _color(s)
        rad = _double(s, "RADIANCE")
        if rad is None:
            pw = _double(s, "POWER")
            rad = None if pw is None else pw / (4.0 * 3.1415927 * 3.1415927 * r * r)
        sphere_sources.append(dict(center=c, radius=r, color=col, radiance=rad))
        # LightSource::addSSource appends the source's visible sphere
        # (lightSources.h:176-182): colour 0, reflectance 0, VOLUME emitter
        m = max(col)
        emit = tuple(ch / m * rad for ch in col)
        spheres.append(dict(center=c, radius=r, color=(0.0, 0.0, 0.0),
                            material=scenes._material(scenes.BRDF_NONE, int_ior=0.0, ext_ior=0.0,
                                                      reflectance=0.0, color=(0, 0, 0),
                                                      light=scenes.LIGHT_VOLUME, radiance=emit)))
    t64 = np.array([[c for v in t for c in v] for t in tris], np.float64)
    n_tri_mats = len(mats)
    sph = np.zeros(len(spheres), scenes.SPHERE_DTYPE)
    for i, q in enumerate(spheres):
        sph[i]["center"] = q["center"]
        sph[i]["radius"] = q["radius"]
        sph[i]["mat"] = len(mats)
        mats.append(q["material"])
    n_src = len(sphere_sources)
    pts = np.zeros(len(points), scenes.POINT_DTYPE)
    for i, (c, w) in enumerate(points):
        pts[i]["pos"] = c
        pts[i]["power"] = w
    mats_arr = np.zeros(len(mats), scenes.MATERIAL_DTYPE)
    for i, m in enumerate(mats):
        mats_arr[i] = m
    camera = scenes.Camera(tuple(pos), tuple(look), tuple(up), float(fovx), float(fovy), int(W),
                           int(H))
    meta = {"tris64": t64, "sphere_sources": sphere_sources, "source": src[:200]}
    return scenes.Scene(t64.astype(np.float32), np.arange(n_tri_mats, dtype=np.uint32), mats_arr,
                        np.array(lights, scenes.LIGHT_DTYPE), camera, "xml", meta, spheres=sph,
                        tri_lights=np.array(tri_lights, np.uint32),
                        sphere_sources=np.arange(len(spheres) - n_src, len(spheres),
                                                 dtype=np.uint32),
                        point_sources=pts)
