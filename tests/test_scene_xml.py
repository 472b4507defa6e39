"""chakumrender_amd.scene_xml against the reference's own front end
(SceneParser + prepareScreenCamera / prepareObjects / prepareLightSources,
built from its sources with the vendored tinyxml2) on its two scene files
(tests/golden/scene_xml.npz holds the files and what the reference built).

Bars: FP64 vertices equal to 1e-12 relative (same operations; the quaternion
rotations go through libm sin/cos); materials and light radiance exactly as
the C ABI's FP32 record holds them; mesh-light sizes, spheres, sphere
sources and camera basis equal."""
import numpy as np
import pytest

from chakumrender_amd import abi, scene_xml, scenes


@pytest.fixture(scope="module")
def G(golden):
    return golden("scene_xml.npz")


def _mat_rows(sc):
    m = sc.mats[sc.tri_mat]
    return np.c_[m["brdf"], m["light"], m["ks"], m["alpha"], m["phong_exp"], m["int_ior"],
                 m["ext_ior"], m["reflectance"], m["color"], m["radiance"]].astype(np.float32)


@pytest.mark.parametrize("name", ["scene0", "scene1"])
def test_triangles_materials_and_lights_equal_reference(G, name):
    sc = scene_xml.load_scene_xml(str(G[f"{name}_xml"]))
    ref_tris = G[f"{name}_tris"]
    assert sc.n_tris == ref_tris.shape[0] == int(G[f"{name}_counts"][0])
    np.testing.assert_allclose(sc.meta["tris64"], ref_tris, rtol=1e-12, atol=1e-9)
    np.testing.assert_array_equal(sc.tris, ref_tris.astype(np.float32))
    ref = G[f"{name}_tmat"][:, :14].astype(np.float32)
    np.testing.assert_array_equal(_mat_rows(sc), ref)
    # mesh lights: the reference's MeshLight sizes, appended after the objects
    sizes = G[f"{name}_mesh"]
    assert [int(c) for c in sc.lights["n_tris"]] == [int(c) for c in sizes]
    if len(sizes):
        assert int(sc.lights["first_tri"][0]) == sc.n_tris - int(sizes.sum())
        emit = sc.mats["light"][sc.tri_mat] != scenes.LIGHT_NONE
        assert emit.sum() == sizes.sum()


@pytest.mark.parametrize("name", ["scene0", "scene1"])
def test_camera_basis_equals_reference(G, name):
    sc = scene_xml.load_scene_xml(str(G[f"{name}_xml"]))
    cam = G[f"{name}_cam"].reshape(4, 3)
    c = sc.camera
    assert (c.width, c.height) == tuple(int(v) for v in G[f"{name}_counts"][6:8])
    pos, d, cx, cy = c.basis()
    np.testing.assert_allclose(pos, cam[0], rtol=0, atol=0)
    np.testing.assert_allclose(d, cam[1], rtol=1e-15)
    np.testing.assert_allclose(cx, cam[2], rtol=1e-13)
    np.testing.assert_allclose(cy, cam[3], rtol=1e-13)


def test_spheres_and_sources(G):
    """scene1.xml: sphere objects, the dummy sphere and the sphere source's
    visible sphere in vSphereList's order with their materials, and the
    sphere source, equal to the reference's parsed scene."""
    sc = scene_xml.load_scene_xml(str(G["scene1_xml"]))
    ref = G["scene1_spheres"]
    assert len(sc.spheres) == ref.shape[0]
    for s, r in zip(sc.spheres, ref):
        np.testing.assert_allclose(list(s["center"]) + [s["radius"]], r[:4], rtol=1e-15)
        m = sc.mats[s["mat"]]
        got = np.r_[m["brdf"], m["light"], m["ks"], m["alpha"], m["phong_exp"], m["int_ior"],
                    m["ext_ior"], m["reflectance"], m["color"]].astype(np.float32)
        np.testing.assert_array_equal(got, r[4:15].astype(np.float32))
    ss = G["scene1_sphere_sources"]
    assert len(sc.sphere_sources) == ss.shape[0] == 1
    q = sc.spheres[sc.sphere_sources[0]]
    np.testing.assert_allclose(list(q["center"]) + [q["radius"]], ss[0, :4], rtol=1e-15)
    np.testing.assert_allclose(sc.mats[q["mat"]]["radiance"], ss[0, 4:7].astype(np.float32),
                               rtol=0)
    assert sc.mats[q["mat"]]["light"] == scenes.LIGHT_VOLUME
    assert len(sc.tri_lights) == len(sc.point_sources) == 0
    s0 = scene_xml.load_scene_xml(str(G["scene0_xml"]))
    assert len(s0.spheres) == 1 and len(s0.sphere_sources) == 0  # the dummy sphere only


def test_point_and_triangle_sources():
    """POINTSOURCE (parsePointSource, sceneParser.cpp:76-90) and TRISOURCE
    (prepareScene.cpp:87-111): power and radiance are col.maxnorm() times
    the value; the triangle light's triangle follows the objects."""
    text = ("<SCREEN><WIDTH>4</WIDTH><HEIGHT>4</HEIGHT></SCREEN><CAMERA><POSITION>0 0 0</POSITION>"
            "<LOOKAT>0 0 -1</LOOKAT><UP>0 1 0</UP><FOVX>60</FOVX><FOVY>60</FOVY></CAMERA>"
            "<POINTSOURCE><POSITION>1 2 3</POSITION><COLOR>0.5 1.0 0.25</COLOR><POWER>8</POWER>"
            "</POINTSOURCE>"
            "<TRISOURCE><TR><V1>0 0 -1</V1><V2>1 0 -1</V2><V3>0 1 -1</V3><COLOR>2 1 1</COLOR>"
            "<RADIANCE>3</RADIANCE></TR></TRISOURCE>")
    sc = scene_xml.load_scene_xml(text)
    np.testing.assert_array_equal(sc.point_sources["pos"][0], [1, 2, 3])
    np.testing.assert_array_equal(sc.point_sources["power"][0], [4.0, 8.0, 2.0])
    assert sc.tri_lights.tolist() == [sc.n_tris - 1] and len(sc.lights) == 0
    m = sc.mats[sc.tri_mat[sc.tri_lights[0]]]
    assert m["light"] == scenes.LIGHT_PLANAR
    np.testing.assert_array_equal(m["radiance"], np.float32([3.0, 1.5, 1.5]))


def test_temperature_colours():
    """tempToColor (mathPrimitives.cpp:32-60): clamps to [1000, 40000] K,
    white near 6600 K, red-dominant below and blue-dominant above."""
    assert scene_xml.temp_to_color(500) == scene_xml.temp_to_color(1000)
    assert scene_xml.temp_to_color(1e6) == scene_xml.temp_to_color(40000)
    r, g, b = scene_xml.temp_to_color(6600)
    assert r == 1.0 and g > 0.99 and b > 0.99
    r, g, b = scene_xml.temp_to_color(1900)
    assert r == 1.0 and g < 0.6 and b == 0.0
    r, g, b = scene_xml.temp_to_color(40000)
    assert b == 1.0 and r < 0.7


def test_parse_errors():
    with pytest.raises(ValueError):
        scene_xml.load_scene_xml("<SCREEN><WIDTH>4</WIDTH><HEIGHT>4</HEIGHT></SCREEN>")
    bad = ("<SCREEN><WIDTH>4</WIDTH><HEIGHT>4</HEIGHT></SCREEN><CAMERA><POSITION>0 0 0</POSITION>"
           "<LOOKAT>0 0 -1</LOOKAT><UP>0 1 0</UP><FOVX>60</FOVX><FOVY>60</FOVY></CAMERA>"
           "<MESHOBJECT><TR><V1>0 0 0</V1><V2>1 0 0</V2><V3>0 1 0</V3><COLOR>1 1 1</COLOR>"
           "</TR></MESHOBJECT>")
    with pytest.raises(ValueError, match="ALPHA"):
        scene_xml.load_scene_xml(bad)


def test_ppm_writer_format(tmp_path):
    """P3 output of main.cpp:98-117: header "P3\\n%d %d\\n%d\\n", row 0 (top)
    first, "%d %d %d " per pixel, toDisplayValue = int(clamp(x)^(1/2.2) * 255
    + .5) (main.cpp:27): 0 -> 0, 1 -> 255, 0.5 -> 186, 0.2 -> 123, and values
    outside [0, 1] clamp."""
    from chakumrender_amd import host

    img = np.array([[[0.0, 1.0, 0.5], [-1.0, 2.0, 0.2], [0.5, 0.5, 0.5]],
                    [[1.0, 0.0, 0.0], [0.2, 0.2, 0.2], [3.0, -3.0, 1.0]]])
    p = tmp_path / "image.ppm"
    host.write_ppm(p, img)
    txt = p.read_text()
    assert txt.startswith("P3\n3 2\n255\n")
    body = txt[len("P3\n3 2\n255\n"):]
    assert body == ("0 255 186 0 255 123 186 186 186 "
                    "255 0 0 123 123 123 255 0 255 ")
