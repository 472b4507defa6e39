"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute: there is no GPU here)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from chakumrender_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(chk_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_built():
    assert abi.LIB_PATH.exists(), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    L = C.CDLL(str(abi.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signature table covers all of them
    assert set(names) <= set(abi.SIGNATURES), set(names) - set(abi.SIGNATURES)


def test_status_strings_without_device():
    L = abi.load()
    assert L.chk_status_string(0) == b"ok"
    assert L.chk_status_string(4) == b"capacity exceeded"
    assert b"gfx950" in L.chk_version()


def test_null_arguments_rejected_without_device():
    L = abi.load()
    assert L.chk_create(0, None) == 1
    assert L.chk_scene_get_info(None, None) == 1
    assert L.chk_ltc_eval(None, None, None, None, None, None, 0, None, None, 0, None, None) == 1
    assert L.chk_brdf_eval(None, 0, None, None, None, None, None, 5, 0, None, None, None) == 1
    assert b"null" in L.chk_last_error()
    assert L.chk_direct_light(None, None, None, None, None, 4, None, None, None) == 1
    assert L.chk_mis_weights(None, None, None, None, None, None, 4, None, 80, None, None) == 1
    assert L.chk_render_direct(None, None, None, None, None, None, None, None) == 1
    assert b"null" in L.chk_last_error()


def test_struct_layouts_match_header():
    txt = (ROOT / "include" / "chakum_types.h").read_text()
    assert "56 bytes" in txt
    assert C.sizeof(abi.chk_material) == 56
    assert C.sizeof(abi.chk_camera) == 4 * 11 + 8
    assert C.sizeof(abi.chk_render_params) == 48
    assert C.sizeof(abi.chk_direct_params) == 24
    assert C.sizeof(abi.chk_sphere) == 40 and C.sizeof(abi.chk_point_source) == 48
    assert C.sizeof(abi.chk_env_source) == 48


LAYOUT_STRUCTS = ["chk_material", "chk_mesh_light", "chk_sphere", "chk_point_source",
                  "chk_env_source", "chk_scene_desc", "chk_camera", "chk_render_params",
                  "chk_direct_params", "chk_scene_info", "chk_render_stats", "chk_request",
                  "chk_batch_tuning", "chk_batch_stats", "chk_train_params", "chk_train_record"]


def test_ctypes_layouts_equal_the_c_compilers(tmp_path):
    """sizeof and every field offset of the ctypes records equal what gcc
    makes of include/*.h."""
    import subprocess

    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "chakum_c.h"', "int main(void) {"]
    for name in LAYOUT_STRUCTS:
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in getattr(abi, name)._fields_:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for name in LAYOUT_STRUCTS:
        T = getattr(abi, name)
        assert int(got[name]) == C.sizeof(T), name
        for f, _ in T._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(T, f).offset, (name, f)


@pytest.mark.parametrize("flag,val", [("CHK_RENDER_PHYSICAL_EMITTER", abi.RENDER_PHYSICAL_EMITTER),
                                      ("CHK_RENDER_SUMSQ", abi.RENDER_SUMSQ),
                                      ("CHK_RENDER_PROFILE", abi.RENDER_PROFILE),
                                      ("CHK_RENDER_ONE_STREAM", abi.RENDER_ONE_STREAM),
                                      ("CHK_DIRECT_VISIBILITY", abi.DIRECT_VISIBILITY),
                                      ("CHK_MIS_LIGHT_SAMPLES", abi.MIS_LIGHT_SAMPLES),
                                      ("CHK_MIS_BSDF_SAMPLES", abi.MIS_BSDF_SAMPLES)])
def test_flag_values_match_header(flag, val):
    txt = "".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    m = re.search(rf"#define {flag}\s+(0x[0-9a-fA-F]+|\d+)u?", txt)
    assert m and int(m.group(1), 0) == val
