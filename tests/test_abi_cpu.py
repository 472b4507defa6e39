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
    assert C.sizeof(abi.chk_render_params) == 32
    assert C.sizeof(abi.chk_direct_params) == 24


@pytest.mark.parametrize("flag,val", [("CHK_RENDER_PHYSICAL_EMITTER", abi.RENDER_PHYSICAL_EMITTER),
                                      ("CHK_RENDER_SUMSQ", abi.RENDER_SUMSQ),
                                      ("CHK_RENDER_PROFILE", abi.RENDER_PROFILE),
                                      ("CHK_DIRECT_VISIBILITY", abi.DIRECT_VISIBILITY),
                                      ("CHK_MIS_LIGHT_SAMPLES", abi.MIS_LIGHT_SAMPLES),
                                      ("CHK_MIS_BSDF_SAMPLES", abi.MIS_BSDF_SAMPLES)])
def test_flag_values_match_header(flag, val):
    txt = "".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    m = re.search(rf"#define {flag}\s+(0x[0-9a-fA-F]+|\d+)u?", txt)
    assert m and int(m.group(1), 0) == val
