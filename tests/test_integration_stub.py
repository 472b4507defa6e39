"""The C++ binding a maintainer adds to the reference (INTEGRATION.md's cpp
blocks) compiles against include/chakum_c.h and the reference's own headers
(geometryPrimitives.h, objects.h for Triangle, Sphere, Vec and the scene
lists).  The helpers the stub names but leaves to the maintainer
(to_chk_material, mesh_light_ranges, ...) are declared, not defined: this is
a syntax and type check of the binding, no program is built or run."""
import os
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path(os.environ.get("CHK_REFERENCE", "/root/reference"))

PRELUDE = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <stdint.h>
#include "mathPrimitives.h"
#include "geometryPrimitives.h"
#include "objects.h"
#include "chakum_c.h"
// the maintainer's helpers the binding names
chk_material to_chk_material(const Triangle &t);
chk_material to_chk_material(const Sphere &s);
std::vector<chk_mesh_light> mesh_light_ranges();
std::vector<uint32_t> sphere_source_indices();
std::vector<uint32_t> tri_light_indices();
extern const double *ltc_table;
struct Pending { chk_scene_t scene; chk_camera camera; chk_render_params params; double *accum_dev; };
extern std::vector<Pending> pending;
extern uint32_t nSpheres, n_mesh_lights, total_spp, pass_spp;
"""


def stub_source():
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    assert len(blocks) >= 3, "INTEGRATION.md lost its cpp blocks"
    first = blocks[0]
    marker = "// ... after prepareLightSources() and loadAccels():"
    assert marker in first
    head, body = first.split(marker, 1)
    # the later blocks are separate snippets in the same function: a scope each
    rest = "\n".join("{\n" + b + "}\n" for b in blocks[1:])
    return (PRELUDE + head +
            "\nvoid render_through_libchakum(Vec pos, Vec dir, Vec up, float fovx, float fovy,"
            " uint32_t w, uint32_t h, Vec *pixelColors) {\n" + body + rest + "\n}\n")


@pytest.mark.skipif(not (REF / "headers" / "objects.h").exists(), reason="reference headers absent")
def test_integration_stub_compiles(tmp_path):
    src = tmp_path / "stub.cpp"
    src.write_text(stub_source())
    hip_inc = Path("/opt/rocm/include")
    # the reference's headers are C++98 (in-class initialisers of static
    # const doubles): -fpermissive accepts them in C++11, which HIP needs
    cmd = ["g++", "-std=gnu++11", "-fpermissive", "-fsyntax-only", "-w", "-D__HIP_PLATFORM_AMD__",
           "-I", str(ROOT / "include"), "-I", str(REF / "headers"),
           "-I", str(REF / "externLib" / "objectLoader"), "-I", str(REF / "externLib" / "XMLParser"),
           "-I", str(hip_inc), str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
