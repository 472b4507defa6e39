"""The quantized BVH nodes (chk_bvh.h Node4Q) and the traversal's ray / node
test (chk_node_test.h node_hits_q), on the CPU: every quantized child box
contains its exact box, and every child box an exact slab test accepts
(quad precision, on the exact box) the kernels' test accepts too, with an
entry distance no later than the exact one up to FP32 rounding
(tests/native/bvh_quant_check.cpp).  The test is conservative by
construction; BvhAccel::intersectT's boxes (bvhAccel.cpp:200-249) only
prune, so a conservative test keeps the closest hit unchanged.  The grid
cases (boxes exactly on their grid, rays through corners, edges and faces,
some from the grid origin) fail if the outward move of b_a or the kGrow
factor is left out."""
import os
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CSRC = ROOT / "chakumrender_amd" / "csrc"


def test_quantized_nodes_are_conservative(tmp_path):
    bvh_o = tmp_path / "chk_bvh.o"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", f"-I{CSRC}", f"-I{ROOT / 'include'}",
                    "-c", str(CSRC / "chk_bvh.cpp"), "-o", str(bvh_o)], check=True)
    exe = tmp_path / "bvh_quant_check"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2", f"-I{CSRC}",
                    f"-I{ROOT / 'include'}", str(ROOT / "tests" / "native" / "bvh_quant_check.cpp"),
                    "-x", "none", str(bvh_o), "-lgomp", "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe), "600"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    m = re.search(r"grid_accepts (\d+) boxes (\d+) contain_bad (\d+) id_bad (\d+) tests (\d+) "
                  r"accepts (\d+) missed (\d+) key_late (\d+)", out.stdout)
    assert m, out.stdout + out.stderr
    grid, boxes, contain_bad, id_bad, tests, accepts, missed, late = map(int, m.groups())
    assert boxes > 50000 and tests > 1000000
    assert grid > 50000 and accepts > 100000  # the boundary cases are exercised
    assert contain_bad == 0 and id_bad == 0
    assert missed == 0 and late == 0
