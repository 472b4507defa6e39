"""Extract the 64x64 GGX LTC fit table into a flat little-endian float64 file.

The renderer's LTC lookup needs the same fitted table the reference ships as a
C initializer (headers/ltc.inc:1-12298: tabM at :3-4100, tabMinv at
:4102-8199, tabAmplitude at :8201-12298).  The table is numeric data (the
Heitz et al. 2016 GGX fit); this script parses the text and writes

    ltc_ggx_64.bin : 4096 records x 19 float64 =
                     M[9] (file order), Minv[9] (file order), amplitude

Record index is a + t*64 (a = sqrt(alpha) bin, t = theta bin), i.e. the same
index the reference's M_GGX uses (headers/ltc.h:18-28).  "File order" is the
order of the nine numbers inside each brace; the reference feeds them to the
column-major 9-argument mat3 constructor (headers/mathPrimitives.h:149-163),
so m[0], m[1], m[2] are the first column.  Amplitudes are float literals in
the reference and are stored here after rounding to float32 and widening, so
the file holds exactly the values the reference computes with.

Run once from a checkout that has the reference mounted:
    python chakumrender_amd/data/make_ltc_table.py /root/reference/headers/ltc.inc
"""
import re
import sys
from pathlib import Path

import numpy as np

NUM = re.compile(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?")


def parse(path):
    txt = Path(path).read_text()
    parts = txt.split("static const")
    tabs = {}
    for p in parts[1:]:
        name = p.split("[")[0].split()[-1]
        if "=" not in p or "{" not in p:
            continue
        body = p.split("=", 1)[1]
        tabs[name] = np.array([float(x) for x in NUM.findall(body)], dtype=np.float64)
    m = tabs["tabM"].reshape(4096, 9)
    minv = tabs["tabMinv"].reshape(4096, 9)
    amp = tabs["tabAmplitude"].astype(np.float32).astype(np.float64).reshape(4096, 1)
    return np.concatenate([m, minv, amp], axis=1)


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/headers/ltc.inc"
    table = parse(src)
    assert table.shape == (4096, 19)
    out = Path(__file__).with_name("ltc_ggx_64.bin")
    table.astype("<f8").tofile(out)
    nz = np.nonzero(np.any(table[:, :9] != 0, axis=0))[0].tolist()
    nzi = np.nonzero(np.any(table[:, 9:18] != 0, axis=0))[0].tolist()
    print(f"wrote {out} ({out.stat().st_size} bytes); nonzero M slots {nz}, Minv slots {nzi}")


if __name__ == "__main__":
    main()
