"""Convert the reference's regression outputs (PRECOMP/REGexample*.txt) into a JSON fixture.

Reads the reference's own golden files as data (rhosig mu nu lambda col1 col2 per line, printed
with %lg, i.e. 6 significant digits; amu_forlattice.c:47-61) and records next to each column the
call that produced it (amu_forlattice.c:121-379).  Run in a container that has the reference:
    python tests/golden/make_reg_fixtures.py /root/reference/PRECOMP
"""
import json
import math
import os
import sys


def read_cols(path):
    blocks, cur = [], []
    for line in open(path):
        if line.startswith("rhosig="):
            f = line.split()
            cur.append((float(f[8]), float(f[9])))
            if len(cur) == 384:
                blocks.append(cur)
                cur = []
    return blocks


def main(src):
    b, co, si = 0.209243, math.cos(0.209243), math.sin(0.209243)
    x2 = [0.0, 0.0, 1.228 * si, 1.228 * co]
    y2 = [0.0, 0.0, 0.0, 0.0484223]
    z = [0.0, 0.0, 0.0, 0.0]
    fac = 0.046241301098278
    cases = {
        2: [(("L0", x2, y2), ("L0", z, y2)), (("L0", y2, x2), ("L0", x2, z))],
        4: [(("L0", [1, 1, 1, 1.1], [1, 1, 1, 1]), ("L0", [1, 1, 1, 1], [1, 1, 1, 1.1]))],
        5: [(("L0", [0.001] * 4, [1, 1, 1, 1]), ("L0", [1, 1, 1, 1], [0.001] * 4))],
        6: [(("L0", [fac, 3 * fac, 3 * fac, 52 * fac], [fac] * 4), ("L0", [fac] * 4, [fac, 3 * fac, 3 * fac, 52 * fac]))],
        7: [(("L0", [0.001] * 4, [0.001, 0.001, 0.001, -0.001]), ("L0", [0.001, 0.001, 0.001, -0.001], [0.001] * 4))],
        10: [(("L0", z, [0.1] * 4), ("L0", [0.1] * 4, z))],
    }
    out = []
    for ex, blocks in cases.items():
        vals = read_cols(os.path.join(src, f"REGexample{ex}.txt"))
        assert len(vals) == len(blocks), (ex, len(vals))
        for bi, (c1, c2) in enumerate(blocks):
            for ci, call in enumerate((c1, c2)):
                name, x, y = call
                out.append({"example": ex, "block": bi, "column": ci + 1, "kernel": name,
                            "x": [float(v) for v in x], "y": [float(v) for v in y],
                            "expected": [v[ci] for v in vals[bi]]})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reg_examples.json")
    json.dump({"source": "reference PRECOMP/REGexample{2,4,5,6,7,10}.txt, %lg (6 significant digits)",
               "cases": out}, open(dst, "w"), indent=0)
    print("wrote", dst, len(out), "columns")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/PRECOMP")
