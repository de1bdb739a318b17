"""Extract the NaCl supercell-invariance fixtures held by the reference test
src/nnp/test_nnp_symm.cpp (make_NaCl_small :25-45, make_NaCl_large :47-289)
into tests/golden/nacl.json: lattice constants, lattice matrices and the
atom coordinates (numbers only).  Run in a container where /root/reference
exists: python tests/golden/make_nacl.py
"""
import json
import os
import re

SRC = "/root/reference/src/nnp/test_nnp_symm.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nacl.json")


def main():
    text = open(SRC).read().splitlines()
    small = {"a0": 5.6199998856, "lv_rows": [[0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]],
             "species": [0, 1], "pos_units_a0": []}
    large = {"a0": 16.8599987030, "lv_rows": [[1, 0, 0], [0, 1, 0], [0, 0, 1]], "species": [],
             "pos_units_a0": []}
    pat = re.compile(r"struc\.posn\((\d+)\)<<([-0-9.e]+),([-0-9.e]+),([-0-9.e]+);")
    for ln, line in enumerate(text, start=1):
        m = pat.search(line)
        if not m:
            continue
        xyz = [float(m.group(k)) for k in (2, 3, 4)]
        if 25 <= ln <= 45:
            small["pos_units_a0"].append(xyz)
        elif 47 <= ln <= 289:
            large["pos_units_a0"].append(xyz)
    n = len(large["pos_units_a0"])
    large["species"] = [0] * (n // 2) + [1] * (n - n // 2)
    assert len(small["pos_units_a0"]) == 2 and n == 216, (len(small["pos_units_a0"]), n)
    json.dump({"source": "src/nnp/test_nnp_symm.cpp", "small": small, "large": large}, open(OUT, "w"))
    print(OUT)


if __name__ == "__main__":
    main()
