"""Extract the Sobol direction matrices and the Sobol pixel-enumeration tables
from the reference's data tables into small binary fixtures.

The reference stores the Joe-Kuo (2008) direction numbers as a 1024x52 u32 table
(Gyunity/src/sampler/Sobol.cpp:38-10689, MIT licence, L. Gruenschloss) and the
per-resolution enumeration tables `vdc_sobol_matrices` / `vdc_sobol_matrices_inv`
(Gyunity/src/sampler/SobolEnum.cpp:5-1465).  Those tables are data: this script
reads the hexadecimal literals as text and writes them as little-endian binary
files.  Nothing from the reference is compiled, imported or executed.

Outputs (all little-endian):
  sobol_matrices.u32     NDIM x 52 uint32  (first NDIM dimensions)
  sobol_enum_vdc.u64     25 x 52 uint64    (vdc_sobol_matrices, m = 1..25)
  sobol_enum_vdc_inv.u64 26 x 52 uint64    (vdc_sobol_matrices_inv, m = 1..26)

Run:  python tests/golden/make_sobol_fixtures.py [/root/reference]
"""
import os
import re
import sys

import numpy as np

NDIM = 256
HERE = os.path.dirname(os.path.abspath(__file__))


def _hex_literals(text):
    return [int(h, 16) for h in re.findall(r"0x([0-9a-fA-F]+)U?L?L?", text)]


def _blocks(text, name):
    """Return the list of brace blocks (one per m) of a 2-D C array `name`."""
    start = text.index(name)
    body = text[text.index("{", start) + 1:]
    depth, cur, out = 0, [], []
    for ch in body:
        if ch == "{":
            depth += 1
            cur = []
            continue
        if ch == "}":
            if depth == 0:
                break
            depth -= 1
            out.append("".join(cur))
            continue
        if depth == 1:
            cur.append(ch)
    return out


def main(ref_root):
    src = os.path.join(ref_root, "Gyunity", "src", "sampler")
    with open(os.path.join(src, "Sobol.cpp")) as f:
        sobol_txt = f.read()
    tab = sobol_txt[sobol_txt.index("Matrices::matrices[") :]
    vals = _hex_literals(tab)
    assert len(vals) == 1024 * 52, len(vals)
    mats = np.array(vals, dtype=np.uint64).reshape(1024, 52)
    assert mats.max() < 2**32
    mats = mats.astype(np.uint32)
    mats[:NDIM].astype("<u4").tofile(os.path.join(HERE, "sobol_matrices.u32"))

    with open(os.path.join(src, "SobolEnum.cpp")) as f:
        enum_txt = f.read()
    vdc = _blocks(enum_txt, "vdc_sobol_matrices[][52]")
    inv = _blocks(enum_txt, "vdc_sobol_matrices_inv[][52]")
    assert len(vdc) == 25 and len(inv) == 26, (len(vdc), len(inv))
    vdc_arr = np.zeros((25, 52), dtype=np.uint64)
    inv_arr = np.zeros((26, 52), dtype=np.uint64)
    for i, b in enumerate(vdc):
        v = _hex_literals(b)
        vdc_arr[i, : len(v)] = v
    for i, b in enumerate(inv):
        v = _hex_literals(b)
        inv_arr[i, : len(v)] = v
    vdc_arr.astype("<u8").tofile(os.path.join(HERE, "sobol_enum_vdc.u64"))
    inv_arr.astype("<u8").tofile(os.path.join(HERE, "sobol_enum_vdc_inv.u64"))
    print("wrote", NDIM, "Sobol dims, 25 vdc rows, 26 inverse rows")
    return mats


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
