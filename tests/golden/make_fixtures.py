"""Writes the golden fixtures the oracle and the GPU path are checked against.

Run in a container that has the reference source at /root/reference:

    python tests/golden/make_fixtures.py

It reads the reference's literal tables as text (numbers only, nothing is
executed) and stores them as numpy arrays, plus the known-answer vectors of
the reference's own tests as JSON.  Outputs:

  tests/golden/galois_tables.npz   galois.go:29 logTable, :71 expTable,
                                   :78 invTable, :80 mulTable,
                                   :337 mulTableLow, :593 mulTableHigh,
                                   :948 gf2p811dMulMatrices
  tests/golden/leo8_gfni.npy       leopard8.go:1271 gf2p811dMulMatricesLeo8
  tests/golden/kat.json            known-answer vectors (sources cited inside)
  tests/golden/leopard_consts.json leopard.go:100-105 bitwidth/order/modulus/
                                   polynomial and :1105-1110 cantorBasis;
                                   leopard8.go:96-104 and :1030-1032 likewise
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _block(src: str, name: str) -> str:
    """Text of the composite literal `var name = [...]T{ ... }`."""
    start = src.index(f"var {name} = ")
    brace = src.index("{", start)
    depth = 0
    for i in range(brace, len(src)):
        if src[i] == "{":
            depth += 1
        elif src[i] == "}":
            depth -= 1
            if depth == 0:
                return src[brace:i + 1]
    raise ValueError(name)


def _local_block(src: str, decl: str) -> str:
    """Text of the composite literal `decl ... { ... }` inside a function body."""
    start = src.index(decl)
    brace = src.index("{", start)
    return src[brace:src.index("}", brace) + 1]


def _const(src: str, name: str) -> str:
    m = re.search(r"^\s*" + re.escape(name) + r"\s*=\s*(.+?)\s*$", src, flags=re.M)
    if not m:
        raise ValueError(name)
    return m.group(1)


def leopard_consts(leo16: str, leo8: str) -> dict:
    """Field constants of the two Leopard engines, read from the source text."""
    out = {"_source": "leopard.go:100-105,1105-1110; leopard8.go:96-104,1030-1032"}
    for tag, src, sfx, basis_decl in (("gf16", leo16, "", "cantorBasis := [bitwidth]ffe{"),
                                      ("gf8", leo8, "8", "cantorBasis := [bitwidth8]ffe8{")):
        bitwidth = int(_const(src, "bitwidth" + sfx), 0)
        assert _const(src, "order" + sfx) == f"1 << bitwidth{sfx}"
        assert _const(src, "modulus" + sfx) == f"order{sfx} - 1"
        basis = _numbers(_local_block(src, basis_decl))
        assert len(basis) == bitwidth, (tag, basis)
        out[tag] = {"bitwidth": bitwidth, "order": 1 << bitwidth, "modulus": (1 << bitwidth) - 1,
                    "polynomial": int(_const(src, "polynomial" + sfx), 0), "cantor_basis": basis}
    return out


def _numbers(text: str) -> list[int]:
    return [int(t, 0) for t in re.findall(r"0x[0-9a-fA-F]+|\d+", text)]


def main() -> int:
    if not os.path.isdir(REF):
        print("reference source not present; fixtures are committed", file=sys.stderr)
        return 1
    galois = open(os.path.join(REF, "galois.go")).read()
    leo8 = open(os.path.join(REF, "leopard8.go")).read()
    leo16 = open(os.path.join(REF, "leopard.go")).read()
    with open(os.path.join(HERE, "leopard_consts.json"), "w") as f:
        json.dump(leopard_consts(leo16, leo8), f, indent=1)

    def arr(name, n, dtype=np.uint8):
        vals = _numbers(_block(galois, name))
        assert len(vals) == n, (name, len(vals))
        return np.array(vals, dtype=dtype)

    tables = dict(
        log=arr("logTable", 256),
        exp=arr("expTable", 256),
        inv=arr("invTable", 256),
        mul=arr("mulTable", 65536).reshape(256, 256),
        low=arr("mulTableLow", 4096).reshape(256, 16),
        high=arr("mulTableHigh", 4096).reshape(256, 16),
        gfni=np.array([int(t, 0) for t in re.findall(r"0x[0-9a-fA-F]+|\b\d+\b",
                                                     _block(galois, "gf2p811dMulMatrices"))],
                      dtype=np.uint64),
    )
    assert tables["gfni"].shape == (256,)
    np.savez_compressed(os.path.join(HERE, "galois_tables.npz"), **tables)

    leo = [int(t, 16) for t in re.findall(r"0x([0-9a-fA-F]+)", _block(leo8, "gf2p811dMulMatricesLeo8"))]
    assert len(leo) == 256
    np.save(os.path.join(HERE, "leo8_gfni.npy"), np.array(leo, dtype=np.uint64))

    in34 = [0, 1, 2, 3, 4, 5, 6, 10, 50, 100, 150, 174, 201, 255, 99, 32, 67, 85, 200, 199, 198,
            197, 196, 195, 194, 193, 192, 191, 190, 189, 188, 187, 186, 185]
    kat = {
        "_source": "reference test files; each entry names file:line",
        "gal_multiply": {"src": "galois_test.go:128-136",
                         "cases": [[3, 4, 12], [7, 7, 21], [23, 45, 41]]},
        "gal_exp": {"src": "galois_test.go:164-172",
                    "cases": [[2, 2, 4], [5, 20, 235], [13, 7, 43]]},
        "gal_mul_slice": {
            "src": "galois_test.go:139-162",
            "input": in34,
            "steps": [
                {"c": 25, "xor": False,
                 "expect": [0x0, 0x19, 0x32, 0x2b, 0x64, 0x7d, 0x56, 0xfa, 0xb8, 0x6d, 0xc7, 0x85,
                            0xc3, 0x1f, 0x22, 0x7, 0x25, 0xfe, 0xda, 0x5d, 0x44, 0x6f, 0x76, 0x39,
                            0x20, 0xb, 0x12, 0x11, 0x8, 0x23, 0x3a, 0x75, 0x6c, 0x47]},
                {"c": 52, "xor": True,
                 "expect": [0x0, 0x2d, 0x5a, 0x77, 0xb4, 0x99, 0xee, 0x2f, 0x79, 0xf2, 0x7, 0x51,
                            0xd4, 0x19, 0x31, 0xc9, 0xf8, 0xfc, 0xf9, 0x4f, 0x62, 0x15, 0x38, 0xfb,
                            0xd6, 0xa1, 0x8c, 0x96, 0xbb, 0xcc, 0xe1, 0x22, 0xf, 0x78]},
                {"c": 177, "xor": False,
                 "expect": [0x0, 0xb1, 0x7f, 0xce, 0xfe, 0x4f, 0x81, 0x9e, 0x3, 0x6, 0xe8, 0x75,
                            0xbd, 0x40, 0x36, 0xa3, 0x95, 0xcb, 0xc, 0xdd, 0x6c, 0xa2, 0x13, 0x23,
                            0x92, 0x5c, 0xed, 0x1b, 0xaa, 0x64, 0xd5, 0xe5, 0x54, 0x9a]},
                {"c": 117, "xor": True,
                 "expect": [0x0, 0xc4, 0x95, 0x51, 0x37, 0xf3, 0xa2, 0xfb, 0xec, 0xc5, 0xd0, 0xc7,
                            0x53, 0x88, 0xa3, 0xa5, 0x6, 0x78, 0x97, 0x9f, 0x5b, 0xa, 0xce, 0xa8,
                            0x6c, 0x3d, 0xf9, 0xdf, 0x1b, 0x4a, 0x8e, 0xe8, 0x2c, 0x7d]},
            ],
        },
        "matrix_multiply": {"src": "matrix_test.go:90-114",
                            "a": [[1, 2], [3, 4]], "b": [[5, 6], [7, 8]],
                            "expect": [[11, 22], [19, 42]]},
        "matrix_inverse": {
            "src": "matrix_test.go:117-217",
            "cases": [
                {"m": [[56, 23, 98], [3, 100, 200], [45, 201, 123]],
                 "expect": [[175, 133, 33], [130, 13, 245], [112, 35, 126]]},
                {"m": [[1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0, 0, 0, 1, 0], [0, 0, 0, 0, 1],
                       [7, 7, 6, 6, 1]],
                 "expect": [[1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [123, 123, 1, 122, 122],
                            [0, 0, 1, 0, 0], [0, 0, 0, 1, 0]]},
                {"m": [[56, 23], [3, 100], [45, 201]], "error": "not_square"},
                {"m": [[4, 2], [12, 6]], "error": "singular"},
            ],
        },
        "jerasure_8_4": {"src": "reedsolomon_test.go:129-156",
                         "parity_rows": [[1, 1, 1, 1, 1, 1, 1, 1],
                                         [1, 55, 39, 73, 84, 181, 225, 217],
                                         [1, 39, 217, 161, 92, 60, 172, 90],
                                         [1, 172, 70, 235, 143, 34, 200, 101]]},
        "one_encode_5_5": {"src": "reedsolomon_test.go:1075-1125",
                           "data": [[0, 1], [4, 5], [2, 3], [6, 7], [8, 9]],
                           "parity": [[12, 13], [10, 11], [14, 15], [90, 91], [94, 95]]},
        "leo16_add_mod": {"src": "leopard_test.go:167-190",
                          "cases": [[1, 2, 3], [65533, 1, 65534], [65534, 2, 1], [65535, 1, 1],
                                    [65534, 1, 65535], [65535, 0, 65535], [65535, 65535, 65535]]},
        "leo8_add_mod": {"src": "leopard8_test.go:10-33",
                         "cases": [[1, 2, 3], [253, 1, 254], [254, 2, 1], [255, 1, 1],
                                   [254, 1, 255], [255, 0, 255], [255, 255, 255]]},
        "new_errors": {
            "src": "reedsolomon_test.go:2070-2095 (TestNew; the int-overflow row is not representable in C)",
            "cases": [
                {"k": 127, "m": 127, "err": None},
                {"k": 128, "m": 128, "err": None},
                {"k": 255, "m": 1, "err": None},
                {"k": 255, "m": 0, "err": None},
                {"k": 1, "m": 0, "err": None},
                {"k": 65536, "m": 65536, "err": "ErrMaxShardNum"},
                {"k": 0, "m": 1, "err": "ErrInvShardNum"},
                {"k": 1, "m": -1, "err": "ErrInvShardNum"},
                {"k": 65636, "m": 1, "err": "ErrMaxShardNum"},
            ],
        },
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print("fixtures written to", HERE)
    return 0


if __name__ == "__main__":
    sys.exit(main())
