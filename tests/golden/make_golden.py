"""Writes tests/golden/hilbert.json from the reference's own hilbert.h.

The reference header is compiled unmodified by `make -C oracle ref` into
oracle/_ref/libhilbert_ref.so (this container only: the reference tree does
not travel).  The fixture holds inputs and outputs only:
  - edge cases, including the 17-bit value 65536 that the reference's
    shift-and-OR packing spills into the next axis;
  - 4096 seeded random coordinate triples in [0, 65536];
  - 512 seeded random float particles quantised inside a fixed AABB;
  - checksums (sum mod 2^64 and xor) of c2i over the exhaustive blocks
    [0,64)^3 and [65472,65536]^3;
  - 1024 seeded random indices and their i2c coordinates.
Run: python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def main():
    oracle.build(ref=True)
    R = oracle.ref_lib()
    A = C.c_uint64 * 3

    def c2i(x, y, z):
        return int(R.ref_hilbert_c2i(A(x, y, z)))

    edge = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (65535, 65535, 65535), (65536, 0, 0), (0, 65536, 0),
            (0, 0, 65536), (65536, 65536, 65536), (65536, 65535, 0), (32768, 32768, 32768), (12345, 54321, 40000)]
    rng = np.random.default_rng(20240607)
    coords = [tuple(int(v) for v in rng.integers(0, 65537, 3)) for _ in range(4096)]
    out = {"edge": [[*c, c2i(*c)] for c in edge], "random": [[*c, c2i(*c)] for c in coords]}

    aabb = np.array([-1.25, 0.0, 2.0, 3.5, 1.0, 2.75], dtype=np.float32)
    pts = np.empty((512, 4), dtype=np.float32)
    for d in range(3):
        pts[:, d] = rng.uniform(aabb[d], aabb[3 + d], 512).astype(np.float32)
    pts[:, 3] = rng.normal(size=512).astype(np.float32)
    pts[0, :3] = aabb[:3]
    pts[1, :3] = aabb[3:]
    codes = np.empty(512, dtype=np.uint64)
    R.ref_hilbert_codes(pts.ctypes.data_as(C.POINTER(C.c_float)), 512, aabb.ctypes.data_as(C.POINTER(C.c_float)),
                        codes.ctypes.data_as(C.POINTER(C.c_uint64)))
    out["particles"] = {"aabb": aabb.tolist(), "xyzw": pts.tolist(), "codes": [int(c) for c in codes]}

    def block(lo, hi):
        s, x = 0, 0
        for a in range(lo, hi):
            for b in range(lo, hi):
                for c in range(lo, hi):
                    v = c2i(a, b, c)
                    s = (s + v) & 0xFFFFFFFFFFFFFFFF
                    x ^= v
        return [lo, hi, s, x]

    out["blocks"] = [block(0, 64), block(65472, 65537)]
    idx = [int(v) for v in rng.integers(0, 2 ** 48, 1024, dtype=np.uint64)]
    i2c = []
    for i in idx:
        c = A()
        R.ref_hilbert_i2c(C.c_uint64(i), c)
        i2c.append([i, *[int(v) for v in c]])
    out["i2c"] = i2c
    with open(os.path.join(HERE, "hilbert.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "hilbert.json"))


if __name__ == "__main__":
    main()
