"""Writes tests/golden/pcg32.json: draws of the reference's own pcg32
(rainbow-cpu/shared/random_generator.cpp, compiled from the reference tree by
oracle/Makefile into oracle/_ref/libref_pcg.so).  Run in the container that
holds /root/reference:  make -C oracle ref && python tests/golden/make_pcg32_golden.py
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libref_pcg.so")

SEEDS = [0, 1, 2, 7, 255, 4095, 65535, 2**32 + 5, 0x853c49e6748fea9b, 2**63 + 12345]
N = 64


def main():
    lib = C.CDLL(LIB)
    lib.ref_pcg32_reals.argtypes = [C.c_ulonglong, C.c_uint, C.POINTER(C.c_float)]
    lib.ref_pcg32_u32.argtypes = [C.c_ulonglong, C.c_uint, C.POINTER(C.c_uint)]
    out = {"source": "rainbow-cpu/shared/random_generator.cpp pcg32(seed): uniform_uint32(), uniform_real() and uint32(min, max)",
           "n": N, "vectors": []}
    for s in SEEDS:
        u = np.zeros(N, np.uint32)
        r = np.zeros(N, np.float32)
        lib.ref_pcg32_u32(s, N, u.ctypes.data_as(C.POINTER(C.c_uint)))
        lib.ref_pcg32_reals(s, N, r.ctypes.data_as(C.POINTER(C.c_float)))
        out["vectors"].append({"seed": s, "u32": [int(x) for x in u],
                               "real_bits": [int(x) for x in r.view(np.uint32)]})
    # pcg32::uint32(min, max) (random_generator.cpp:73-85), bounds that reject often
    # ((2^32 - v) % v close to 2^31 for v = 2^31 + 1) and the shuffle's small ranges
    lo = np.array([0, 0, 5, 0, 0, 0, 3, 0] * 8, np.uint32)
    hi = np.array([2**31, 15, 5, 0, 2**32 - 2, 6, 2**31 + 2, 255] * 8, np.uint32)
    lib.ref_pcg32_range.argtypes = [C.c_ulonglong, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                    C.POINTER(C.c_uint)]
    out["range_lo"] = [int(x) for x in lo]
    out["range_hi"] = [int(x) for x in hi]
    for v in out["vectors"]:
        r = np.zeros(len(lo), np.uint32)
        lib.ref_pcg32_range(v["seed"], len(lo), lo.ctypes.data_as(C.POINTER(C.c_uint)),
                            hi.ctypes.data_as(C.POINTER(C.c_uint)), r.ctypes.data_as(C.POINTER(C.c_uint)))
        v["range"] = [int(x) for x in r]
    with open(os.path.join(HERE, "pcg32.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
