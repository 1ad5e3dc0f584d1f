"""Write tests/golden/golden.json from the reference's own code where it builds.

Run in the build container (needs oracle/_ref, which oracle/Makefile compiles
from the reference's src/fast_prng.cuh and src/bit_iterator.cuh):

    make -C oracle && python tests/golden/make_golden.py

Every vector here is produced by the reference's PRNG / bit_iterator (through
oracle/_ref) plus numpy closed forms -- never by oracle.c -- so the oracle is
checked against it in tests/test_oracle.py.
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def ref_prng(R, seed, n, skip=0):
    out = np.empty(n, dtype=np.uint32)
    R.amref_prng_fill(seed, skip, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
    return out


def slab_pack_closed_form(vals, bits):
    eps = 64 // bits
    words = np.zeros((len(vals) + eps - 1) // eps, dtype=np.uint64)
    for s in range(eps):
        part = vals[s::eps].astype(np.uint64)
        words[:len(part)] |= part << np.uint64(s * bits)
    return words


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libamref.so not built (needs the reference tree)")
    g = {"source": "reference src/fast_prng.cuh + src/bit_iterator.cuh via oracle/_ref; numpy closed forms"}
    g["prng"] = {str(s): ref_prng(R, s, 64).tolist() for s in (42, 0, 1, 123456789)}
    g["prng_skip"] = {"seed": 42, "skip": 1000, "draws": ref_prng(R, 42, 16, 1000).tolist()}

    # bit_iterator reads of a fixed byte pattern
    pattern = np.array([0x01, 0x80, 0xA5, 0x3C, 0xFF, 0x00, 0x5A, 0xC3], dtype=np.uint8)
    oi = np.zeros(64, np.uint8)
    ow = np.zeros(64, np.uint8)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    R.amref_bit_iterator_read(pattern.ctypes.data_as(u8p), 64, oi.ctypes.data_as(u8p), ow.ctypes.data_as(u8p))
    g["bit_iterator"] = {"bytes": pattern.tolist(), "index": oi.tolist(), "walk": ow.tolist()}

    # generate_mask_uniform (src/main.cu:33-46): draws of fast_prng(42), bit j
    # of each byte from 7 down to 0, set iff draw < (uint64)(s * UINT32_MAX)
    nbytes, sel = 32, 0.3
    d = ref_prng(R, 42, nbytes * 8).astype(np.uint64)
    thr = int(sel * 4294967295.0)
    bits = (d < np.uint64(thr)).reshape(nbytes, 8)  # draw order j = 7..0
    g["mask_uniform_ref"] = {"selectivity": sel, "bytes": np.packbits(bits, axis=1, bitorder="big").ravel().tolist()}

    # benchmark-1 path in miniature: values = rand() & 0x7FF (seed 42),
    # slab at 11 bits, predicate v < 1024, uint32 output + LSB-first mask
    n = 300
    vals = ref_prng(R, 42, n).astype(np.uint64) & np.uint64(0x7FF)
    sel = vals < 1024
    mask = np.packbits(sel, bitorder="little")
    mask = np.concatenate([mask, np.zeros((-len(mask)) % 8, np.uint8)]).view(np.uint64)
    g["c1_small"] = {
        "n": n, "bits": 11, "pred": ["LT", 1024],
        "slab_words": [str(int(x)) for x in slab_pack_closed_form(vals, 11)],
        "mask_words": [str(int(x)) for x in mask],
        "out_u32": vals[sel].astype(np.uint32).tolist(),
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
