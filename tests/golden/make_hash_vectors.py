"""Writes hash_vectors.json: MD5 (hashlib) and CRC-32 (zlib) of seeded byte
strings at edge-case lengths, plus IFSC entries of zero-padded slices.

hashlib and zlib are independent, standard implementations of RFC 1321 and
CRC-32/IEEE; these vectors pin the oracle and the GPU hash kernels.
Run: python tests/golden/make_hash_vectors.py
"""
import hashlib
import json
import os
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LENGTHS = [0, 1, 3, 55, 56, 57, 63, 64, 65, 119, 120, 127, 128, 129, 1000, 4096, 5376, 65535, 65536]


def main():
    rng = np.random.default_rng(20261203)
    vec = []
    for n in LENGTHS:
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        vec.append({"seed_len": n, "hex": b.hex() if n <= 129 else None,
                    "md5": hashlib.md5(b).hexdigest(), "crc32": zlib.crc32(b) & 0xFFFFFFFF})
    # the long ones are regenerated from the generator state below
    rng2 = np.random.default_rng(7)
    slices = []
    slice_size = 4096
    for ln in [4096, 4095, 1, 0, 2048, 4094, 17]:
        b = rng2.integers(0, 256, ln, dtype=np.uint8).tobytes()
        padded = b + bytes(slice_size - ln)
        slices.append({"len": ln, "md5": hashlib.md5(padded).hexdigest(),
                       "crc32": zlib.crc32(padded) & 0xFFFFFFFF})
    out = {"note": "bytes for lengths > 129 come from numpy default_rng(20261203) in LENGTHS order; "
                   "slices from default_rng(7) in order, zero-padded to slice_size",
           "vectors": vec, "slice_size": slice_size, "slices": slices}
    with open(os.path.join(HERE, "hash_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
