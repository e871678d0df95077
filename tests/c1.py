"""The CPU reference configuration (C1): one 4 MiB file of seeded uniform
bytes, 64 KiB slices, 5% redundancy -> 64 input slices and 4 recovery slices
with exponential volumes; then 3 seeded slices zeroed and repaired.

Bytes come from splitmix64 (a fixed published generator) so the committed
digests do not depend on a library's random stream."""
from __future__ import annotations

import numpy as np

C1_NAME = "c1_input.bin"
C1_STEM = "c1_input"
C1_BYTES = 4 << 20
C1_BLOCK = 65536
C1_REDUNDANCY = 5.0
C1_SEED = 0x50415232  # "PAR2"
C1_ERASE_SEED = 0x52455041  # "REPA"


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count outputs of splitmix64 started at `seed` (Vigna's constants)."""
    i = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def seeded_bytes(seed: int, nbytes: int) -> np.ndarray:
    return splitmix64(seed, (nbytes + 7) // 8).view(np.uint8)[:nbytes].copy()


def c1_input() -> bytes:
    return seeded_bytes(C1_SEED, C1_BYTES).tobytes()


def c1_erasures(n_slices: int = C1_BYTES // C1_BLOCK, k: int = 3) -> list[int]:
    """k distinct slice indices from a separate splitmix64 stream."""
    out = []
    for v in splitmix64(C1_ERASE_SEED, 64):
        p = int(v % np.uint64(n_slices))
        if p not in out:
            out.append(p)
        if len(out) == k:
            break
    return sorted(out)
