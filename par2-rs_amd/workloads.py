"""Synthetic workloads of the batched and served configurations (seeded), used
by bench.py and by the GPU tests that run those configurations at full size.

Batched: 256 independent PAR2 sets, input slice counts log-uniform 8..4096,
slice sizes log-uniform 4 KiB..1 MiB (multiples of 4 B), redundancy uniform
1..20 % with the recovery count by the creator's rule (src/creator.rs:324-331),
half creates and half repairs with ceil(U[1,10] % * N) erasures clamped to the
recovery count, total source <= 48 GiB.

Served: Poisson arrivals of create/repair requests; per-request source size
lognormal (median 64 MiB, sigma 1, cap 8 GiB), slice size 256 KiB..1 MiB in
4-byte steps, 30 % repairs with ceil(U[1,10] % * N) erasures.
"""
from __future__ import annotations

import numpy as np


def random_bytes(seed: int, n: int) -> np.ndarray:
    return np.frombuffer(np.random.default_rng(seed).bytes(n), np.uint8)


def draw_sets(seed, count=256, cap_bytes=48 << 30):
    """The batched workload: count sets as dicts n, s, r (recovery slices of
    an encode), kind ("encode"/"repair") and m (erasures of a repair, clamped
    to r); redrawn until the source total fits cap_bytes.  Returns (sets,
    total source bytes)."""
    rng = np.random.default_rng(seed)
    while True:
        sets = []
        for i in range(count):
            n = int(round(np.exp(rng.uniform(np.log(8), np.log(4096)))))
            s = int(round(np.exp(rng.uniform(np.log(4096), np.log(1 << 20))) / 4)) * 4
            pct = rng.uniform(1, 20)
            r = -(-int(n * s * pct / 100.0) // s)
            r = max(r, 1)
            kind = "repair" if i % 2 else "encode"
            m = 0
            if kind == "repair":
                m = min(r, max(1, int(np.ceil(rng.uniform(0.01, 0.10) * n))))
            sets.append({"n": n, "s": s, "r": r, "kind": kind, "m": m})
        total = sum(x["n"] * x["s"] for x in sets)
        if total <= cap_bytes:
            return sets, total


def draw_requests(seed, count, sigma=1.0):
    """The served workload: count requests as dicts t (arrival time in units
    of the mean gap; Exp(1) gaps, scaled by 1/rate when replayed), kind, n, s
    and k (recovery slices of a create, erasures of a repair)."""
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for _ in range(count):
        t += float(rng.exponential(1.0))
        size = min(float(np.exp(np.log(64 * 2**20) + sigma * rng.standard_normal())), 8.0 * 2**30)
        s = int(rng.integers(65536, 262145)) * 4
        n = int(min(32768, max(1, -(-int(size) // s))))
        if rng.random() < 0.3:
            m = max(1, int(np.ceil(rng.uniform(0.01, 0.10) * n)))
            reqs.append({"t": t, "kind": "repair", "n": n, "s": s, "k": m})
        else:
            r = max(1, -(-int(n * s * rng.uniform(1, 20) / 100.0) // s))
            reqs.append({"t": t, "kind": "encode", "n": n, "s": s, "k": r})
    return reqs
