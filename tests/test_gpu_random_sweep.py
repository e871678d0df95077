"""A seeded random sweep over the entry points and contraction variants: for
each case a kind (host encode, device encode on 2-byte-offset buffers with the
IFSC hashes, the allocation-free encode, reconstruct with a random erasure set
and recovery subset, the strategy's region muladd, a small batch), a variant
(automatic, LDS tables, FP4, int8), and random sizes (1-600 slices of 2 B to
70 000 B, 0-40 recovery slices with exponents below 70 000, so some exceed
the field order and wrap).  Every output is compared with the oracle
(encode_avx2, pinned against the scalar oracle; muladd_region; hashlib/zlib).
A reconstruct that finds its matrix singular must find it singular in the
oracle's restatement of the reference transform too (exponents that are equal
modulo 65535, or a singular generalised Vandermonde matrix)."""
from __future__ import annotations

import hashlib
import os
import sys
import zlib

import numpy as np
import pytest

import oracle
import par2hip

pytestmark = pytest.mark.gpu
THREADS = min(16, len(os.sched_getaffinity(0)))


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    if par2hip.available() <= 0:
        pytest.fail("no gfx950 device visible to libpar2hip.so")
    c = par2hip.Context(0)
    yield c
    c.close()


def oracle_encode(data, exps):
    r, s = len(exps), data.shape[1]
    if r == 0:
        return np.zeros((0, s), np.uint8)
    words = np.ascontiguousarray(data.view(np.uint16))
    return oracle.encode_avx2(words, exps, threads=THREADS).view(np.uint8).reshape(r, s)


def one_case(ctx, rng, kind, kern):
    n = int(rng.choice([1, 2, 3, 7, 16, 33, 64, 100, 257, 600]))
    s = 2 * int(rng.choice([1, 8, 24, 500, 1024, 2048, 4104, 8192, 16392, 35000]))
    r = int(rng.integers(0, 41))
    while r and n * s * r > 2e8:
        r //= 2
    exps = np.sort(rng.choice(70000, r, replace=False)).astype(np.uint32)
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    say(kind, "kernel", kern, "n", n, "s", s, "r", r)
    want = oracle_encode(data, exps)
    if kind == "encode":
        assert np.array_equal(ctx.encode(data, exps).reshape(r, s), want)
    elif kind in ("encode_dev", "noalloc"):
        off = 2 * int(rng.integers(0, 9))
        d, dr = ctx.alloc(n * s + 64), ctx.alloc(max(r, 1) * s + 64)
        m5, cr = ctx.alloc(n * 16), ctx.alloc(n * 4)
        try:
            d.upload(data.reshape(-1), offset=off)
            fn = ctx.encode_dev_noalloc if kind == "noalloc" else ctx.encode_dev
            fn(d.view(off), n, s, exps, dr.view(off), m5, cr)
            ctx.sync()
            if r:
                assert np.array_equal(dr.download(r * s, offset=off).reshape(r, s), want)
            md, cc = m5.download().reshape(n, 16), cr.download().view(np.uint32)
            for i in rng.choice(n, min(n, 3), replace=False):
                assert md[i].tobytes() == hashlib.md5(data[i].tobytes()).digest()
                assert int(cc[i]) == zlib.crc32(data[i].tobytes())
        finally:
            for b in (d, dr, m5, cr):
                b.free()
    elif kind == "reconstruct":
        if r == 0 or n < 2:
            return
        m = int(rng.integers(1, min(r, n) + 1))
        missing = np.sort(rng.choice(n, m, replace=False)).astype(np.uint32)
        use = np.sort(rng.choice(r, m, replace=False))
        damaged = data.copy()
        damaged[missing] = 0x33
        try:
            out = ctx.reconstruct(damaged, missing, want[use], exps[use])
        except par2hip.RepairFailed:
            present = np.setdiff1d(np.arange(n), missing)
            with pytest.raises(ValueError):
                oracle.transform(n, missing, present, exps[use])
            return
        assert np.array_equal(out, data[missing])
    elif kind == "region":
        st = par2hip.HipStrategy(ctx)
        words = s // 2
        ns, nd = int(rng.integers(1, 20)), int(rng.integers(1, 10))
        srcs = [rng.integers(0, 65536, words).astype(np.uint16) for _ in range(ns)]
        dsts = [rng.integers(0, 65536, words).astype(np.uint16) for _ in range(nd)]
        coeff = rng.integers(0, 65536, (nd, ns)).astype(np.uint16)
        o = int(rng.integers(0, words))
        ln = int(rng.integers(0, words - o + 1))
        ref = [x.copy() for x in dsts]
        oracle.muladd_region(ref, srcs, coeff, o, ln)
        st.muladd_region(dsts, srcs, coeff, o, ln)
        assert all(np.array_equal(a, b) for a, b in zip(dsts, ref))
    elif kind == "batch":
        jobs, wants = [], []
        try:
            for _ in range(int(rng.integers(1, 6))):
                n2 = int(rng.choice([1, 5, 40, 130]))
                s2 = 2 * int(rng.choice([4, 512, 2052, 8192]))
                r2 = int(rng.integers(1, 12))
                d2 = rng.integers(0, 256, (n2, s2), dtype=np.uint8)
                e2 = np.sort(rng.choice(2000, r2, replace=False)).astype(np.uint32)
                db = ctx.alloc(n2 * s2)
                db.upload(d2.reshape(-1))
                jobs.append({"kind": "encode", "slices": db, "n": n2, "slice_size": s2, "exps": e2,
                             "recovery": ctx.alloc(r2 * s2), "md5": ctx.alloc(n2 * 16), "crc": ctx.alloc(n2 * 4)})
                wants.append(oracle_encode(d2, e2))
            ctx.batch_run_dev(jobs)
            ctx.sync()
            for j, w in zip(jobs, wants):
                assert np.array_equal(j["recovery"].download().reshape(w.shape), w)
        finally:
            for j in jobs:
                for k in ("slices", "recovery", "md5", "crc"):
                    j[k].free()


@pytest.mark.parametrize("seed", [101, 202])
def test_random_sweep(ctx, seed):
    rng = np.random.default_rng(seed)
    kinds = ["encode", "encode_dev", "reconstruct", "region", "batch", "noalloc"]
    for case in range(400):
        kind = str(rng.choice(kinds))
        kern = int(rng.integers(0, 4))
        prev = ctx.set_gf_kernel(kern)
        try:
            say("seed", seed, "case", case)
            one_case(ctx, rng, kind, kern)
        finally:
            ctx.set_gf_kernel(prev)
