"""The full sizes of the plain-encode and plain-repair workloads, byte for
byte: every recovery slice of the encode against the oracle's multithreaded
restatement of the reference's AVX2 strategy (itself pinned against the scalar
oracle), every IFSC digest against the oracle's MD5/CRC32, then encode -> erase
-> rebuild gives back every erased slice exactly and the repair finds exactly
the erased slices (par2_rs.rs:785-902 encoder, repair.rs:536-604 damaged set
and lowest exponents, par2_rs.rs:319-675 rebuild)."""
from __future__ import annotations

import hashlib
import sys
import zlib

import os

import numpy as np
import pytest

import oracle
import par2hip

pytestmark = pytest.mark.gpu


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


def _seeded(seed, n, s):
    return np.random.default_rng(seed).integers(0, 256, (n, s), dtype=np.uint8)


@pytest.mark.parametrize("n,s,r", [(2000, 512 << 10, 100), (1024, 1 << 20, 102)])
def test_encode_erase_repair_full_size(ctx, n, s, r):
    """2000 x 512 KiB with 100 recovery slices, and 1024 x 1 MiB with 102
    (the encode step's automatic matrix-core plan runs there): every erased
    slice comes back through repair_set_dev, which verifies all slices first."""
    data = _seeded(n + r, n, s)
    say("full-size round trip", n, s, r)
    d_sl, d_rec = ctx.alloc(n * s), ctx.alloc(r * s)
    d_md5, d_crc = ctx.alloc(n * 16), ctx.alloc(n * 4)
    d_sl.upload(data)
    ctx.encode_dev(d_sl, n, s, np.arange(r, dtype=np.uint32), d_rec, d_md5, d_crc)
    ctx.sync()
    md5 = d_md5.download().reshape(n, 16)
    crc = d_crc.download().view(np.uint32)
    for i in (0, n // 3, n - 1):
        assert md5[i].tobytes() == hashlib.md5(data[i].tobytes()).digest(), i
        assert int(crc[i]) == zlib.crc32(data[i].tobytes()), i
    # all of it against the oracle
    threads = min(16, len(os.sched_getaffinity(0)))
    say("oracle: AVX2 encode + hashes of the whole set, threads", threads)
    want_rec = oracle.encode_avx2(data.view(np.uint16), np.arange(r), threads=threads)
    rec = d_rec.download().reshape(r, s)
    bad = [e for e in range(r) if not np.array_equal(rec[e].view(np.uint16), want_rec[e])]
    assert not bad, f"recovery slices differing from the oracle: {bad[:10]}"
    del want_rec
    wm, wc = oracle.hash_slices_mt(data, n, s, threads=threads)
    assert np.array_equal(md5, wm) and np.array_equal(crc, wc)
    rng = np.random.default_rng(7 * n)
    erased = np.sort(rng.choice(n, r, replace=False))
    damaged = data.copy()
    damaged[erased] = 0
    d_sl.upload(damaged)
    d_out = ctx.alloc(r * s)
    got = ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, np.arange(r, dtype=np.uint32), d_out)
    assert got.tolist() == erased.tolist()
    out = d_out.download().reshape(r, s)
    assert np.array_equal(out, data[erased])
    for b in (d_sl, d_rec, d_md5, d_crc, d_out):
        b.free()
