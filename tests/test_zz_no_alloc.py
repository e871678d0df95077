"""par2hip_encode_dev_noalloc (the encode entry point that allocates nothing;
its working memory is the library's per-device array) through the C ABI:
recovery slices and IFSC digests equal par2hip_encode_dev's on two calls in a
row, the second smaller than the first (the array keeps the first call's
contents), and a problem beyond the array's size is refused before anything
is queued."""
from __future__ import annotations

import sys

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


def _encode_both(ctx, data, exps, hashes):
    n, s = data.shape
    r = len(exps)
    d_sl = ctx.alloc(n * s)
    d_sl.upload(data)
    outs = []
    for noalloc in (False, True):
        d_rec = ctx.alloc(max(r, 1) * s)
        d_rec.upload(np.full(max(r, 1) * s, 0x5A, np.uint8))
        d_md5 = ctx.alloc(n * 16) if hashes else None
        d_crc = ctx.alloc(n * 4) if hashes else None
        fn = ctx.encode_dev_noalloc if noalloc else ctx.encode_dev
        fn(d_sl, n, s, exps, d_rec, d_md5, d_crc)
        ctx.sync()
        outs.append((d_rec.download(r * s), d_md5.download() if hashes else None,
                     d_crc.download() if hashes else None))
    return outs


@pytest.mark.parametrize("first,second", [((1024, 1 << 20, 102), (300, 8194, 17)),
                                          ((2000, 16384, 40), (7, 4096, 3))])
def test_noalloc_equals_encode_dev_twice(ctx, first, second):
    rng = np.random.default_rng(first[0])
    for k, (n, s, r) in enumerate((first, second)):
        data = rng.integers(0, 256, (n, s), dtype=np.uint8)
        exps = np.arange(r, dtype=np.uint32) if k == 0 else np.sort(rng.choice(70000, r, replace=False)).astype(np.uint32)
        say("noalloc call", k, "n", n, "s", s, "r", r)
        (rec_a, md5_a, crc_a), (rec_b, md5_b, crc_b) = _encode_both(ctx, data, exps, hashes=True)
        assert np.array_equal(rec_a, rec_b), f"call {k}: recovery slices differ from par2hip_encode_dev"
        assert np.array_equal(md5_a, md5_b) and np.array_equal(crc_a, crc_b), f"call {k}: digests differ"
        w = min(64, s // 2)
        cols = np.ascontiguousarray(data.view(np.uint16)[:, :w])
        want = oracle.encode(cols, exps)
        assert np.array_equal(rec_b.reshape(r, s).view(np.uint16)[:, :w], want), f"call {k}: vs oracle"


def test_noalloc_without_hashes_and_without_recovery(ctx):
    rng = np.random.default_rng(5)
    data = rng.integers(0, 256, (33, 1030), dtype=np.uint8)  # 2-byte aligned rows
    (rec_a, _, _), (rec_b, _, _) = _encode_both(ctx, data, np.arange(9, dtype=np.uint32), hashes=False)
    assert np.array_equal(rec_a, rec_b)
    (_, md5_a, crc_a), (_, md5_b, crc_b) = _encode_both(ctx, data, np.zeros(0, np.uint32), hashes=True)
    assert np.array_equal(md5_a, md5_b) and np.array_equal(crc_a, crc_b)


def test_noalloc_refuses_beyond_its_array(ctx):
    n, s = 32768, 4096
    d_sl = ctx.alloc(n * s)
    d_rec = ctx.alloc(64 * s)
    say("noalloc too large")
    with pytest.raises(par2hip.InvalidFormat):  # 32768 x 64 needs 256 MiB of tables
        ctx.encode_dev_noalloc(d_sl, n, s, np.arange(64, dtype=np.uint32), d_rec)
    with pytest.raises(par2hip.InvalidFormat):
        ctx.encode_dev_noalloc(d_sl, 32769, 2, np.arange(1, dtype=np.uint32), d_rec)
