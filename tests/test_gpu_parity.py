"""GPU parity: libpar2hip.so (through the C ABI) against the oracle and the
reference's golden fixture.  Every comparison is bit-exact.

Case names are printed before each GPU step so a failing run names its case.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import zlib

import numpy as np
import pytest

import oracle
import par2hip
from fixtures import GOLDEN, quickpar

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


KERNELS = {"lds": 1, "mfma_fp4": 2, "mfma_i8": 3}


@pytest.fixture(params=list(KERNELS))
def kctx(request, ctx):
    """The context with one contraction kernel forced (LDS tables / bit-sliced
    MFMA with FP4 or int8 operands)."""
    prev = ctx.set_gf_kernel(KERNELS[request.param])
    yield ctx
    ctx.set_gf_kernel(prev)


# ------------------------------------------------------------ golden set --
def test_quickpar_encode_and_ifsc(kctx):
    ctx = kctx
    qp = quickpar()
    say("quickpar encode n=200 S=5376 r=60")
    rec, md5, crc = ctx.encode(qp.data, qp.exponents, hashes=True)
    assert np.array_equal(rec, qp.recovery), "recovery slices differ from QuickPar"
    for i, (m, c) in enumerate(qp.ifsc):
        assert md5[i].tobytes() == m and int(crc[i]) == c, i


@pytest.mark.parametrize("missing", [[0, 1], [3, 4, 5], [16, 17], list(range(103, 136)), [199]])
def test_quickpar_repair(kctx, missing):
    ctx = kctx
    # tests/external_multiblock_repair.rs:49-173 shapes
    qp = quickpar()
    m = len(missing)
    damaged = qp.data.copy()
    damaged[missing] = 0
    say("quickpar repair", missing[:4], "m", m)
    out = ctx.reconstruct(damaged, missing, qp.recovery[:m], qp.exponents[:m])
    assert np.array_equal(out, qp.data[missing])


def test_quickpar_repair_exponent_gaps(ctx):
    qp = quickpar()
    missing, exps = [10, 50, 90, 150], [0, 1, 3, 5]
    say("quickpar repair with exponent gaps")
    out = ctx.reconstruct(qp.data, missing, qp.recovery[exps], exps)
    assert np.array_equal(out, qp.data[missing])


# ----------------------------------------------------- encode vs oracle --
ENCODE_SHAPES = [
    (1, 2, 1), (1, 6, 3), (2, 18, 5), (3, 34, 4), (17, 1000, 9), (37, 4098, 17),
    (16, 4096, 16), (64, 65536, 4), (5, 65538, 33), (200, 5376, 1), (130, 2050, 21), (8, 16, 0),
]


@pytest.mark.parametrize("n,s,r", ENCODE_SHAPES)
def test_encode_matches_oracle(kctx, n, s, r):
    ctx = kctx
    rng = np.random.default_rng(n * 1000 + s + r)
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    say("encode", n, s, r)
    rec, md5, crc = ctx.encode(data, np.arange(r), hashes=True)
    want = oracle.encode(data.view(np.uint16), np.arange(r)).view(np.uint8)
    assert np.array_equal(rec, want.reshape(r, s))
    wm, wc = oracle.hash_slices(data, n, s)
    assert np.array_equal(md5, wm) and np.array_equal(crc, wc)


def test_encode_arbitrary_exponents(kctx):
    ctx = kctx
    rng = np.random.default_rng(99)
    data = rng.integers(0, 256, (29, 1026), dtype=np.uint8)
    exps = [0, 3, 7, 100, 4096, 65534, 65535, 70000]
    say("encode exps", exps)
    rec = ctx.encode(data, exps)
    assert np.array_equal(rec, oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(len(exps), -1))


def test_encode_rejects_bad_sizes(ctx):
    say("encode invalid")
    with pytest.raises(par2hip.InvalidFormat):
        ctx.encode(np.zeros((2, 7), np.uint8), [0])
    with pytest.raises(par2hip.InvalidFormat):
        ctx.encode(np.zeros((32769, 2), np.uint8), [0])


_C2 = {}


def _c2_set():
    """1024 x 1 MiB seeded slices and their 102 recovery slices from the
    oracle's multithreaded AVX2 restatement (pinned against the scalar
    oracle), computed once per session."""
    if not _C2:
        n, s, r = 1024, 1 << 20, 102
        data = np.random.default_rng(2).integers(0, 256, (n, s), dtype=np.uint8)
        threads = min(16, len(os.sched_getaffinity(0)))
        _C2["data"] = data
        _C2["rec"] = oracle.encode_avx2(data.view(np.uint16), np.arange(r), threads=threads)
    return _C2["data"], _C2["rec"]


def test_encode_full_size_every_byte(kctx):
    """1024 x 1 MiB slices, R=102 (the plain-encode workload), with each
    contraction variant: all 102 recovery slices byte for byte."""
    ctx = kctx
    data, want = _c2_set()
    n, s = data.shape
    r = want.shape[0]
    say("encode full size", n, s, r)
    rec = ctx.encode(data, np.arange(r))
    bad = [e for e in range(r) if not np.array_equal(rec[e].view(np.uint16), want[e])]
    assert not bad, f"recovery slices differing from the oracle: {bad[:10]}"


# --------------------------------------------------------------- hashes --
def test_hash_golden_vectors(ctx):
    with open(os.path.join(GOLDEN, "hash_vectors.json")) as f:
        g = json.load(f)
    rng = np.random.default_rng(20261203)
    for v in g["vectors"]:
        b = rng.integers(0, 256, v["seed_len"], dtype=np.uint8)
        say("hash len", len(b))
        buf = b if len(b) else np.zeros(1, np.uint8)
        md5, crc = ctx.hash_slices(buf, 1, len(b), lens=[len(b)])
        assert md5[0].tobytes().hex() == v["md5"] and int(crc[0]) == v["crc32"], len(b)
    rng2 = np.random.default_rng(7)
    ss = g["slice_size"]
    buf = np.zeros((len(g["slices"]), ss), np.uint8)
    lens = []
    for i, s in enumerate(g["slices"]):
        buf[i, : s["len"]] = rng2.integers(0, 256, s["len"], dtype=np.uint8)
        lens.append(s["len"])
    say("hash padded slices")
    md5, crc = ctx.hash_slices(buf, len(lens), ss, lens=lens)
    for i, s in enumerate(g["slices"]):
        assert md5[i].tobytes().hex() == s["md5"] and int(crc[i]) == s["crc32"]


@pytest.mark.parametrize("n,s,stride", [(1, 64, 0), (3, 55, 60), (65, 120, 0), (130, 4096, 4100), (7, 1 << 20, 0),
                                          (5, 262146, 262150), (4, 99999, 100001)])
def test_hash_shapes_vs_stdlib(ctx, n, s, stride):
    rng = np.random.default_rng(n + s)
    st = stride or s
    buf = rng.integers(0, 256, n * st, dtype=np.uint8)
    say("hash", n, s, st)
    md5, crc = ctx.hash_slices(buf, n, s, slice_stride=st)
    for i in range(n):
        b = buf[i * st:i * st + s].tobytes()
        assert md5[i].tobytes() == hashlib.md5(b).digest() and int(crc[i]) == zlib.crc32(b)


@pytest.mark.parametrize("shift,n,s,stride", [(1, 6, 70000, 70001), (2, 9, 4100, 4102), (3, 3, (1 << 20) + 4, (1 << 20) + 7)])
def test_hash_dev_unaligned_bases(ctx, shift, n, s, stride):
    """Slices at any byte address (the device entry point takes them where
    they lie): both hashes equal the stdlib's."""
    rng = np.random.default_rng(shift * 1000 + n)
    buf = rng.integers(0, 256, shift + n * stride, dtype=np.uint8)
    d = ctx.alloc(len(buf))
    d.upload(buf)
    d_md5, d_crc = ctx.alloc(16 * n), ctx.alloc(4 * n)
    say("hash dev unaligned", shift, n, s, stride)
    ctx.hash_slices_dev(d.view(shift), n, s, d_md5, d_crc, slice_stride=stride)
    ctx.sync()
    md5 = d_md5.download().reshape(n, 16)
    crc = d_crc.download().view(np.uint32)
    for i in range(n):
        b = buf[shift + i * stride:shift + i * stride + s].tobytes()
        assert md5[i].tobytes() == hashlib.md5(b).digest(), f"md5 slice {i}"
        assert int(crc[i]) == zlib.crc32(b), f"crc slice {i}"


@pytest.mark.parametrize("n,s", [(4, 16384), (3, 65536), (5, 196608), (2, 1 << 21), (3, 16388), (4, 70000),
                                 (2, 3 * 16384 + 10000), (3, 16380)])
def test_hash_whole_lane_chunks_aligned_and_offset(ctx, n, s):
    """At a 16-byte aligned base the CRC-32 runs the slice's whole 16 KiB rows
    in the row layout (lane l folds bytes [64 l, 64 l + 64) of every row and
    skips the rest with the x^(8 * 16320) tables) and the bytes after them in
    one chunk per lane, joined as raw(rows) x^(8 tail) + raw(tail); a slice
    shorter than a row, and any slice at a 4-byte offset, runs one chunk per
    lane.  Both hashes equal the stdlib's."""
    rng = np.random.default_rng(n * 7 + s)
    buf = rng.integers(0, 256, 16 + n * s, dtype=np.uint8)
    d = ctx.alloc(len(buf))
    d.upload(buf)
    for shift in (0, 4):
        d_md5, d_crc = ctx.alloc(16 * n), ctx.alloc(4 * n)
        say("hash whole lane chunks", n, s, shift)
        ctx.hash_slices_dev(d.view(shift), n, s, d_md5, d_crc)
        ctx.sync()
        md5 = d_md5.download().reshape(n, 16)
        crc = d_crc.download().view(np.uint32)
        for i in range(n):
            b = buf[shift + i * s:shift + (i + 1) * s].tobytes()
            assert int(crc[i]) == zlib.crc32(b), f"crc slice {i} shift {shift}"
            assert md5[i].tobytes() == hashlib.md5(b).digest(), f"md5 slice {i} shift {shift}"


def test_hash_ragged_lengths_zero_padded(ctx):
    rng = np.random.default_rng(3)
    s = 5376
    lens = [0, 1, 55, 56, 63, 64, 65, 5375, 5376, 2688]
    buf = rng.integers(0, 256, len(lens) * s, dtype=np.uint8)
    say("hash ragged")
    md5, crc = ctx.hash_slices(buf, len(lens), s, lens=lens)
    for i, ln in enumerate(lens):
        b = buf[i * s:i * s + ln].tobytes() + bytes(s - ln)
        assert md5[i].tobytes() == hashlib.md5(b).digest() and int(crc[i]) == zlib.crc32(b), ln


def test_hash_md5_loop_boundaries(ctx):
    """Every count of whole 64-byte blocks from 0 to 40, plus 0, 8 or 56
    bytes, at 16-, 4- and 1-byte aligned bases, with one length for all the
    slices and with a length per slice: the MD5 chain's unclamped main loop
    (six blocks per pass while the prefetch stays inside the slice), its
    clamped loop and its tail hand over at different block counts, and each
    hand-over is crossed.  Both hashes equal the stdlib's."""
    rng = np.random.default_rng(64)
    sizes = [64 * k + r for k in range(41) for r in (0, 8, 56) if 64 * k + r]
    stride = 64 * 42  # a multiple of 16
    n = len(sizes)
    for shift in (0, 4, 1):
        buf = rng.integers(0, 256, shift + n * stride, dtype=np.uint8)
        d = ctx.alloc(len(buf))
        d.upload(buf)
        d_md5, d_crc = ctx.alloc(16 * n), ctx.alloc(4 * n)
        say("md5 loop boundaries, a length per slice, base shift", shift)
        ctx.hash_slices_dev(d.view(shift), n, stride, d_md5, d_crc, slice_stride=stride, lens=sizes)
        ctx.sync()
        md5 = d_md5.download().reshape(n, 16)
        crc = d_crc.download().view(np.uint32)
        for i, s in enumerate(sizes):
            b = buf[shift + i * stride:shift + i * stride + s].tobytes() + bytes(stride - s)
            assert md5[i].tobytes() == hashlib.md5(b).digest(), (shift, s)
            assert int(crc[i]) == zlib.crc32(b), (shift, s)
        say("md5 loop boundaries, one length for all, base shift", shift)
        for s in sizes:
            ctx.hash_slices_dev(d.view(shift), 3, s, d_md5, d_crc, slice_stride=stride)
            ctx.sync()
            md5 = d_md5.download()[:48].reshape(3, 16)
            crc = d_crc.download()[:12].view(np.uint32)
            for i in range(3):
                b = buf[shift + i * stride:shift + i * stride + s].tobytes()
                assert md5[i].tobytes() == hashlib.md5(b).digest(), (shift, s, i)
                assert int(crc[i]) == zlib.crc32(b), (shift, s, i)


@pytest.mark.parametrize("s", [256, 4160])
def test_hash_more_chain_waves_than_cus(ctx, s):
    """16385 full, 16-byte aligned slices: more 64-chain waves than the
    device has CUs, so the chains run without a helper wave (the form that
    needs a CU per chain wave is not taken); both hashes equal the stdlib's
    on a sample of slices, the last one included."""
    rng = np.random.default_rng(s)
    n = 64 * ctx.device_info()["compute_units"] + 1
    buf = rng.integers(0, 256, n * s, dtype=np.uint8)
    d = ctx.alloc(len(buf))
    d.upload(buf)
    d_md5, d_crc = ctx.alloc(16 * n), ctx.alloc(4 * n)
    say("hash", n, "slices of", s, "bytes")
    ctx.hash_slices_dev(d, n, s, d_md5, d_crc)
    ctx.sync()
    md5 = d_md5.download().reshape(n, 16)
    crc = d_crc.download().view(np.uint32)
    for i in list(range(0, n, 997)) + [n - 1]:
        b = buf[i * s:(i + 1) * s].tobytes()
        assert md5[i].tobytes() == hashlib.md5(b).digest(), i
        assert int(crc[i]) == zlib.crc32(b), i


# --------------------------------------------------------------- repair --
@pytest.mark.parametrize("m", [1, 2, 17, 100, 300])
def test_gf_invert_vs_oracle(ctx, m):
    rng = np.random.default_rng(m)
    c = oracle.par2_constants(2000)
    cols = rng.choice(2000, m, replace=False)
    exps = rng.choice(3000, m, replace=False)
    A = np.array([[oracle.gf_pow(int(c[j]), int(e)) for j in cols] for e in exps], np.uint16)
    say("gf_invert m", m)
    assert np.array_equal(ctx.gf_invert(A), oracle.gf_invert(A))


def test_gf_invert_singular(ctx):
    A = np.array([[1, 2, 1], [3, 4, 3], [5, 6, 5]], np.uint16)
    say("gf_invert singular")
    with pytest.raises(par2hip.RepairFailed):
        ctx.gf_invert(A)
    A = np.array([[0, 1], [0, 7]], np.uint16)
    with pytest.raises(par2hip.RepairFailed):
        ctx.gf_invert(A)


@pytest.mark.parametrize("n,s,r,m", [(10, 64, 3, 3), (64, 1000, 8, 5), (300, 4098, 40, 40), (33, 2, 4, 2)])
def test_encode_erase_reconstruct(kctx, n, s, r, m):
    ctx = kctx
    rng = np.random.default_rng(n + s + r + m)
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    say("roundtrip", n, s, r, m)
    rec = ctx.encode(data, np.arange(r))
    missing = sorted(rng.choice(n, m, replace=False).tolist())
    use = sorted(rng.choice(r, m, replace=False).tolist())
    damaged = data.copy()
    damaged[missing] = 0xA5
    out = ctx.reconstruct(damaged, missing, rec[use], use)
    assert np.array_equal(out, data[missing])
    # the oracle's faithful transform+replay agrees
    present = [d for d in range(n) if d not in missing]
    want, _ = oracle.reconstruct(n, missing, present, use, rec[use].view(np.uint16), data[present].view(np.uint16))
    assert np.array_equal(out.view(np.uint16), want)


def test_reconstruct_singular_and_insufficient(ctx):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, (8, 32), dtype=np.uint8)
    rec = ctx.encode(data, [0, 1, 2])
    say("reconstruct errors")
    with pytest.raises(par2hip.RepairFailed):
        ctx.reconstruct(data, [1, 2], rec[[0, 0]], [0, 0])
    with pytest.raises(par2hip.InsufficientRecovery):
        ctx.reconstruct(data, [1, 2, 3], rec[:2], [0, 1])


def test_reconstruct_dev_in_place(kctx):
    ctx = kctx
    rng = np.random.default_rng(5)
    n, s, r = 50, 4096, 6
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    rec = ctx.encode(data, np.arange(r))
    missing = [0, 7, 49]
    damaged = data.copy()
    damaged[missing] = 0
    d_sl = ctx.alloc(n * s)
    d_rec = ctx.alloc(3 * s)
    d_sl.upload(damaged)
    d_rec.upload(rec[[1, 3, 5]])
    say("reconstruct_dev in place")
    ctx.reconstruct_dev(d_sl, n, s, missing, d_rec, [1, 3, 5], in_place=True)
    ctx.sync()
    assert np.array_equal(d_sl.download().reshape(n, s), data)


# ----------------------------------------------------- strategy methods --
SIZES = [0, 1, 7, 15, 16, 17, 31, 33, 100, 1000]
SCALARS = [0, 1, 2, 12345, 65535]


def test_strategy_identity(ctx):
    s = par2hip.HipStrategy(ctx)
    assert s.name() == "HIP-gfx950" and s.priority() == 40 and s.is_available()
    assert not s.supports_shuffle2x()


@pytest.mark.parametrize("size", SIZES)
def test_strategy_mul_slice_and_muladd(ctx, size):
    # src/galois/tests.rs:255-298 sweep, GPU strategy against the scalar oracle
    s = par2hip.HipStrategy(ctx)
    rng = np.random.default_rng(size)
    for k in SCALARS:
        src = rng.integers(0, 65536, size).astype(np.uint16)
        dst = rng.integers(0, 65536, size).astype(np.uint16)
        want = dst.copy()
        oracle.muladd(want, src, k)
        got = dst.copy()
        say("muladd", size, k)
        s.muladd(got, src, k)
        assert np.array_equal(got, want)
        d1, d2 = src.copy(), src.copy()
        oracle.mul_slice(k, d1)
        s.mul_slice(k, d2)
        assert np.array_equal(d1, d2)


def test_strategy_muladd_region(kctx):
    ctx = kctx
    s = par2hip.HipStrategy(ctx)
    rng = np.random.default_rng(3)
    nd, ns, L = 5, 9, 1031
    srcs = [rng.integers(0, 65536, L).astype(np.uint16) for _ in range(ns)]
    dsts = [rng.integers(0, 65536, L).astype(np.uint16) for _ in range(nd)]
    coeff = rng.integers(0, 65536, (nd, ns + 3)).astype(np.uint16)
    coeff[1, 2] = 0
    want = [d.copy() for d in dsts]
    oracle.muladd_region(want, srcs, coeff, 13, 1000)
    say("muladd_region")
    s.muladd_region(dsts, srcs, list(coeff), 13, 1000)
    for a, b in zip(dsts, want):
        assert np.array_equal(a, b)


def test_strategy_muladd_region_ragged_rows(kctx):
    """Coefficient rows shorter and longer than the source list: a short row
    applies only its own entries (src/galois/scalar.rs:100)."""
    s = par2hip.HipStrategy(kctx)
    rng = np.random.default_rng(11)
    ns, L = 7, 777
    srcs = [rng.integers(0, 65536, L).astype(np.uint16) for _ in range(ns)]
    lens = [0, 1, 3, 6, 7, 9, 12]
    rows = [rng.integers(1, 65536, k).astype(np.uint16) for k in lens]
    dsts = [rng.integers(0, 65536, L).astype(np.uint16) for _ in lens]
    want = [d.copy() for d in dsts]
    oracle.muladd_region_rows(want, srcs, rows, 5, 700)
    # the short rows really leave sources out: the full-width product differs
    full = [d.copy() for d in dsts]
    padded = np.zeros((len(lens), ns), np.uint16)
    for i, r in enumerate(rows):
        padded[i, :min(len(r), ns)] = r[:ns]
    padded[0] = 1
    oracle.muladd_region(full, srcs, padded, 5, 700)
    assert not np.array_equal(full[0], want[0])
    say("muladd_region ragged rows", lens)
    s.muladd_region(dsts, srcs, rows, 5, 700)
    for a, b in zip(dsts, want):
        assert np.array_equal(a, b)


# ---------------------------------------------------- verify + repair set --
CRC_SILENT = bytes.fromhex("410671db01")  # the CRC-32 polynomial: XOR-ing it never changes CRC-32


def _set_with_damage(n, s, r, seed):
    rng = np.random.default_rng(seed)
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    md5, crc = oracle.hash_slices(data, n, s)
    rec = oracle.encode(data.view(np.uint16), np.arange(r)).view(np.uint8).reshape(r, s)
    return rng, data, md5, crc, rec


@pytest.mark.parametrize("inplace", [True, False])
def test_repair_set_crc_detected(kctx, inplace):
    ctx = kctx
    n, s, r = 120, 4096, 12
    rng, data, md5, crc, rec = _set_with_damage(n, s, r, 31)
    erased = sorted(rng.choice(n, 7, replace=False).tolist())
    damaged = data.copy()
    damaged[erased] = 0
    perm = rng.permutation(r)  # recovery rows in shuffled order: the lowest exponents must be chosen
    d_sl, d_rec = ctx.alloc(n * s), ctx.alloc(r * s)
    d_sl.upload(damaged)
    d_rec.upload(rec[perm])
    d_out = None if inplace else ctx.alloc(r * s)
    say("repair_set crc", erased, "inplace", inplace)
    got = ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, perm, d_out)
    assert got.tolist() == erased
    if inplace:
        assert np.array_equal(d_sl.download().reshape(n, s), data)
    else:
        out = d_out.download().reshape(r, s)[: len(erased)]
        assert np.array_equal(out, data[erased])
        assert np.array_equal(d_sl.download().reshape(n, s), damaged)


def test_repair_set_md5_only_damage(kctx):
    """A CRC-silent corruption is caught by the MD5 confirmation and the
    speculative rebuild is redone with the extra slice (verify.rs:621-645)."""
    ctx = kctx
    n, s, r = 64, 2048, 6
    rng, data, md5, crc, rec = _set_with_damage(n, s, r, 32)
    damaged = data.copy()
    damaged[[5, 40]] = 0
    silent = 17
    damaged[silent, 100:105] ^= np.frombuffer(CRC_SILENT, np.uint8)
    assert zlib.crc32(damaged[silent].tobytes()) == int(crc[silent])
    d_sl, d_rec = ctx.alloc(n * s), ctx.alloc(r * s)
    d_sl.upload(damaged)
    d_rec.upload(rec)
    say("repair_set md5-only damage")
    got = ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, np.arange(r))
    assert got.tolist() == [5, silent, 40]
    assert np.array_equal(d_sl.download().reshape(n, s), data)


def test_repair_set_clean_and_insufficient(ctx):
    n, s, r = 40, 1024, 3
    rng, data, md5, crc, rec = _set_with_damage(n, s, r, 33)
    d_sl, d_rec = ctx.alloc(n * s), ctx.alloc(r * s)
    d_sl.upload(data)
    d_rec.upload(rec)
    say("repair_set clean")
    assert ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, np.arange(r)).tolist() == []
    damaged = data.copy()
    damaged[[1, 2, 3, 4]] = 0
    d_sl.upload(damaged)
    say("repair_set insufficient")
    with pytest.raises(par2hip.InsufficientRecovery):
        ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, np.arange(r))
    # nothing was written back
    assert np.array_equal(d_sl.download().reshape(n, s), damaged)


def test_repair_set_quickpar(ctx):
    """The golden set: delete a whole 33-slice file and two slices elsewhere,
    repair from the QuickPar recovery volumes using their IFSC entries."""
    qp = quickpar()
    n, s = qp.data.shape
    md5 = np.frombuffer(b"".join(m for m, _ in qp.ifsc), np.uint8).reshape(n, 16)
    crc = np.array([c for _, c in qp.ifsc], np.uint32)
    lost = list(range(103, 136)) + [0, 199]
    damaged = qp.data.copy()
    damaged[lost] = 0
    d_sl, d_rec = ctx.alloc(n * s), ctx.alloc(qp.recovery.nbytes)
    d_sl.upload(damaged)
    d_rec.upload(qp.recovery)
    say("repair_set quickpar", len(lost))
    got = ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, qp.exponents)
    assert got.tolist() == sorted(lost)
    assert np.array_equal(d_sl.download().reshape(n, s), qp.data)


@pytest.mark.parametrize("exps", [[0, 1, 2, 3, 4], [3, 4, 5, 6, 7], [10, 11], [0, 2, 3, 4, 5], [65530, 65531, 65532]])
def test_reconstruct_inverse_paths(kctx, exps):
    """Consecutive exponents take the closed-form (Lagrange) inverse, others
    Gauss-Jordan; both must rebuild the same bytes."""
    ctx = kctx
    rng = np.random.default_rng(sum(exps))
    n, s = 90, 1026
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    rec = oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(len(exps), s)
    missing = sorted(rng.choice(n, len(exps), replace=False).tolist())
    say("inverse path", exps)
    out = ctx.reconstruct(data, missing, rec, exps)
    assert np.array_equal(out, data[missing])


def test_reconstruct_many_erasures(kctx):
    """m = 300 erasures: closed-form inverse beyond the LDS Gauss-Jordan size."""
    ctx = kctx
    rng = np.random.default_rng(300)
    n, s, m = 700, 512, 300
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    rec = ctx.encode(data, np.arange(m))
    missing = sorted(rng.choice(n, m, replace=False).tolist())
    say("reconstruct m=300")
    out = ctx.reconstruct(data, missing, rec, np.arange(m))
    assert np.array_equal(out, data[missing])


# ------------------------------------------------- automatic kernel choice --
def _encode_dev_profiled(ctx, data, r):
    n, s = data.shape
    d_sl, d_rec, d_md5, d_crc = ctx.alloc(n * s), ctx.alloc(r * s), ctx.alloc(n * 16), ctx.alloc(n * 4)
    d_sl.upload(data.reshape(-1))
    prev = ctx.set_gf_kernel(0)
    ctx.profile(True)
    ctx.profile_read()
    try:
        ctx.encode_dev(d_sl, n, s, np.arange(r, dtype=np.uint32), d_rec, d_md5, d_crc)
        ctx.sync()
        stats = ctx.profile_read()
    finally:
        ctx.profile(False)
        ctx.set_gf_kernel(prev)
    rec = d_rec.download().reshape(r, s)
    md5 = d_md5.download().reshape(n, 16)
    for b in (d_sl, d_rec, d_md5, d_crc):
        b.free()
    return rec, md5, stats


def test_auto_choice_under_md5_chains(ctx):
    """1024 x 1 MiB, R=102 with the IFSC hashes: the matrix-core contraction
    fits under the MD5 chains, so the automatic choice takes it (on a device
    with reserved hash CUs), and the step stays bit-exact."""
    n, s, r = 1024, 1 << 20, 102
    data = np.random.default_rng(21).integers(0, 256, (n, s), dtype=np.uint8)
    say("auto choice, encode", n, s, r)
    rec, md5, stats = _encode_dev_profiled(ctx, data, r)
    if ctx.device_info()["hash_cus"] > 0:
        assert "gf_muladd_mfma" in stats and "gf_muladd_lds" not in stats, stats
    for w0 in (0, 123457):
        cols = np.ascontiguousarray(data.view(np.uint16)[:, w0:w0 + 64])
        assert np.array_equal(rec.view(np.uint16)[:, w0:w0 + 64], oracle.encode(cols, np.arange(r))), w0
    for i in (0, 517, n - 1):
        assert md5[i].tobytes() == hashlib.md5(data[i].tobytes()).digest(), i


def test_auto_choice_under_md5_chains_ragged(ctx):
    """512 slices of 512 KiB + 48 bytes, R=60 with the IFSC hashes: the
    matrix-core contraction beside the chains, its last tile column running
    past the end of the rows; bit-exact at the start and at the ragged end."""
    n, s, r = 512, (1 << 19) + 48, 60
    data = np.random.default_rng(24).integers(0, 256, (n, s), dtype=np.uint8)
    say("auto choice, ragged encode", n, s, r)
    rec, md5, stats = _encode_dev_profiled(ctx, data, r)
    if ctx.device_info()["hash_cus"] > 0:
        assert "gf_muladd_mfma" in stats and "gf_muladd_lds" not in stats, stats
    words = s // 2
    for w0 in (0, words - 300, words - 64):
        cols = np.ascontiguousarray(data.view(np.uint16)[:, w0:w0 + 64])
        assert np.array_equal(rec.view(np.uint16)[:, w0:w0 + 64], oracle.encode(cols, np.arange(r))), w0
    for i in (0, n - 1):
        assert md5[i].tobytes() == hashlib.md5(data[i].tobytes()).digest(), i


def test_auto_choice_short_chains_keep_lds(ctx):
    """2000 x 8 KiB, R=200: the chains are far shorter than a matrix-core
    contraction of this size, so the LDS-table kernel runs."""
    n, s, r = 2000, 8192, 200
    data = np.random.default_rng(22).integers(0, 256, (n, s), dtype=np.uint8)
    say("auto choice, short chains", n, s, r)
    rec, _, stats = _encode_dev_profiled(ctx, data, r)
    assert "gf_muladd_lds" in stats and "gf_muladd_mfma" not in stats, stats
    cols = np.ascontiguousarray(data.view(np.uint16)[:, 1000:1064])
    assert np.array_equal(rec.view(np.uint16)[:, 1000:1064], oracle.encode(cols, np.arange(r)))


def test_auto_choice_few_outputs_take_fp4_tile(ctx):
    """At most 4 outputs with 16-byte aligned rows: the automatic choice is the
    FP4 form's 4-output tile (each input word read once), beside the MD5
    chains too; bit-exact against the oracle."""
    rng = np.random.default_rng(23)
    for n, s, r in [(300, 65536, 4), (64, 8192, 1)]:
        data = rng.integers(0, 256, (n, s), dtype=np.uint8)
        say("auto choice, few outputs", n, s, r)
        d_sl, d_rec = ctx.alloc(n * s), ctx.alloc(r * s)
        d_sl.upload(data.reshape(-1))
        prev = ctx.set_gf_kernel(0)
        ctx.profile(True)
        ctx.profile_read()
        try:
            ctx.encode_dev(d_sl, n, s, np.arange(r, dtype=np.uint32), d_rec)
            ctx.sync()
            stats = ctx.profile_read()
        finally:
            ctx.profile(False)
            ctx.set_gf_kernel(prev)
        assert "gf_muladd_mfma" in stats and "gf_muladd_lds" not in stats, stats
        rec = d_rec.download().reshape(r, s)
        assert np.array_equal(rec, oracle.encode(data.view(np.uint16), np.arange(r)).view(np.uint8).reshape(r, s))
        rec2, _, _ = _encode_dev_profiled(ctx, data, r)  # with the hashes beside it
        assert np.array_equal(rec2, rec)


def test_max_sources_32768(kctx):
    """The largest set PAR2's constants allow (32768 input slices, par2_rs.rs:64-76
    with n up to 32768): encode with exponents up to the field's order against the
    oracle, then rebuild erased slices from the recovery slices.  64-byte slices
    keep the oracle fast and put every word tile past the row end."""
    ctx = kctx
    n, s = 32768, 64
    exps = [0, 1, 65533, 65534]
    data = np.random.default_rng(32768).integers(0, 256, (n, s), dtype=np.uint8)
    say("encode n=32768", s, exps)
    rec = ctx.encode(data, exps)
    want = oracle.encode(np.ascontiguousarray(data.view(np.uint16)), np.asarray(exps))
    assert np.array_equal(rec.view(np.uint16), want)
    missing = [0, 4097, 20000, 32767]
    out = ctx.reconstruct(data, missing, rec, exps)
    assert np.array_equal(out, data[missing])
