"""Pins the CPU oracle before it is trusted as the checker.

Each check is independent of the oracle's own formula: the reference's golden
fixture (QuickPar set, all 60 recovery slices and 200 IFSC entries), the
reference's known-answer tests, hashlib/zlib, carry-less multiplication mod the
PAR2 polynomial computed here from first principles, and round-trip invariants.
"""
from __future__ import annotations

import hashlib
import json
import os
import zlib

import numpy as np
import pytest

import oracle
from fixtures import GOLDEN, quickpar

POLY = 0x1100B


def clmul_mod(a: int, b: int) -> int:
    """Schoolbook GF(2)[x] product reduced mod x^16+x^12+x^3+x+1."""
    p = 0
    for i in range(16):
        if (b >> i) & 1:
            p ^= a << i
    for bit in range(30, 15, -1):
        if (p >> bit) & 1:
            p ^= POLY << (bit - 16)
    return p


def clmul_mod_np(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = a.astype(np.uint32)
    b = b.astype(np.uint32)
    p = np.zeros_like(a)
    for i in range(16):
        p ^= np.where((b >> i) & 1, a << i, 0).astype(np.uint32)
    for bit in range(30, 15, -1):
        p ^= np.where((p >> bit) & 1, np.uint32(POLY << (bit - 16)), 0).astype(np.uint32)
    return p.astype(np.uint16)


# ---------------------------------------------------------------- field --
def test_reference_field_kats():
    # tests/galois/core.rs:8-131 and src/galois/tests.rs:18-87
    assert oracle.gf_mul(1, 5) == 5 and oracle.gf_mul(5, 1) == 5
    assert oracle.gf_mul(0, 5) == 0 and oracle.gf_mul(5, 0) == 0
    assert oracle.gf_mul(3, 7) == oracle.gf_mul(7, 3)
    assert oracle.gf_mul(123, 456) == oracle.gf_mul(456, 123)
    assert oracle.gf_div(5, 5) == 1 and oracle.gf_div(1234, 1234) == 1
    assert oracle.gf_div(5, 1) == 5 and oracle.gf_div(0, 5) == 0
    assert oracle.gf_div(5, 0) is None and oracle.gf_div(0, 0) is None
    assert oracle.gf_pow(5, 0) == 1 and oracle.gf_pow(0, 0) == 1
    assert oracle.gf_pow(5, 1) == 5 and oracle.gf_pow(0, 5) == 0 and oracle.gf_pow(0, 100) == 0
    for n in range(5):
        assert oracle.gf_pow(2, n) == 1 << n
    a = 12345
    assert oracle.gf_pow(a, 2) == oracle.gf_mul(a, a)
    assert oracle.gf_pow(a, 7) == oracle.gf_mul(oracle.gf_pow(a, 3), oracle.gf_pow(a, 4))
    inv = oracle.gf_div(1, a)
    assert oracle.gf_mul(a, inv) == 1


def test_tables_match_polynomial_walk():
    log, exp = oracle.tables()
    # exp is the orbit of the generator 2 under clmul_mod: a permutation of 1..65535
    assert sorted(exp[:65535].tolist()) == list(range(1, 65536))
    assert np.array_equal(exp[:65535], exp[65535:])
    nxt = clmul_mod_np(exp[:65535], np.full(65535, 2, np.uint16))
    assert np.array_equal(nxt[:-1], exp[1:65535]) and nxt[-1] == 1
    assert log[0] == 0
    assert np.array_equal(log[exp[:65535]], np.arange(65535, dtype=np.uint16))


def test_gf_mul_equals_carryless_product():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 65536, 4000).astype(np.uint16)
    b = rng.integers(0, 65536, 4000).astype(np.uint16)
    a[:4] = [0, 1, 65535, 0x8000]
    b[:4] = [65535, 65535, 65535, 0x8000]
    want = clmul_mod_np(a, b)
    got = np.array([oracle.gf_mul(int(x), int(y)) for x, y in zip(a, b)], np.uint16)
    assert np.array_equal(got, want)


def test_gf_pow_by_repeated_product():
    for a in (2, 3, 0x1100B & 0xFFFF, 65535, 40000):
        acc = 1
        for n in range(0, 40):
            assert oracle.gf_pow(a, n) == acc
            acc = clmul_mod(acc, a)
    # order of 2 is 65535; exponents reduce mod 65535 (core.rs:256)
    assert oracle.gf_pow(2, 65535) == 1
    assert oracle.gf_pow(2, 65536) == 2


def test_par2_constants():
    # n_k = 1,2,4,7,8,11,13,14,16,19,... (no multiple of 3,5,17,257), c = 2^n
    ns = [n for n in range(1, 5000) if n % 3 and n % 5 and n % 17 and n % 257][:2000]
    assert ns[:10] == [1, 2, 4, 7, 8, 11, 13, 14, 16, 19]
    c = oracle.par2_constants(2000)
    want = []
    x = 1
    k = 0
    for n in range(1, ns[-1] + 1):
        x = clmul_mod(x, 2)
        if k < len(ns) and n == ns[k]:
            want.append(x)
            k += 1
    assert c.tolist() == want
    # every constant has order 65535: c^(65535/p) != 1 for p in {3,5,17,257}
    for cc in c[:64]:
        for p in (3, 5, 17, 257):
            assert oracle.gf_pow(int(cc), 65535 // p) != 1


def test_coeff_matrix_contiguous_and_general_agree():
    n = 300
    a = oracle.coeff_matrix(n, np.arange(20))
    b = oracle.coeff_matrix(n, np.array([0] + list(range(1, 20)), np.uint32) + np.uint32(0))
    assert np.array_equal(a, b)
    c = oracle.par2_constants(n)
    for r in (0, 1, 5, 19):
        for d in (0, 7, 299):
            assert a[r, d] == oracle.gf_pow(int(c[d]), r)


# --------------------------------------------------------- slice ops --
def test_slice_ops_brute_force():
    rng = np.random.default_rng(2)
    for size in (0, 1, 7, 15, 16, 17, 31, 33, 100, 1000):
        for k in (0, 1, 2, 12345, 65535):
            src = rng.integers(0, 65536, size).astype(np.uint16)
            dst = rng.integers(0, 65536, size).astype(np.uint16)
            want = dst ^ clmul_mod_np(src, np.full(size, k, np.uint16))
            got = dst.copy()
            oracle.muladd(got, src, k)
            assert np.array_equal(got, want)
            d2 = src.copy()
            oracle.mul_slice(k, d2)
            assert np.array_equal(d2, clmul_mod_np(src, np.full(size, k, np.uint16)))


def test_region_brute_force():
    rng = np.random.default_rng(3)
    nd, ns, L = 5, 9, 257
    srcs = [rng.integers(0, 65536, L).astype(np.uint16) for _ in range(ns)]
    dsts = [rng.integers(0, 65536, L).astype(np.uint16) for _ in range(nd)]
    coeff = rng.integers(0, 65536, (nd, ns + 3)).astype(np.uint16)
    coeff[1, 2] = 0
    want = [d.copy() for d in dsts]
    off, ln = 13, 200
    for i in range(nd):
        for j in range(ns):  # .take(num_srcs): the 3 extra coefficients are ignored
            want[i][off:off + ln] ^= clmul_mod_np(srcs[j][off:off + ln], np.full(ln, coeff[i, j], np.uint16))
    oracle.muladd_region(dsts, srcs, coeff, off, ln)
    for a, b in zip(dsts, want):
        assert np.array_equal(a, b)


def test_region_ragged_rows_brute_force():
    """Rows shorter than the source list apply only their own entries
    (src/galois/scalar.rs:100): against the carry-less product per entry."""
    rng = np.random.default_rng(4)
    ns, L = 6, 131
    srcs = [rng.integers(0, 65536, L).astype(np.uint16) for _ in range(ns)]
    lens = [0, 2, 5, 6, 9]
    rows = [rng.integers(0, 65536, k).astype(np.uint16) for k in lens]
    dsts = [rng.integers(0, 65536, L).astype(np.uint16) for _ in lens]
    want = [d.copy() for d in dsts]
    off, ln = 7, 100
    for i, row in enumerate(rows):
        for j in range(min(len(row), ns)):
            want[i][off:off + ln] ^= clmul_mod_np(srcs[j][off:off + ln], np.full(ln, row[j], np.uint16))
    oracle.muladd_region_rows(dsts, srcs, rows, off, ln)
    for a, b in zip(dsts, want):
        assert np.array_equal(a, b)


# ------------------------------------------------------ golden fixture --
def test_quickpar_recovery_slices_bit_exact():
    qp = quickpar()
    assert qp.slice_size == 5376 and qp.data.shape[0] == 200
    assert qp.exponents == list(range(60))
    rec = oracle.encode(qp.words, qp.exponents)
    assert np.array_equal(rec, qp.rec_words), "all 60 recovery slices must match QuickPar"


def test_quickpar_ifsc_bit_exact():
    qp = quickpar()
    md5s, crcs = oracle.hash_slices(qp.data, qp.data.shape[0], qp.slice_size)
    for i, (m, c) in enumerate(qp.ifsc):
        assert md5s[i].tobytes() == m, i
        assert int(crcs[i]) == c, i


def test_quickpar_unpadded_tail_hash_equals_padded():
    # hashing the unpadded tail with zero-extension == hashing the padded slice
    qp = quickpar()
    last = [k for k, (fi, s) in enumerate(qp.owner) if s == -(-qp.files[fi].length // qp.slice_size) - 1]
    lens = np.full(qp.data.shape[0], qp.slice_size, np.uintp)
    for k in last:
        fi, s = qp.owner[k]
        lens[k] = qp.files[fi].length - s * qp.slice_size
    m1, c1 = oracle.hash_slices(qp.data, qp.data.shape[0], qp.slice_size, lens=lens)
    m2, c2 = oracle.hash_slices(qp.data, qp.data.shape[0], qp.slice_size)
    assert np.array_equal(m1, m2) and np.array_equal(c1, c2)


@pytest.mark.parametrize("case", [
    # tests/external_multiblock_repair.rs:49-173: 2 slices in one file, 3 slices,
    # slices across 2 files, a whole 33-slice file, 1 slice
    ("two_in_one_file", [0, 1]),
    ("three", [3, 4, 5]),
    ("two_files", [16, 17]),
    ("whole_file", list(range(103, 136))),
    ("one", [199]),
])
def test_quickpar_repair_reference_transform(case):
    qp = quickpar()
    _, missing = case
    n = qp.data.shape[0]
    present = [d for d in range(n) if d not in set(missing)]
    m = len(missing)
    exps = qp.exponents[:m]  # lowest exponents first (repair.rs:584-604)
    out, row_order = oracle.reconstruct(n, missing, present, exps,
                                        qp.rec_words[:m], qp.words[present])
    assert list(row_order) == list(range(m)), "contiguous exponents never pivot-swap"
    assert np.array_equal(out, qp.words[missing])


def test_quickpar_repair_with_exponent_gaps():
    # recovery exponents {0,1,3,5}: a genuine choice when volumes are missing
    qp = quickpar()
    n = qp.data.shape[0]
    missing = [10, 50, 90, 150]
    present = [d for d in range(n) if d not in missing]
    exps = [0, 1, 3, 5]
    out, _ = oracle.reconstruct(n, missing, present, exps, qp.rec_words[exps], qp.words[present])
    assert np.array_equal(out, qp.words[missing])


# ------------------------------------------------------------- hashes --
def test_md5_rfc1321_suite():
    kats = {
        b"": "d41d8cd98f00b204e9800998ecf8427e",
        b"a": "0cc175b9c0f1b6a831c399e269772661",
        b"abc": "900150983cd24fb0d6963f7d28e17f72",
        b"message digest": "f96b697d7cb7938d525a2f31aaf161d0",
        b"abcdefghijklmnopqrstuvwxyz": "c3fcd3d76192e4007dfb496cca67e13b",
        b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789": "d174ab98d277d9f5a5611c2c9f419d9f",
        b"1234567890" * 8: "57edf4a22be3c955ac49da2e2107b67a",
    }
    for msg, h in kats.items():
        assert oracle.md5(msg).hex() == h


def test_crc32_kats():
    assert oracle.crc32(b"123456789") == 0xCBF43926
    assert oracle.crc32(b"abc") == 0x352441C2  # src/hash/mod.rs KAT
    assert oracle.crc32(b"") == 0


def test_hash_golden_vectors():
    with open(os.path.join(GOLDEN, "hash_vectors.json")) as f:
        g = json.load(f)
    rng = np.random.default_rng(20261203)
    for v in g["vectors"]:
        b = rng.integers(0, 256, v["seed_len"], dtype=np.uint8).tobytes()
        if v["hex"] is not None:
            assert b.hex() == v["hex"]
        assert oracle.md5(b).hex() == v["md5"]
        assert oracle.crc32(b) == v["crc32"]
        assert hashlib.md5(b).hexdigest() == v["md5"] and zlib.crc32(b) == v["crc32"]
    rng2 = np.random.default_rng(7)
    ss = g["slice_size"]
    buf = np.zeros((len(g["slices"]), ss), np.uint8)
    lens = []
    for i, s in enumerate(g["slices"]):
        buf[i, : s["len"]] = rng2.integers(0, 256, s["len"], dtype=np.uint8)
        lens.append(s["len"])
    md5s, crcs = oracle.hash_slices(buf, len(lens), ss, lens=np.array(lens, np.uintp))
    for i, s in enumerate(g["slices"]):
        assert md5s[i].tobytes().hex() == s["md5"] and int(crcs[i]) == s["crc32"]


def test_hash_random_vs_stdlib():
    rng = np.random.default_rng(11)
    for ln in list(range(0, 200)) + [4095, 4096, 4097, 100003]:
        b = rng.integers(0, 256, ln, dtype=np.uint8).tobytes()
        assert oracle.md5(b) == hashlib.md5(b).digest()
        assert oracle.crc32(b) == zlib.crc32(b)


# -------------------------------------------------------------- repair --
def test_gf_invert_identity_and_singular():
    rng = np.random.default_rng(5)
    for m in (1, 2, 7, 33):
        c = oracle.par2_constants(500)
        cols = rng.choice(500, m, replace=False)
        exps = rng.choice(200, m, replace=False)
        A = np.array([[oracle.gf_pow(int(c[j]), int(e)) for j in cols] for e in exps], np.uint16)
        inv = oracle.gf_invert(A)
        # A * inv == I, computed with the independent carry-less product
        prod = np.zeros((m, m), np.uint16)
        for i in range(m):
            for j in range(m):
                acc = 0
                for k in range(m):
                    acc ^= clmul_mod(int(A[i, k]), int(inv[k, j]))
                prod[i, j] = acc
        assert np.array_equal(prod, np.eye(m, dtype=np.uint16))
    A = np.array([[1, 2, 1], [3, 4, 3], [5, 6, 5]], np.uint16)  # duplicated column
    with pytest.raises(ValueError):
        oracle.gf_invert(A)
    with pytest.raises(ValueError):
        oracle.transform(10, [1, 1], [0, 2, 3], [0, 1])


def test_encode_erase_reconstruct_roundtrip():
    rng = np.random.default_rng(9)
    n, words, r = 64, 517, 6
    data = rng.integers(0, 65536, (n, words)).astype(np.uint16)
    rec = oracle.encode(data, range(r))
    missing = sorted(rng.choice(n, 5, replace=False).tolist())
    present = [d for d in range(n) if d not in missing]
    use = [0, 1, 2, 4, 5]
    out, _ = oracle.reconstruct(n, missing, present, use, rec[use], data[present])
    assert np.array_equal(out, data[missing])


def test_avx2_baseline_matches_scalar():
    if not oracle.have_avx2():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(4)
    for n, words, r in ((3, 16, 2), (17, 1000, 9), (40, 4099, 12), (130, 64, 5)):
        data = rng.integers(0, 65536, (n, words)).astype(np.uint16)
        want = oracle.encode(data, range(r))
        for t in (1, 3):
            assert np.array_equal(oracle.encode_avx2(data, range(r), threads=t), want)
    qp = quickpar()
    assert np.array_equal(oracle.encode_avx2(qp.words, qp.exponents, threads=4), qp.rec_words)
