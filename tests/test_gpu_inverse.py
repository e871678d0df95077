"""The device inverse at large erasure counts: the blocked multi-workgroup
Gauss-Jordan (panels of 64 pivots, updates as region contractions) for gapped
exponents, and the closed-form inverse up to 1024 consecutive exponents
(compute_reconstruction_transform, src/par2_rs.rs:319-422, has no size limit
below the 65535-block field limit).

m = 257 and 1000: equal to the oracle's Gauss-Jordan.  m = 5000: A * A^-1 = I
on sampled rows and columns with an independent carry-less product (no table
lookups), then a rebuild of 5000 erased slices out of 6000 from gapped
exponents equals the original bytes.  Singular matrices are reported at the
blocked sizes too."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle
import par2hip

pytestmark = pytest.mark.gpu
POLY = 0x1100B
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


def clmul_mod_np(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Carry-less product mod x^16+x^12+x^3+x+1, elementwise (no tables)."""
    a = a.astype(np.uint32)
    b = b.astype(np.uint32)
    p = np.zeros(np.broadcast(a, b).shape, np.uint32)
    for i in range(16):
        p ^= np.where((b >> i) & 1, a << i, 0).astype(np.uint32)
    for bit in range(30, 15, -1):
        p ^= np.where((p >> bit) & 1, np.uint32(POLY << (bit - 16)), 0).astype(np.uint32)
    return p.astype(np.uint16)


def erasure_matrix(m, n_total, rng, gapped=True):
    """A[i][j] = c_{miss_j}^{e_i} (par2_rs.rs:347-359) for m distinct erased
    slices out of n_total and m distinct exponents, from the constants' logs."""
    log, exp = oracle.tables()
    c = oracle.par2_constants(n_total)
    cols = rng.choice(n_total, m, replace=False)
    exps = (np.sort(rng.choice(3 * m + 100, m, replace=False)) if gapped else np.arange(m)).astype(np.uint64)
    lc = log[c[cols]].astype(np.uint64)
    A = exp[(exps[:, None] * lc[None, :]) % 65535]
    return A.astype(np.uint16), cols, exps


def check_identity(A, inv, rows, cols):
    m = A.shape[0]
    for i in rows:  # row i of A * A^-1
        acc = np.zeros(m, np.uint16)
        for j0 in range(0, m, 512):
            blk = clmul_mod_np(A[i, j0:j0 + 512][:, None], inv[j0:j0 + 512, :])
            acc ^= np.bitwise_xor.reduce(blk, axis=0)
        want = np.zeros(m, np.uint16)
        want[i] = 1
        assert np.array_equal(acc, want), f"row {i} of A * A^-1"
    for j in cols:  # column j of A^-1 * A
        acc = np.zeros(m, np.uint16)
        for k0 in range(0, m, 512):
            blk = clmul_mod_np(inv[:, k0:k0 + 512], A[k0:k0 + 512, j][None, :])
            acc ^= np.bitwise_xor.reduce(blk, axis=1)
        want = np.zeros(m, np.uint16)
        want[j] = 1
        assert np.array_equal(acc, want), f"column {j} of A^-1 * A"


@pytest.mark.parametrize("m", [257, 1000])
def test_blocked_inverse_equals_oracle(ctx, m):
    rng = np.random.default_rng(m)
    A, _, _ = erasure_matrix(m, 4000, rng)
    say("blocked gf_invert m", m)
    inv = ctx.gf_invert(A)
    assert np.array_equal(inv, oracle.gf_invert(A))


def test_blocked_inverse_m5000_identity(ctx):
    rng = np.random.default_rng(5000)
    A, _, _ = erasure_matrix(5000, 20000, rng)
    say("blocked gf_invert m", 5000)
    inv = ctx.gf_invert(A)
    check_identity(A, inv, rows=[0, 2345, 4999], cols=[64, 4999])


@pytest.mark.parametrize("m", [300, 1000])
def test_blocked_inverse_singular(ctx, m):
    rng = np.random.default_rng(m + 1)
    A, _, _ = erasure_matrix(m, 4000, rng)
    A[:, m // 3] = A[:, 5]  # two equal columns
    say("blocked singular m", m)
    with pytest.raises(par2hip.RepairFailed):
        ctx.gf_invert(A)
    B, _, _ = erasure_matrix(m, 4000, rng)
    B[m - 7] = B[2]  # two equal rows, found in a later panel
    with pytest.raises(par2hip.RepairFailed):
        ctx.gf_invert(B)


@pytest.mark.parametrize("m", [600, 1024])
def test_closed_form_inverse_large_m(ctx, m):
    """Consecutive exponents 0..m-1 up to the closed form's 1024: the rebuild
    (which takes the closed form) gives back the erased slices."""
    rng = np.random.default_rng(m + 2)
    n, s = m + 200, 64
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    rec = ctx.encode(data, np.arange(m))
    missing = np.sort(rng.choice(n, m, replace=False))
    damaged = data.copy()
    damaged[missing] = 0x5A
    say("closed form rebuild m", m)
    out = ctx.reconstruct(damaged, missing, rec, np.arange(m))
    assert np.array_equal(out, data[missing])


def test_rebuild_5000_of_6000_gapped(ctx):
    rng = np.random.default_rng(6000)
    n, s, m = 6000, 512, 5000
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    exps = np.sort(rng.choice(20000, m, replace=False)).astype(np.uint32)
    say("encode", n, s, "recovery", m, "gapped")
    rec = ctx.encode(data, exps)
    w = np.ascontiguousarray(data.view(np.uint16)[:, :16])
    assert np.array_equal(rec.view(np.uint16)[:, :16], oracle.encode_avx2(w, exps, threads=THREADS))
    missing = np.sort(rng.choice(n, m, replace=False))
    damaged = data.copy()
    damaged[missing] = rng.integers(0, 256, (m, s), dtype=np.uint8)
    say("rebuild", m, "of", n)
    out = ctx.reconstruct(damaged, missing, rec, exps)
    bad = np.nonzero(~np.all(out == data[missing], axis=1))[0]
    assert bad.size == 0, f"{bad.size} rebuilt slices differ, first {missing[bad[:5]]}"
