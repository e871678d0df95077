"""Every tile shape of the LDS contraction (words per thread, rows per
workgroup, sources per LDS chunk) gives the oracle's bytes, and autotuned
runs stay bit-exact."""
from __future__ import annotations

import os
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
    c.set_gf_kernel(1)
    yield c
    c.force_tile_shape(None)
    c.close()


# (n, slice bytes, r): tails of every word multiple, r below / at / above a
# group of 8 and a workgroup of 16 or 32 rows, more sources than one chunk
PROBLEMS = [(5, 64, 1), (37, 4100, 7), (70, 2050, 9), (33, 16386, 17), (40, 32768, 33), (3, 2, 2)]


@pytest.mark.parametrize("shape", par2hip.Context.TILE_SHAPES)
def test_every_tile_shape_matches_oracle(ctx, shape):
    rng = np.random.default_rng(sum(shape))
    ctx.force_tile_shape(shape)
    try:
        for n, s, r in PROBLEMS:
            data = rng.integers(0, 256, (n, s), dtype=np.uint8)
            exps = np.sort(rng.choice(400, r, replace=False)).astype(np.uint32)
            say("shape", shape, "problem", n, s, r)
            got = ctx.encode(data, exps)
            want = oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(r, s)
            assert np.array_equal(got, want), f"shape {shape} problem {(n, s, r)}"
            # repair too: the fold and the rebuild contractions use the same shape
            if r >= 2 and n > r:
                missing = np.sort(rng.choice(n, r, replace=False)).astype(np.uint32)
                out = ctx.reconstruct(data, missing, got, exps)
                assert np.array_equal(out, data[missing]), f"shape {shape} repair {(n, s, r)}"
    finally:
        ctx.force_tile_shape(None)


def test_autotune_stays_exact_and_fills_table(ctx, tmp_path):
    rng = np.random.default_rng(5)
    prev = ctx.set_autotune(True)
    try:
        for n, s, r in [(64, 65536, 6), (200, 8192, 40), (16, 4098, 3), (64, 65536, 6)]:
            data = rng.integers(0, 256, (n, s), dtype=np.uint8)
            exps = np.arange(r, dtype=np.uint32)
            say("autotune", n, s, r)
            got = ctx.encode(data, exps)
            assert np.array_equal(got, oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(r, s))
        path = tmp_path / "tuned.txt"
        par2hip.autotune_save(str(path))
        rows = [l.split() for l in path.read_text().splitlines() if not l.startswith("#")]
        assert len(rows) >= 3
        assert all(tuple(int(x) for x in row[1:]) in par2hip.Context.TILE_SHAPES for row in rows)
        assert par2hip.autotune_load(str(path)) == len(rows)
    finally:
        ctx.set_autotune(prev)


# (n, slice bytes, r) for the matrix-core form's tile shapes: 1-2 outputs
# take the 2-output x 1024-word shape, 3-4 the 4-output x 512-word one (16
# sources per stage), more the 8-output x 512-word one; ragged last tiles,
# sources not a multiple of a stage, one source, and the exact tile sizes
MFMA_PROBLEMS = [(1, 2048, 1), (5, 2048, 2), (9, 2080, 1), (21, 4128, 2), (16, 4096, 3), (17, 2080, 4),
                 (100, 6176, 4), (33, 1040, 3), (40, 3088, 5), (7, 1024, 8), (50, 4112, 9)]


@pytest.mark.parametrize("kernel", [2, 3], ids=["fp4", "int8"])
def test_mfma_tile_shapes_match_oracle(ctx, kernel):
    rng = np.random.default_rng(kernel)
    prev = ctx.set_gf_kernel(kernel)
    try:
        for n, s, r in MFMA_PROBLEMS:
            data = rng.integers(0, 256, (n, s), dtype=np.uint8)
            exps = np.sort(rng.choice(400, r, replace=False)).astype(np.uint32)
            say("mfma", kernel, "problem", n, s, r)
            got = ctx.encode(data, exps)
            want = oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(r, s)
            assert np.array_equal(got, want), f"kernel {kernel} problem {(n, s, r)}"
            if n > r:
                missing = np.sort(rng.choice(n, r, replace=False)).astype(np.uint32)
                out = ctx.reconstruct(data, missing, got, exps)
                assert np.array_equal(out, data[missing]), f"kernel {kernel} repair {(n, s, r)}"
        # accumulating contraction (the strategy's region muladd) with <= 4 rows
        s = par2hip.HipStrategy(ctx)
        srcs = [rng.integers(0, 65536, 3000).astype(np.uint16) for _ in range(19)]
        dsts = [rng.integers(0, 65536, 3000).astype(np.uint16) for _ in range(3)]
        coeff = rng.integers(0, 65536, (3, 19)).astype(np.uint16)
        want = [d.copy() for d in dsts]
        oracle.muladd_region(want, srcs, coeff, 8, 2056)
        s.muladd_region(dsts, srcs, list(coeff), 8, 2056)
        for a, b in zip(dsts, want):
            assert np.array_equal(a, b)
    finally:
        ctx.set_gf_kernel(prev)


# The FP4 ring form over many stages: 600 sources (19-38 stages, the last
# one padded), 64 KiB + 2064 B slices (16-byte rows, as the matrix-core form
# needs: whole tiles in the ring form and a ragged 8-word tile column in the
# staged form), one to eight outputs (single, small and split tiles), every
# byte against the multithreaded AVX2 restatement, then a rebuild of as many
# erased slices from those outputs.
@pytest.mark.parametrize("r", [1, 2, 4, 8])
def test_fp4_ring_many_stages(ctx, r):
    rng = np.random.default_rng(600 + r)
    n, s = 600, 65536 + 2064
    data = rng.integers(0, 256, (n, s), dtype=np.uint8)
    exps = np.sort(rng.choice(2000, r, replace=False)).astype(np.uint32)
    prev = ctx.set_gf_kernel(2)
    try:
        say("fp4 ring", n, s, r)
        got = ctx.encode(data, exps)
        want = oracle.encode_avx2(data.view(np.uint16), exps).view(np.uint8).reshape(r, s)
        assert np.array_equal(got, want), f"fp4 ring r={r}"
        missing = np.sort(rng.choice(n, r, replace=False)).astype(np.uint32)
        out = ctx.reconstruct(data, missing, got, exps)
        assert np.array_equal(out, data[missing]), f"fp4 ring repair r={r}"
    finally:
        ctx.set_gf_kernel(prev)
