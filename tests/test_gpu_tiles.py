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
