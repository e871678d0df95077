"""Batched runs (many sets of uneven sizes, encode and repair mixed, one
persistent contraction launch): every job's bytes equal the single-set path
and the oracle."""
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


def make_jobs(ctx, rng, specs):
    """specs: (kind, n, slice_size, r_or_m, consecutive, hashes)"""
    jobs, host = [], []
    for kind, n, s, k, consec, hashes in specs:
        data = rng.integers(0, 256, (n, s), dtype=np.uint8)
        d_sl = ctx.alloc(n * s)
        job = {"kind": kind, "slices": d_sl, "n": n, "slice_size": s}
        h = {"data": data, "kind": kind}
        if kind == "encode":
            exps = np.arange(k, dtype=np.uint32) if consec else np.sort(rng.choice(200, k, replace=False)).astype(np.uint32)
            d_sl.upload(data)
            job.update(exps=exps, recovery=ctx.alloc(max(k, 1) * s))
            h["exps"] = exps
        else:
            exps = np.arange(k, dtype=np.uint32) + (3 if consec == "shift" else 0)
            if not consec:
                exps = np.sort(rng.choice(60, k, replace=False)).astype(np.uint32)
            rec = oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(k, s)
            missing = np.sort(rng.choice(n, k, replace=False)).astype(np.uint32)
            damaged = data.copy()
            damaged[missing] = 0
            d_sl.upload(damaged)
            d_rec = ctx.alloc(k * s)
            d_rec.upload(rec)
            job.update(exps=exps, recovery=d_rec, missing=missing, out=ctx.alloc(k * s))
            h.update(exps=exps, missing=missing, damaged=damaged)
        if hashes:
            job.update(md5=ctx.alloc(n * 16), crc=ctx.alloc(n * 4))
        jobs.append(job)
        host.append(h)
    return jobs, host


def check(ctx, jobs, host):
    for i, (j, h) in enumerate(zip(jobs, host)):
        n, s = j["n"], j["slice_size"]
        if j["kind"] == "encode":
            r = len(h["exps"])
            got = j["recovery"].download(r * s).reshape(r, s)
            want = oracle.encode(h["data"].view(np.uint16), h["exps"]).view(np.uint8).reshape(r, s)
            assert np.array_equal(got, want), f"job {i} encode"
            hashed = h["data"]
        else:
            m = len(h["missing"])
            got = j["out"].download(m * s).reshape(m, s)
            assert np.array_equal(got, h["data"][h["missing"]]), f"job {i} repair"
            hashed = h["damaged"]
        if "md5" in j:
            md5 = j["md5"].download().reshape(n, 16)
            crc = j["crc"].download().view(np.uint32)
            wm, wc = oracle.hash_slices(hashed, n, s)
            assert np.array_equal(md5, wm) and np.array_equal(crc, wc), f"job {i} hashes"


def test_batch_mixed_uneven(ctx):
    rng = np.random.default_rng(2026)
    specs = [
        ("encode", 300, 4096, 30, True, True),
        ("encode", 7, 5376, 3, True, True),
        ("encode", 1, 2, 1, True, True),        # one word
        ("encode", 40, 1026, 17, False, True),  # 2-byte aligned rows, arbitrary exponents
        ("encode", 129, 65540, 9, True, True),   # 4-byte aligned rows, workgroup-per-slice CRC
        ("repair", 6, (1 << 18) + 20, 3, True, True),
        ("repair", 200, 2048, 20, True, True),
        ("repair", 60, 4100, 5, "shift", True),  # consecutive exponents from 3
        ("repair", 33, 18, 4, False, True),      # gapped exponents: Gauss-Jordan path
        ("repair", 5, 512, 5, True, False),      # every data slice lost
        ("encode", 64, 8192, 0, True, True),     # hashes only
    ]
    jobs, host = make_jobs(ctx, rng, specs)
    say("batch mixed", len(jobs))
    ctx.batch_run_dev(jobs)
    ctx.sync()
    check(ctx, jobs, host)


def test_batch_matches_single_set_path(ctx):
    rng = np.random.default_rng(7)
    specs = [("encode", int(rng.integers(8, 400)), int(rng.integers(2, 3000)) * 4, int(rng.integers(1, 40)), True, True)
             for _ in range(24)]
    jobs, host = make_jobs(ctx, rng, specs)
    say("batch 24 encodes")
    ctx.batch_run_dev(jobs)
    ctx.sync()
    for j, h in zip(jobs, host):
        n, s, r = j["n"], j["slice_size"], len(h["exps"])
        single = ctx.encode(h["data"], h["exps"])
        assert np.array_equal(j["recovery"].download(r * s).reshape(r, s), single)


def test_batch_repeat_and_empty(ctx):
    rng = np.random.default_rng(9)
    jobs, host = make_jobs(ctx, rng, [("encode", 50, 2048, 6, True, True), ("repair", 50, 2048, 6, True, True)])
    say("batch repeat")
    for _ in range(3):
        ctx.batch_run_dev(jobs)
    ctx.sync()
    check(ctx, jobs, host)
    ctx.batch_run_dev([])
    with pytest.raises(par2hip.InvalidFormat):
        bad = dict(jobs[0])
        bad["slice_size"] = 7
        ctx.batch_run_dev([bad])


def clear_outputs(jobs):
    for j in jobs:
        buf = j["recovery"] if j["kind"] == "encode" else j["out"]
        buf.upload(np.zeros(buf.nbytes, np.uint8))
        for k in ("md5", "crc"):
            if k in j:
                j[k].upload(np.zeros(j[k].nbytes, np.uint8))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_plan_replays(ctx, graph):
    """A planned batch launched three times (outputs cleared and unrelated
    work on the same context between launches) gives the oracle's bytes
    every time, and follows new input contents in the same buffers."""
    rng = np.random.default_rng(41 + int(graph))
    specs = [
        ("encode", 120, 4096, 12, True, True),
        ("encode", 9, 5376, 3, False, True),
        ("repair", 80, 2052, 6, True, True),
        ("repair", 40, 4100, 4, "shift", True),
        ("encode", 70, 65540, 5, True, True),
        ("encode", 16, 8192, 0, True, True),  # hashes only
    ]
    jobs, host = make_jobs(ctx, rng, specs)
    say("plan", "graph" if graph else "eager")
    plan = ctx.batch_plan(jobs, graph=graph)
    try:
        side = rng.integers(0, 256, (20, 2048), dtype=np.uint8)
        for rep in range(3):
            clear_outputs(jobs)
            other = ctx.encode(side, np.arange(3, dtype=np.uint32))  # eager work between replays
            assert np.array_equal(other, oracle.encode(side.view(np.uint16), np.arange(3)).view(np.uint8).reshape(3, -1))
            plan.launch()
            ctx.sync()
            check(ctx, jobs, host)
        # new contents in the encode jobs' input buffers
        for j, h in zip(jobs, host):
            if j["kind"] == "encode":
                h["data"] = rng.integers(0, 256, h["data"].shape, dtype=np.uint8)
                j["slices"].upload(h["data"])
        clear_outputs(jobs)
        plan.launch()
        ctx.sync()
        check(ctx, jobs, host)
    finally:
        plan.close()


def test_batch_plan_rejects_gapped_repair(ctx):
    rng = np.random.default_rng(43)
    jobs, _ = make_jobs(ctx, rng, [("encode", 20, 2048, 4, True, False), ("repair", 33, 18, 4, False, True)])
    with pytest.raises(par2hip.InvalidFormat):
        ctx.batch_plan(jobs, graph=True)
    # the context still works afterwards
    ctx.batch_run_dev(jobs)
    ctx.sync()


SHAPES = [(8, 2, 16), (8, 1, 16), (4, 2, 16), (4, 1, 16), (4, 4, 16), (8, 2, 32), (2, 2, 16), (1, 1, 16)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "v%d_g%d_d%d" % s)
def test_batch_every_tile_shape(ctx, shape):
    """Each tile shape the batched contraction runs (v in {8, 4} x g in {1, 2};
    g = 4 and 32-source shapes map to g = 2, narrower v to the default), forced
    on a mixed batch (ragged tiles, 2-, 4- and 16-byte aligned rows, folds and
    main contractions), eager and as a planned graph: the oracle's bytes."""
    rng = np.random.default_rng(sum(shape))
    specs = [
        ("encode", 150, 4096 + 64, 20, True, True),
        ("encode", 37, 1026, 9, False, True),    # 2-byte aligned rows
        ("encode", 70, 8196, 17, True, True),    # 4-byte aligned rows
        ("repair", 90, 6144, 12, True, True),
        ("repair", 12, 2048, 12, True, True),    # every data slice lost: recovery rows only
        ("repair", 200, 3000, 9, "shift", True),
    ]
    jobs, host = make_jobs(ctx, rng, specs)
    ctx.force_tile_shape(shape)
    try:
        say("batch shape", shape)
        ctx.batch_run_dev(jobs)
        ctx.sync()
        check(ctx, jobs, host)
        clear_outputs(jobs)
        plan = ctx.batch_plan(jobs, graph=True)
        try:
            plan.launch()
            ctx.sync()
            check(ctx, jobs, host)
        finally:
            plan.close()
    finally:
        ctx.force_tile_shape(None)


def test_batch_uses_tuned_shapes(ctx, tmp_path):
    """With autotuning on, a batched task takes the tuned table's shape for its
    class (loaded from text here), and the bytes stay the oracle's."""
    rng = np.random.default_rng(77)
    specs = [("encode", 300, 16384, 24, True, True), ("repair", 120, 8192, 10, True, True)]
    jobs, host = make_jobs(ctx, rng, specs)
    # key = vmax | row groups << 8 | bit_length(nsrc) << 16 | bit_length(words) << 24
    keys = [8 | 3 << 8 | (300).bit_length() << 16 | (8192).bit_length() << 24,
            8 | 2 << 8 | (120).bit_length() << 16 | (4096).bit_length() << 24]
    table = tmp_path / "tuned.txt"
    table.write_text("".join("%x 4 1 16\n" % k for k in keys))
    assert par2hip.autotune_load(str(table)) == 2
    was = ctx.set_autotune(True)
    try:
        say("batch tuned shapes")
        ctx.batch_run_dev(jobs)
        ctx.sync()
        check(ctx, jobs, host)
    finally:
        ctx.set_autotune(was)


def test_batch_plan_autotuned_policies(ctx):
    """With autotuning on, the plan times every tile-shape policy on the batch
    itself and keeps the fastest; each policy's plan gives the oracle's bytes."""
    rng = np.random.default_rng(78)
    specs = [("encode", 200, 8192, 30, True, True), ("encode", 40, 4096, 5, True, True),
             ("repair", 150, 16384, 12, True, True), ("repair", 30, 2048, 4, "shift", True)]
    jobs, host = make_jobs(ctx, rng, specs)
    was = ctx.set_autotune(True)
    try:
        say("plan autotuned")
        plan = ctx.batch_plan(jobs, graph=True)
        try:
            t = plan.tuning()
            assert t["policies"] == 4 and 0 <= t["policy"] < 4 and all(ms > 0 for ms in t["ms"])
            clear_outputs(jobs)
            plan.launch()
            ctx.sync()
            check(ctx, jobs, host)
        finally:
            plan.close()
    finally:
        ctx.set_autotune(was)
    plan = ctx.batch_plan(jobs, graph=False)  # untuned: no policies timed
    try:
        assert plan.tuning()["policies"] == 0
    finally:
        plan.close()
