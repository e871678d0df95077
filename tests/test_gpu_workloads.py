"""The batched and served configurations at their full sizes.

Batched: the 256-set workload (log-uniform 8..4096 slices of 4 KiB..1 MiB,
redundancy 1..20 %, half repairs with 1..10 % erasures; ~25 GiB of source as
windows of a 1 GiB seeded pool) planned once and replayed as a HIP graph.
Every encode is checked against the oracle's AVX2 restatement on a 1 KiB
column window of every recovery slice, every repair rebuilds the original
bytes of every erased slice (the erased rows hold garbage), and the IFSC
digests of sampled slices equal hashlib/zlib.

Served: 200 requests of the served workload (lognormal sizes, median 64 MiB,
30 % repairs) submitted with Poisson gaps to a server with a 200 GB budget;
every result equals the plain single-set path on the same inputs, the plain
path equals the oracle on a column window, and every repair gives back the
original slices.
"""
from __future__ import annotations

import hashlib
import os
import sys
import threading
import time
import zlib

import numpy as np
import pytest

import oracle
import par2hip
import workloads

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


class PoolSet:
    """A set whose bytes are a window of a host pool: bytes [off, off + nb)
    when the set fits the pool, else the pool repeated from its start."""

    def __init__(self, pool, nb, off):
        self.pool, self.nb = pool, nb
        self.off = off if nb <= len(pool) else 0
        self.wrap = nb > len(pool)

    def bytes(self, a, b):
        if not self.wrap:
            return self.pool[self.off + a:self.off + b]
        L = len(self.pool)
        idx = np.arange(a, b) % L
        return self.pool[idx]

    def upload(self, d):
        if not self.wrap:
            d.upload(self.pool[self.off:self.off + self.nb])
        else:
            for o in range(0, self.nb, len(self.pool)):
                d.upload(self.pool[:min(len(self.pool), self.nb - o)], offset=o)

    def columns(self, n, s, w0, nw):
        """(n, nw) u16 words [w0, w0 + nw) of every slice."""
        return np.stack([self.bytes(d * s + 2 * w0, d * s + 2 * (w0 + nw)).view(np.uint16) for d in range(n)])


def test_batched_256_sets_full_size(ctx):
    sets, total = workloads.draw_sets(4)
    assert len(sets) == 256 and total >= 24 << 30
    pool = workloads.random_bytes(5, 1 << 30)
    rng = np.random.default_rng(6)
    say("batched workload", len(sets), "sets", round(total / 2**30, 2), "GiB")
    jobs, meta = [], []
    for x in sets:
        n, s = x["n"], x["s"]
        nb = n * s
        src = PoolSet(pool, nb, int(rng.integers(0, len(pool) - min(nb, len(pool)) + 1)))
        d_sl = ctx.alloc(nb)
        src.upload(d_sl)
        job = {"kind": x["kind"], "slices": d_sl, "n": n, "slice_size": s,
               "md5": ctx.alloc(n * 16), "crc": ctx.alloc(n * 4)}
        w0 = int(rng.integers(0, s // 2 - 512 + 1))
        m = {"src": src, "w0": w0}
        if x["kind"] == "encode":
            job.update(exps=np.arange(x["r"], dtype=np.uint32), recovery=ctx.alloc(x["r"] * s))
        else:
            k = x["m"]
            d_rec = ctx.alloc(k * s)
            ctx.encode_dev(d_sl, n, s, np.arange(k, dtype=np.uint32), d_rec)
            missing = np.sort(rng.choice(n, k, replace=False)).astype(np.uint32)
            ctx.sync()
            garbage = rng.integers(0, 256, s, dtype=np.uint8)
            for i in missing:  # the erased rows hold garbage: the rebuild must not read them
                d_sl.upload(garbage, offset=int(i) * s)
            job.update(exps=np.arange(k, dtype=np.uint32), recovery=d_rec, missing=missing, out=ctx.alloc(k * s))
            m.update(garbage=garbage)
        jobs.append(job)
        meta.append(m)
    ctx.sync()
    say("plan + graph replay")
    plan = ctx.batch_plan(jobs, graph=True)
    for rep in range(2):  # outputs cleared between replays: each replay writes everything
        for j in jobs:
            (j["recovery"] if j["kind"] == "encode" else j["out"]).upload(np.zeros(1024, np.uint8))
        plan.launch()
        ctx.sync()
    plan.close()
    checked_rows = 0
    for j, x, m in zip(jobs, sets, meta):
        n, s, src, w0 = j["n"], j["slice_size"], m["src"], m["w0"]
        cols = src.columns(n, s, w0, 512)
        if j["kind"] == "encode":
            want = oracle.encode_avx2(cols, j["exps"], threads=THREADS)
            for e in range(len(j["exps"])):
                got = j["recovery"].download(1024, offset=e * s + 2 * w0).view(np.uint16)
                assert np.array_equal(got, want[e]), f"set n={n} s={s}: recovery {e}"
                checked_rows += 1
        else:
            k = len(j["missing"])
            want_rec = oracle.encode_avx2(cols, np.arange(k), threads=THREADS)
            for e in range(k):
                got = j["recovery"].download(1024, offset=e * s + 2 * w0).view(np.uint16)
                assert np.array_equal(got, want_rec[e]), f"set n={n} s={s}: recovery input {e}"
            out = j["out"].download().reshape(k, s)
            for row, i in enumerate(j["missing"]):
                assert np.array_equal(out[row], src.bytes(int(i) * s, int(i + 1) * s)), f"set n={n}: slice {i}"
        md5 = j["md5"].download().reshape(n, 16)
        crc = j["crc"].download().view(np.uint32)
        miss = set(int(i) for i in j.get("missing", []))
        for i in sorted({0, n // 2, n - 1}):
            b = m["garbage"].tobytes() if i in miss else src.bytes(i * s, (i + 1) * s).tobytes()
            assert md5[i].tobytes() == hashlib.md5(b).digest() and int(crc[i]) == zlib.crc32(b), f"n={n}: hash {i}"
    say("checked", checked_rows, "recovery rows,", sum(len(j.get("missing", [])) for j in jobs), "rebuilt slices")
    for j in jobs:
        for key in ("slices", "md5", "crc", "recovery", "out"):
            if key in j:
                j[key].free()


def test_served_200_requests_full_size(ctx):
    reqs = workloads.draw_requests(4, 200)
    pool_bytes = 16 << 30
    d_pool = ctx.alloc(pool_bytes)
    block = workloads.random_bytes(21, 1 << 30)
    for o in range(0, pool_bytes, len(block)):
        d_pool.upload(block, offset=o)
    rng = np.random.default_rng(104)
    total = 0
    rec_bufs = []
    for q in reqs:
        n, s, k = q["n"], q["s"], q["k"]
        nb = n * s
        total += nb
        q["off"] = int(rng.integers(0, (pool_bytes - nb) // 256)) * 256
        q["exps"] = np.arange(k, dtype=np.uint32)
        q["w0"] = int(rng.integers(0, s // 2 - 64 + 1))
        if q["kind"] == "repair":
            # recovery rows of the intact window; the erased rows are then
            # rebuilt from the others and must equal the window's own bytes
            d_rec = ctx.alloc(k * s)
            ctx.encode_dev(d_pool.view(q["off"]), n, s, q["exps"], d_rec)
            q["rec"] = d_rec
            rec_bufs.append(d_rec)
            q["missing"] = np.sort(rng.choice(n, k, replace=False)).astype(np.uint32)
    ctx.sync()
    say("served workload", len(reqs), "requests", round(total / 2**30, 2), "GiB")

    def job_of(q):
        j = {"kind": q["kind"], "slices": d_pool.view(q["off"]), "n": q["n"], "slice_size": q["s"],
             "exps": q["exps"]}
        if q["kind"] == "repair":
            j.update(recovery=q["rec"], missing=q["missing"])
        return j

    srv = par2hip.Server(0, hbm_budget=200 * 10**9, lanes=2, max_batch=512)
    try:
        tickets = [None] * len(reqs)

        def submitter():
            t0 = time.perf_counter()
            for i, q in enumerate(reqs):
                arr = t0 + q["t"] / 2000.0  # 2000 requests/s offered
                now = time.perf_counter()
                if now < arr:
                    time.sleep(arr - now)
                tickets[i] = srv.submit(job_of(q), hashes=True)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        results = []
        for i, q in enumerate(reqs):
            res = srv.wait(tickets[i], timeout_ms=120000)
            rows = q["k"] * q["s"]
            results.append((par2hip.d2h(res["d_out"], rows, ctx), par2hip.d2h(res["d_md5"], q["n"] * 16, ctx),
                            par2hip.d2h(res["d_crc"], q["n"] * 4, ctx), res["batch_size"]))
            srv.release(tickets[i])
        st = srv.stats()
        say("server", st)
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["charged"] == 0
        assert st["max_batch_seen"] > 1, "no request was batched with another"
    finally:
        srv.close()
    # the plain single-set path on the same inputs, and the oracle beside it
    max_out = max(q["k"] * q["s"] for q in reqs)
    max_n = max(q["n"] for q in reqs)
    d_out, d_md5, d_crc = ctx.alloc(max_out), ctx.alloc(max_n * 16), ctx.alloc(max_n * 4)
    blk = np.frombuffer(block, np.uint8)
    for i, (q, (out, md5, crc, _)) in enumerate(zip(reqs, results)):
        n, s, k = q["n"], q["s"], q["k"]
        v = d_pool.view(q["off"])
        if q["kind"] == "encode":
            ctx.encode_dev(v, n, s, q["exps"], d_out, d_md5, d_crc)
        else:
            ctx.hash_slices_dev(v, n, s, d_md5, d_crc)
            ctx.reconstruct_dev(v, n, s, q["missing"], q["rec"], q["exps"], d_out)
        ctx.sync()
        assert np.array_equal(out, d_out.download(k * s)), f"request {i} ({q['kind']}): output vs plain path"
        assert np.array_equal(md5, d_md5.download(n * 16)) and np.array_equal(crc, d_crc.download(n * 4)), i
        # the window's bytes: the pool is the 1 GiB block repeated

        def window_bytes(a, b, off=q["off"]):
            return blk[(np.arange(off + a, off + b)) % len(blk)]

        if q["kind"] == "encode":
            cols = np.stack([window_bytes(d * s + 2 * q["w0"], d * s + 2 * q["w0"] + 128).view(np.uint16)
                             for d in range(n)])
            want = oracle.encode_avx2(cols, q["exps"], threads=THREADS)
            got = out.reshape(k, s)[:, 2 * q["w0"]:2 * q["w0"] + 128].copy().view(np.uint16)
            assert np.array_equal(got, want), f"request {i}: recovery window vs oracle"
        else:
            o = out.reshape(k, s)
            for row in (0, k - 1):
                idx = int(q["missing"][row])
                assert np.array_equal(o[row], window_bytes(idx * s, (idx + 1) * s)), f"request {i}: rebuilt {idx}"
        j = n // 2
        b = window_bytes(j * s, (j + 1) * s).tobytes()
        assert md5.reshape(n, 16)[j].tobytes() == hashlib.md5(b).digest(), i
        assert int(crc.view(np.uint32)[j]) == zlib.crc32(b), i
    for b in rec_bufs + [d_out, d_md5, d_crc, d_pool]:
        b.free()
