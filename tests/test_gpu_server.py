"""The serving loop: requests admitted by HBM budget and run in continuous
batches give, for every request, the same bytes as the request run alone and
as the single-set path (and the oracle)."""
from __future__ import annotations

import sys
import threading

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


def make_requests(ctx, rng, count, max_n=300, max_s=12288):
    """Random encode/repair requests with device-resident inputs; host copies
    of what each request must produce."""
    reqs = []
    for i in range(count):
        n = int(rng.integers(4, max_n))
        s = int(rng.integers(1, max_s // 4)) * 4
        data = rng.integers(0, 256, (n, s), dtype=np.uint8)
        d_sl = ctx.alloc(n * s)
        if i % 3 == 2:  # repair: erase m slices, recovery rows 0..m-1
            m = int(rng.integers(1, max(2, n // 8) + 1))
            exps = np.arange(m, dtype=np.uint32)
            rec = oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(m, s)
            missing = np.sort(rng.choice(n, m, replace=False)).astype(np.uint32)
            damaged = data.copy()
            damaged[missing] = rng.integers(0, 256, (m, s), dtype=np.uint8)
            d_sl.upload(damaged)
            d_rec = ctx.alloc(m * s)
            d_rec.upload(rec)
            job = {"kind": "repair", "slices": d_sl, "n": n, "slice_size": s, "exps": exps,
                   "recovery": d_rec, "missing": missing}
            want = data[missing]
            hashed = damaged
        else:
            d_sl.upload(data)
            r = int(rng.integers(0, 24))
            exps = np.arange(r, dtype=np.uint32) if i % 2 else np.sort(rng.choice(300, r, replace=False)).astype(np.uint32)
            job = {"kind": "encode", "slices": d_sl, "n": n, "slice_size": s, "exps": exps}
            want = oracle.encode(data.view(np.uint16), exps).view(np.uint8).reshape(r, s) if r else np.zeros((0, s), np.uint8)
            hashed = data
        reqs.append((job, want, hashed))
    return reqs


def fetch(ctx, res, job):
    n, s = job["n"], job["slice_size"]
    rows = len(job["missing"]) if job["kind"] == "repair" else len(job["exps"])
    out = par2hip.d2h(res["d_out"], rows * s, ctx).reshape(rows, s) if rows else np.zeros((0, s), np.uint8)
    md5 = par2hip.d2h(res["d_md5"], n * 16, ctx).reshape(n, 16)
    crc = par2hip.d2h(res["d_crc"], n * 4, ctx).view(np.uint32)
    return out, md5, crc


def test_server_mix_equals_alone_and_oracle(ctx):
    rng = np.random.default_rng(11)
    reqs = make_requests(ctx, rng, 36)
    srv = par2hip.Server(0, hbm_budget=1 << 30, lanes=2, max_batch=64)
    try:
        # submit from another thread while the first batches already run
        tickets = [None] * len(reqs)

        def submitter():
            for i, (job, _, _) in enumerate(reqs):
                tickets[i] = srv.submit(job)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        say("server mix", len(reqs))
        mixed = []
        for i, (job, want, hashed) in enumerate(reqs):
            res = srv.wait(tickets[i])
            out, md5, crc = fetch(ctx, res, job)
            assert np.array_equal(out, want), f"request {i} ({job['kind']}) bytes"
            wm, wc = oracle.hash_slices(hashed, job["n"], job["slice_size"])
            assert np.array_equal(md5, wm) and np.array_equal(crc, wc), f"request {i} hashes"
            mixed.append((out, md5, crc, res["batch_size"]))
            srv.release(tickets[i])
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["charged"] == 0
        say("batches", st["batches"], "largest", st["max_batch_seen"])
        # the same requests one at a time: identical bytes
        for i, (job, _, _) in enumerate(reqs[:12]):
            t = srv.submit(job)
            res = srv.wait(t)
            assert res["batch_size"] == 1
            alone = fetch(ctx, res, job)
            for a, b in zip(alone, mixed[i][:3]):
                assert np.array_equal(a, b), f"request {i} alone vs in the mix"
            srv.release(t)
    finally:
        srv.close()


def test_server_admission_by_budget(ctx):
    """A budget that holds about two requests: the rest wait for memory, all
    finish once results are released, and the budget is never exceeded."""
    rng = np.random.default_rng(12)
    n, s = 64, 65536
    jobs = []
    for i in range(8):
        data = rng.integers(0, 256, (n, s), dtype=np.uint8)
        d = ctx.alloc(n * s)
        d.upload(data)
        jobs.append(({"kind": "encode", "slices": d, "n": n, "slice_size": s,
                      "exps": np.arange(8, dtype=np.uint32)}, data))
    budget = 2 * (n * s + 8 * s + n * 20 + 8 * n * 2 + n * 64 * 16) + 4096
    srv = par2hip.Server(0, hbm_budget=budget, lanes=2, max_batch=8)
    try:
        tickets = [srv.submit(j) for j, _ in jobs]
        st = srv.stats()
        assert st["pending"] >= 5, st  # only two fit at once
        say("admission", st)
        for (job, data), t in zip(jobs, tickets):
            res = srv.wait(t)
            out = par2hip.d2h(res["d_out"], 8 * s, ctx).reshape(8, s)
            assert np.array_equal(out, ctx.encode(data, np.arange(8, dtype=np.uint32)))
            srv.release(t)
        st = srv.stats()
        assert st["peak_charged"] <= budget and st["completed"] == 8 and st["charged"] == 0
    finally:
        srv.close()


def test_server_rejects_bad_requests(ctx):
    srv = par2hip.Server(0, hbm_budget=1 << 20, lanes=1, max_batch=4)
    try:
        d = ctx.alloc(4096)
        with pytest.raises(par2hip.InvalidFormat):
            srv.submit({"kind": "encode", "slices": d, "n": 2, "slice_size": 7, "exps": [0]})
        with pytest.raises(par2hip.InvalidFormat):
            srv.submit({"kind": "repair", "slices": d, "n": 2, "slice_size": 2048, "exps": [0, 1],
                        "recovery": d, "missing": [1, 1]})
        big = ctx.alloc(1 << 21)
        with pytest.raises(par2hip.DeviceError):  # footprint beyond the whole budget
            srv.submit({"kind": "encode", "slices": big, "n": 32, "slice_size": 1 << 16, "exps": [0]})
        assert srv.stats()["submitted"] == 0
    finally:
        srv.close()


def test_server_md5_slice_limit(ctx):
    """Hashed slices of 128 GiB or more are refused at submission (the MD5
    kernel's block counter and bit length stay exact below); one just under
    the limit passes that check and is refused only because it can never fit
    the 1 MiB budget.  Every case is refused before anything is launched."""
    srv = par2hip.Server(0, hbm_budget=1 << 20, lanes=1, max_batch=4)
    try:
        d = ctx.alloc(4096)
        limit = 1 << 37
        with pytest.raises(par2hip.InvalidFormat):
            srv.submit({"kind": "encode", "slices": d, "n": 1, "slice_size": limit, "exps": []}, hashes=True)
        with pytest.raises(par2hip.DeviceError):  # unhashed: only the budget refuses it
            srv.submit({"kind": "encode", "slices": d, "n": 1, "slice_size": limit, "exps": []}, hashes=False)
        with pytest.raises(par2hip.DeviceError):  # just under the limit: the budget refuses it
            srv.submit({"kind": "encode", "slices": d, "n": 1, "slice_size": limit - 2, "exps": []}, hashes=True)
        assert srv.stats()["submitted"] == 0
    finally:
        srv.close()


def test_server_wait_timeout_on_budget_blocked_ticket(ctx):
    """FIFO admission: with the budget held by an unreleased result, a later
    request cannot be admitted; waiting on it with a limit times out instead
    of hanging, the request keeps its place, and it runs once the first
    result is released."""
    rng = np.random.default_rng(13)
    n, s = 64, 65536
    data = rng.integers(0, 256, (2, n, s), dtype=np.uint8)
    bufs = []
    for k in range(2):
        d = ctx.alloc(n * s)
        d.upload(data[k])
        bufs.append(d)
    exps = np.arange(8, dtype=np.uint32)
    one = n * s + 8 * s + n * 20 + 8 * n * 2 + n * 64 * 16
    srv = par2hip.Server(0, hbm_budget=one + 4096, lanes=1, max_batch=4)
    try:
        t0 = srv.submit({"kind": "encode", "slices": bufs[0], "n": n, "slice_size": s, "exps": exps})
        t1 = srv.submit({"kind": "encode", "slices": bufs[1], "n": n, "slice_size": s, "exps": exps})
        res0 = srv.wait(t0, timeout_ms=60000)  # done: the first one fits
        say("timeout: second request pending", srv.stats()["pending"])
        assert srv.stats()["pending"] == 1
        with pytest.raises(TimeoutError):
            srv.wait(t1, timeout_ms=200)
        with pytest.raises(TimeoutError):
            srv.wait(t1, timeout_ms=0)
        out0 = par2hip.d2h(res0["d_out"], 8 * s, ctx).reshape(8, s)
        srv.release(t0)
        res1 = srv.wait(t1, timeout_ms=60000)
        out1 = par2hip.d2h(res1["d_out"], 8 * s, ctx).reshape(8, s)
        for k, out in enumerate((out0, out1)):
            want = oracle.encode(data[k].view(np.uint16), exps).view(np.uint8).reshape(8, s)
            assert np.array_equal(out, want)
        srv.release(t1)
    finally:
        srv.close()
