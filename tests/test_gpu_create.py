"""PAR2 creation over the GPU entry points: every file par2hip_create_par2
writes equals, byte for byte, what the oracle's restatement of
Par2Creator::create writes (and the committed C1 digests); the C1 set then
repairs its damaged input on the GPU."""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import pytest

import c1
import oracle
import par2fmt
import par2hip

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


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


def gpu_vs_oracle(ctx, tmp_path, files, **kw):
    stem = str(tmp_path / kw.pop("stem", "set"))
    scheme = kw.pop("scheme", par2hip.VOLUMES_EXPONENTIAL)
    written = ctx.create_par2(files, stem, scheme=scheme, **kw)
    want = oracle.par2_create(files, stem=stem, scheme=scheme, **kw)
    assert [p for p, _ in written] == [n for n, _ in want]
    for (path, size), (_, ref) in zip(written, want):
        got = open(path, "rb").read()
        assert size == len(got) == len(ref), path
        assert got == ref, f"{path}: first difference at byte {next(i for i in range(len(ref)) if got[i] != ref[i])}"
    return written


def test_c1_bytes_equal_oracle_and_digests(ctx, tmp_path):
    data = c1.c1_input()
    say("c1 create")
    written = gpu_vs_oracle(ctx, tmp_path, [(c1.C1_NAME, data)], block_size=c1.C1_BLOCK,
                            redundancy=c1.C1_REDUNDANCY, stem=c1.C1_STEM)
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_par2.json")))
    got = {os.path.basename(p): hashlib.md5(open(p, "rb").read()).hexdigest() for p, _ in written}
    assert got == {n: v["md5"] for n, v in g["files"].items()}


@pytest.mark.parametrize("scheme", [par2hip.VOLUMES_EXPONENTIAL, par2hip.VOLUMES_SINGLE])
def test_multi_file_sets_equal_oracle(ctx, tmp_path, scheme):
    rng = np.random.default_rng(17 + scheme)
    files = [("b.dat", rng.integers(0, 256, 70000, dtype=np.uint8).tobytes()),
             ("empty.txt", b""),
             ("a.bin", rng.integers(0, 256, 5000, dtype=np.uint8).tobytes()),
             ("c.raw", rng.integers(0, 256, 3 * 2048, dtype=np.uint8).tobytes())]
    say("multi", scheme)
    gpu_vs_oracle(ctx, tmp_path, files, block_size=2048, redundancy=10.0, scheme=scheme)


@pytest.mark.parametrize("size,redundancy", [(3_000_001, 7.5), (100_000, 0.0), (1 << 20, 142.8)])
def test_automatic_block_size_and_edges(ctx, tmp_path, size, redundancy):
    data = c1.seeded_bytes(size, size).tobytes()
    say("auto", size, redundancy)
    gpu_vs_oracle(ctx, tmp_path, [("f.bin", data)], block_size=0, redundancy=redundancy)


def test_create_refuses_what_the_reference_refuses(ctx, tmp_path):
    with pytest.raises(par2hip.InvalidFormat):
        ctx.create_par2([("x", b"abc")], str(tmp_path / "x"), block_size=2046)
    with pytest.raises(par2hip.InvalidFormat):
        ctx.create_par2([("x", b"abc")], str(tmp_path / "x"), block_size=4096, redundancy=0.0,
                        scheme=par2hip.VOLUMES_SINGLE)


def test_c1_repair_with_the_written_set(ctx, tmp_path):
    """C1 end to end: create on the GPU, zero 3 seeded slices of the input,
    verify and rebuild them on the GPU from the written .par2 files."""
    data = c1.c1_input()
    written = ctx.create_par2([(c1.C1_NAME, data)], str(tmp_path / c1.C1_STEM), block_size=c1.C1_BLOCK,
                              redundancy=c1.C1_REDUNDANCY)
    s = par2fmt.load([p for p, _ in written])
    fid = s.file_ids[0]
    bs, ent = s.slice_size, s.ifsc[fid]
    n = len(ent)
    damaged = bytearray(data)
    for k in c1.c1_erasures():
        damaged[k * bs:(k + 1) * bs] = bytes(bs)
    d_sl = ctx.alloc(n * bs)
    d_sl.upload(np.frombuffer(bytes(damaged), np.uint8))
    rec = np.concatenate([np.frombuffer(s.recovery[e], np.uint8) for e in sorted(s.recovery)])
    d_rec = ctx.alloc(len(rec))
    d_rec.upload(rec)
    md5 = np.frombuffer(b"".join(m for m, _ in ent), np.uint8)
    crc = np.array([c for _, c in ent], np.uint32)
    say("c1 repair")
    bad = ctx.repair_set_dev(d_sl, n, bs, md5, crc, d_rec, sorted(s.recovery))
    ctx.sync()
    assert list(bad) == c1.c1_erasures()
    repaired = d_sl.download(n * bs).tobytes()[:s.files[fid].length]
    assert repaired == data and hashlib.md5(repaired).digest() == s.files[fid].md5
