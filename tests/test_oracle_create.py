"""Pins of the oracle's PAR2 creation restatement (src/creator.rs:261-531,
src/writer.rs, src/volumes.rs): the reference's own known answers for block
sizes and volume names, and every byte of the output re-derived
independently (hashlib/zlib, the packet layout of the spec, the encoder
pinned by the QuickPar fixture), plus a repair from the written files.  Exact
creator bytes can only be pinned by running the Rust reference, which this
image cannot build; the committed digests pin the restatement against
regressions."""
from __future__ import annotations

import hashlib
import json
import os
import struct
import zlib

import numpy as np
import pytest

import c1
import oracle
import par2fmt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_block_size_reference_kats():
    # creator.rs:540-586
    assert oracle.par2_block_size(0) == 2048
    assert oracle.par2_block_size(1_000_000) == 2048
    assert oracle.par2_block_size(10_000_000) == 5000
    assert oracle.par2_block_size(1_000_000_000) == 500_000
    for size in [0, 1000, 1_000_000, 100_000_000, 10_000_000_000, 500_000_000_000]:
        bs = oracle.par2_block_size(size)
        assert bs % 4 == 0 and bs >= 2048
        if size <= 260_000_000_000:
            assert bs <= 4 * 1024 * 1024
        if size:
            assert -(-size // bs) <= 65535


def test_volume_names_reference_kats():
    # volumes.rs:214-247
    assert [oracle.par2_padding_width(x) for x in (0, 9, 10, 99, 100, 999, 1000)] == [1, 1, 2, 2, 3, 3, 4]
    assert oracle.par2_volume_name("test", 0, 0, 1) == "test.vol0+1.par2"
    assert oracle.par2_volume_name("test", 0, 1, 1) == "test.vol0+2.par2"
    assert oracle.par2_volume_name("test", 0, 9, 2) == "test.vol00+10.par2"
    assert oracle.par2_volume_name("test", 10, 19, 2) == "test.vol10+10.par2"
    assert oracle.par2_volume_name("archive", 0, 99, 3) == "archive.vol000+100.par2"


def _blocks(files, bs):
    """(canonical file order, padded blocks) re-derived with hashlib."""
    meta = []
    for name, data in files:
        h16 = hashlib.md5(data[:16384]).digest()
        fid = hashlib.md5(h16 + struct.pack("<Q", len(data)) + name.encode()).digest()
        meta.append((fid, name, data, h16))
    meta.sort(key=lambda m: m[0][::-1])  # parser.rs:556-564: byte 15 first
    blocks = []
    for _, _, data, _ in meta:
        for o in range(0, len(data), bs):
            blocks.append(data[o:o + bs].ljust(bs, b"\0"))
    return meta, blocks


def _check_set(vols, files, bs, r_expected, stem):
    meta, blocks = _blocks(files, bs)
    s = par2fmt.load_buffers([b for _, b in vols])
    assert s.bad_packets == 0
    # every packet of every file: framing, MD5, one set ID = MD5(main body)
    main_body = struct.pack("<QI", bs, len(files)) + b"".join(m[0] for m in meta)
    set_id = hashlib.md5(main_body).digest()
    for name, buf in vols:
        i, types = 0, []
        while i < len(buf):
            assert buf[i:i + 8] == par2fmt.MAGIC
            (length,) = struct.unpack_from("<Q", buf, i + 8)
            assert length % 4 == 0 and buf[i + 32:i + 48] == set_id
            assert hashlib.md5(buf[i + 32:i + length]).digest() == buf[i + 16:i + 32]
            types.append(buf[i + 48:i + 64])
            i += length
        assert i == len(buf)
        head = [par2fmt.T_CREATOR, par2fmt.T_MAIN] + [par2fmt.T_DESC] * len(files) + [par2fmt.T_IFSC] * len(files)
        assert types[:len(head)] == head and all(t == par2fmt.T_RECV for t in types[len(head):])
        creator_body = buf[64:64 + struct.unpack_from("<Q", buf, 8)[0] - 64]
        assert creator_body.rstrip(b"\0").decode() == "par2-rs v0.1.0 https://github.com/zephleggett/par2-rs"
    assert s.slice_size == bs and s.file_ids == [m[0] for m in meta]
    k = 0
    for fid, name, data, h16 in meta:
        d = s.files[fid]
        assert (d.name, d.length, d.md5, d.md5_16k) == (name, len(data), hashlib.md5(data).digest(), h16)
        ent = s.ifsc[fid]
        assert len(ent) == -(-len(data) // bs)
        for md5, crc in ent:
            assert md5 == hashlib.md5(blocks[k]).digest() and crc == zlib.crc32(blocks[k])
            k += 1
    assert sorted(s.recovery) == list(range(r_expected))
    if r_expected:
        words = np.frombuffer(b"".join(blocks), np.uint16).reshape(len(blocks), bs // 2)
        want = oracle.encode(words, np.arange(r_expected)).view(np.uint8).reshape(r_expected, bs)
        for e in range(r_expected):
            assert s.recovery[e] == want[e].tobytes()
    return s


def test_c1_output_rederived_and_digests():
    data = c1.c1_input()
    vols = oracle.par2_create([(c1.C1_NAME, data)], block_size=c1.C1_BLOCK, redundancy=c1.C1_REDUNDANCY,
                              stem=c1.C1_STEM)
    assert [n for n, _ in vols] == ["c1_input.par2", "c1_input.vol0+1.par2", "c1_input.vol1+2.par2",
                                    "c1_input.vol3+1.par2"]
    _check_set(vols, [(c1.C1_NAME, data)], c1.C1_BLOCK, 4, c1.C1_STEM)
    # exponents per volume: 1, 2, then the rest
    per = [sorted(par2fmt.load_buffers([b]).recovery) for _, b in vols]
    assert per == [[], [0], [1, 2], [3]]
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_par2.json")))
    assert g["input_md5"] == hashlib.md5(data).hexdigest()
    assert {n: hashlib.md5(b).hexdigest() for n, b in vols} == {n: v["md5"] for n, v in g["files"].items()}


def test_c1_repair_from_written_files():
    """Zero 3 seeded slices of the file, find them with the IFSC entries of
    the written set and rebuild them from its recovery slices."""
    data = bytearray(c1.c1_input())
    vols = oracle.par2_create([(c1.C1_NAME, bytes(data))], block_size=c1.C1_BLOCK, redundancy=c1.C1_REDUNDANCY,
                              stem=c1.C1_STEM)
    s = par2fmt.load_buffers([b for _, b in vols])
    bs = s.slice_size
    erased = c1.c1_erasures()
    damaged = bytearray(data)
    for k in erased:
        damaged[k * bs:(k + 1) * bs] = bytes(bs)
    ent = s.ifsc[s.file_ids[0]]
    bad = [k for k, (md5, crc) in enumerate(ent)
           if zlib.crc32(damaged[k * bs:(k + 1) * bs]) != crc or hashlib.md5(damaged[k * bs:(k + 1) * bs]).digest() != md5]
    assert bad == erased
    words = np.frombuffer(bytes(damaged), np.uint16).reshape(-1, bs // 2)
    n = words.shape[0]
    present = [k for k in range(n) if k not in bad]
    rec = np.stack([np.frombuffer(s.recovery[e], np.uint16) for e in range(len(bad))])
    out, _ = oracle.reconstruct(n, bad, present, np.arange(len(bad)), rec, words[present])
    for i, k in enumerate(bad):
        damaged[k * bs:(k + 1) * bs] = out[i].view(np.uint8).tobytes()
    assert bytes(damaged) == bytes(data)


def test_multi_file_order_and_single_scheme():
    rng = np.random.default_rng(3)
    files = [("b.dat", rng.integers(0, 256, 70000, dtype=np.uint8).tobytes()),
             ("empty.txt", b""),
             ("a.bin", rng.integers(0, 256, 5000, dtype=np.uint8).tobytes()),
             ("c.raw", rng.integers(0, 256, 3 * 2048, dtype=np.uint8).tobytes())]
    for scheme, names in ((oracle.SCHEME_EXPONENTIAL, None), (oracle.SCHEME_SINGLE, ["multi.par2"])):
        vols = oracle.par2_create(files, block_size=2048, redundancy=10.0, scheme=scheme, stem="multi")
        # blocks: 35 + 0 + 3 + 3 = 41; recovery bytes = int(41*2048*0.1) -> 5 slices
        _check_set(vols, files, 2048, 5, "multi")
        if names:
            assert [n for n, _ in vols] == names
        else:
            assert [n for n, _ in vols] == ["multi.par2", "multi.vol0+1.par2", "multi.vol1+2.par2",
                                            "multi.vol3+2.par2"]


def test_volume_split_counts():
    # volumes.rs:262-330: 7 recovery slices -> 1, 2, 4; 100 -> 2-digit names
    data = np.random.default_rng(4).integers(0, 256, 70 * 2048, dtype=np.uint8).tobytes()
    vols = oracle.par2_create([("x", data)], block_size=2048, redundancy=10.0, stem="t")
    assert [n for n, _ in vols] == ["t.par2", "t.vol0+1.par2", "t.vol1+2.par2", "t.vol3+4.par2"]
    vols = oracle.par2_create([("x", data)], block_size=2048, redundancy=142.8, stem="t")
    names = [n for n, _ in vols][1:]
    assert names[:3] == ["t.vol00+01.par2", "t.vol01+02.par2", "t.vol03+04.par2"] and len(names) == 7
    with pytest.raises(ValueError):  # the single scheme without recovery slices (creator.rs:455-457)
        oracle.par2_create([("x", data)], block_size=2048, redundancy=0.0, scheme=oracle.SCHEME_SINGLE)
