"""Repair of whole PAR2 sets through par2hip_repair_files, on the reference's
own repair cases over its QuickPar fixture (tests/external_multiblock_repair.rs:
47-240, restated: the same corruptions, byte-exact results, no clobber when the
rebuilt data cannot verify) and on a set the GPU creator wrote (C1)."""
from __future__ import annotations

import os
import shutil
import sys

import pytest

import c1
import oracle
import par2hip

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
QP = os.path.join(ROOT, "tests", "golden", "quickpar")
BS = 5376


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


def fixture(tmp_path):
    for f in os.listdir(QP):
        if f.endswith(".data") or f.endswith(".par2"):
            shutil.copy(os.path.join(QP, f), tmp_path / f)
    return str(tmp_path / "testdata.par2")


def corrupt_at(path, off, n):
    with open(path, "r+b") as f:
        f.seek(off)
        f.write(b"\xab" * n)


def pristine(name):
    return open(os.path.join(QP, name), "rb").read()


@pytest.mark.parametrize("case", ["two_in_one", "three_in_one", "two_files", "deleted", "single"])
def test_reference_repair_cases(ctx, tmp_path, case):
    idx = fixture(tmp_path)
    t1, t4 = tmp_path / "test-1.data", tmp_path / "test-4.data"
    if case == "two_in_one":
        corrupt_at(t1, 10, 64)
        corrupt_at(t1, BS * 3 + 10, 64)
    elif case == "three_in_one":
        corrupt_at(t1, 10, 50)
        corrupt_at(t1, BS * 2 + 10, 50)
        corrupt_at(t1, BS * 5 + 10, 50)
    elif case == "two_files":
        corrupt_at(t1, BS + 5, 40)
        corrupt_at(t4, BS * 2 + 5, 40)
    elif case == "deleted":
        os.remove(t1)
    else:
        corrupt_at(t1, 10, 32)
    say("fixture", case)
    rep = ctx.repair_files(idx)
    assert rep["repaired"] == 1
    assert open(t1, "rb").read() == pristine("test-1.data")
    assert open(t4, "rb").read() == pristine("test-4.data")
    assert not any(f.endswith(".par2tmp") for f in os.listdir(tmp_path))
    after = ctx.repair_files(idx, verify_only=True)
    assert after["blocks_damaged"] == 0


def test_bad_reconstruction_does_not_clobber(ctx, tmp_path):
    """Recovery data tampered inside the exponent-0 volume (its packet MD5 is
    not checked when loading, as in the reference): the rebuilt block fails
    verification, the original keeps its damaged bytes, no temp file stays."""
    idx = fixture(tmp_path)
    vol = tmp_path / "testdata.vol00+01.par2"
    data = bytearray(open(vol, "rb").read())
    pkt = data.find(b"PAR 2.0\x00RecvSlic")
    start = pkt + 16 + 4
    for i in range(start + 100, start + 100 + BS // 2):
        data[i] ^= 0xFF
    open(vol, "wb").write(bytes(data))
    t1 = tmp_path / "test-1.data"
    corrupt_at(t1, BS + 5, 64)
    snapshot = open(t1, "rb").read()
    say("tampered recovery")
    with pytest.raises(par2hip.RepairFailed):
        ctx.repair_files(idx)
    assert open(t1, "rb").read() == snapshot
    assert not (tmp_path / "test-1.data.par2tmp").exists()


def test_intact_verify_and_insufficient(ctx, tmp_path):
    idx = fixture(tmp_path)
    rep = ctx.repair_files(idx)
    assert (rep["files"], rep["blocks"], rep["blocks_damaged"], rep["repaired"]) == (10, 200, 0, 0)
    assert rep["recovery_slices"] == 60
    # 61 damaged blocks > 60 recovery slices: refused, nothing written
    for name in ("test-1.data", "test-3.data"):
        os.remove(tmp_path / name)
    rep = ctx.repair_files(idx, verify_only=True)
    assert rep["files_missing"] == 2 and rep["blocks_damaged"] == 33 + 27
    corrupt_at(tmp_path / "test-0.data", 0, 8)
    with pytest.raises(par2hip.InsufficientRecovery):
        ctx.repair_files(idx)
    assert not (tmp_path / "test-1.data").exists()
    assert not any(f.endswith(".par2tmp") for f in os.listdir(tmp_path))


def test_size_mismatch_rebuilds_whole_file(ctx, tmp_path):
    idx = fixture(tmp_path)
    t5 = tmp_path / "test-5.data"
    with open(t5, "ab") as f:
        f.write(b"extra tail")
    say("size mismatch")
    rep = ctx.repair_files(idx)
    assert rep["files_damaged"] == 1 and rep["blocks_damaged"] == -(-99950 // BS)
    assert open(t5, "rb").read() == pristine("test-5.data")


def test_c1_created_damaged_and_repaired(ctx, tmp_path):
    """C1 end to end in the product: create the set on the GPU, zero 3 seeded
    slices of the input file, repair it through the temp-file protocol."""
    data = c1.c1_input()
    path = tmp_path / c1.C1_NAME
    path.write_bytes(data)
    written = ctx.create_par2([(c1.C1_NAME, data)], str(tmp_path / c1.C1_STEM), block_size=c1.C1_BLOCK,
                              redundancy=c1.C1_REDUNDANCY)
    damaged = bytearray(data)
    for k in c1.c1_erasures():
        damaged[k * c1.C1_BLOCK:(k + 1) * c1.C1_BLOCK] = bytes(c1.C1_BLOCK)
    path.write_bytes(bytes(damaged))
    say("c1 repair files")
    rep = ctx.repair_files(written[0][0])
    assert (rep["blocks_damaged"], rep["repaired"]) == (3, 1)
    assert path.read_bytes() == data


@pytest.mark.parametrize("name", ["../escape.bin", "/tmp/escape.bin", "sub/../../escape.bin"])
def test_unsafe_stored_name_is_refused(ctx, tmp_path, name):
    """A set whose FileDesc names a path outside the base directory (absolute or
    with a '..' component) is refused before any file is read or written; the
    reference would join the name to the base path as it stands."""
    data = c1.seeded_bytes(5, 20000).tobytes()
    sub = tmp_path / "set"
    sub.mkdir()
    for fname, b in oracle.par2_create([(name, data)], block_size=4096, redundancy=20.0, stem=str(sub / "evil")):
        open(fname, "wb").write(b)
    say("unsafe name", name)
    with pytest.raises(par2hip.InvalidFormat, match="unsafe file name"):
        ctx.repair_files(str(sub / "evil.par2"))
    assert not (tmp_path / "escape.bin").exists() and not (tmp_path / "escape.bin.par2tmp").exists()
    assert all(p.endswith(".par2") for p in os.listdir(sub))


def _snapshot(d):
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d)) if f.endswith(".data")}


def test_cancel_preset_flag_leaves_originals(ctx, tmp_path):
    """tests/info_and_cancel.rs:216-252: a flag set before the call cancels it;
    the damaged files stay exactly as they were and no temp is left."""
    idx = fixture(tmp_path)
    for name in ("test-1.data", "test-6.data"):
        corrupt_at(tmp_path / name, 2048, 8192)
    before = _snapshot(tmp_path)
    flag = par2hip.CancelFlag(True)
    say("cancel preset")
    with pytest.raises(par2hip.Cancelled):
        ctx.repair_files(idx, cancel=flag)
    assert _snapshot(tmp_path) == before
    assert not any(f.endswith(".par2tmp") for f in os.listdir(tmp_path))
    # verify-only is cancelled too, and a clear flag repairs
    with pytest.raises(par2hip.Cancelled):
        ctx.repair_files(idx, verify_only=True, cancel=flag)
    rep = ctx.repair_files(idx, cancel=par2hip.CancelFlag(False))
    assert rep["repaired"] == 1
    assert open(tmp_path / "test-1.data", "rb").read() == pristine("test-1.data")


def test_cancel_from_thread_is_safe(ctx, tmp_path):
    """tests/info_and_cancel.rs:254-308: the flag set from another thread at
    once; a cancelled call left the originals byte-identical, a completed one
    repaired them, and neither leaves a temp."""
    import threading

    idx = fixture(tmp_path)
    corrupt_at(tmp_path / "test-4.data", 512, 16384)
    before = _snapshot(tmp_path)
    flag = par2hip.CancelFlag(False)
    t = threading.Thread(target=flag.set)
    say("cancel from thread")
    t.start()
    try:
        ctx.repair_files(idx, cancel=flag)
        assert open(tmp_path / "test-4.data", "rb").read() == pristine("test-4.data")
    except par2hip.Cancelled:
        assert _snapshot(tmp_path) == before
    t.join()
    assert not any(f.endswith(".par2tmp") for f in os.listdir(tmp_path))


def test_cancel_at_growing_delays(ctx, tmp_path):
    """The flag raised by a timer at growing delays, so that successive calls
    are cancelled at later and later safe points (loading, verification, the
    rebuild, the temps, the renames) until one completes: every cancelled call
    leaves the originals byte-identical and no temp, the completed one
    repairs them."""
    import threading

    idx = fixture(tmp_path)
    corrupt_at(tmp_path / "test-1.data", 10, 64)
    corrupt_at(tmp_path / "test-4.data", BS * 2 + 5, 40)
    before = _snapshot(tmp_path)
    done = False
    for delay_ms in (0.0, 0.5, 1, 2, 4, 8, 16, 32, 64, 128, 256, 1000):
        flag = par2hip.CancelFlag(False)
        tm = threading.Timer(delay_ms / 1e3, flag.set)
        tm.start()
        say("cancel after ms", delay_ms)
        try:
            ctx.repair_files(idx, cancel=flag)
            done = True
        except par2hip.Cancelled:
            assert _snapshot(tmp_path) == before
        tm.cancel()
        tm.join()
        assert not any(f.endswith(".par2tmp") for f in os.listdir(tmp_path))
        if done:
            break
    assert done
    assert open(tmp_path / "test-1.data", "rb").read() == pristine("test-1.data")
    assert open(tmp_path / "test-4.data", "rb").read() == pristine("test-4.data")


def test_overflowing_layout_refused(ctx, tmp_path):
    """An index whose block size (2^62) and file lengths would wrap the block
    layout's byte count is refused with EINVAL before any buffer is sized."""
    import struct
    import hashlib

    sid = bytes([7]) * 16

    def packet(ptype, body):
        body = body + b"\0" * (-len(body) % 4)
        md5 = hashlib.md5(sid + ptype + body).digest()
        return b"PAR2\0PKT" + struct.pack("<Q", 64 + len(body)) + md5 + sid + ptype + body

    ids = [bytes([k + 1]) + b"\0" * 15 for k in range(4)]
    out = packet(b"PAR 2.0\0Main\0\0\0\0", struct.pack("<QI", 1 << 62, 4) + b"".join(ids))
    for k, fid in enumerate(ids):
        out += packet(b"PAR 2.0\0FileDesc", fid + b"\0" * 32 + struct.pack("<Q", 1 << 62) + b"f%d" % k)
    p = tmp_path / "crafted.par2"
    p.write_bytes(out)
    say("crafted overflow index")
    with pytest.raises(par2hip.InvalidFormat):
        ctx.repair_files(str(p), verify_only=True)
    with pytest.raises(par2hip.InvalidFormat):
        ctx.repair_files(str(p))
