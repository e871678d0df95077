"""The product's PAR2 loader (par2hip_par2_info; no device needed) on the
reference's QuickPar fixture and on sets the oracle writes: block size, file
count, recovery exponents and set ID as the packets give them, volumes found
beside the index by recovery set ID, damaged packets skipped."""
from __future__ import annotations

import hashlib
import os
import shutil
import struct

import pytest

import c1
import oracle
import par2fmt
import par2hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
QP = os.path.join(ROOT, "tests", "golden", "quickpar")

pytestmark = pytest.mark.skipif(not os.path.exists(par2hip.LIB_PATH), reason="libpar2hip.so not built")


def test_quickpar_fixture_info():
    info = par2hip.par2_info(os.path.join(QP, "testdata.par2"))
    s = par2fmt.load_dir(QP)
    assert info["block_size"] == s.slice_size == 5376
    assert info["files"] == 10 and info["blocks"] == 200
    assert (info["recovery_slices"], info["min_exponent"], info["max_exponent"]) == (60, 0, 59)
    main = open(os.path.join(QP, "testdata.par2"), "rb").read()
    assert info["set_id"] == main[32:48]


def test_volumes_found_by_set_id_only(tmp_path):
    for f in os.listdir(QP):
        shutil.copy(os.path.join(QP, f), tmp_path / f)
    # a stranger .par2 in the same directory (another recovery set) is ignored
    data = c1.seeded_bytes(9, 50000).tobytes()
    for name, b in oracle.par2_create([("other.bin", data)], block_size=2048, redundancy=20.0, stem="other"):
        open(tmp_path / name, "wb").write(b)
    info = par2hip.par2_info(str(tmp_path / "testdata.par2"))
    assert info["recovery_slices"] == 60 and info["files"] == 10
    # without some volumes fewer exponents are found
    os.remove(tmp_path / "testdata.vol31+29.par2")
    info = par2hip.par2_info(str(tmp_path / "testdata.par2"))
    assert (info["recovery_slices"], info["max_exponent"]) == (31, 30)


def test_oracle_written_set_info(tmp_path):
    data = c1.c1_input()
    for name, b in oracle.par2_create([(c1.C1_NAME, data)], block_size=c1.C1_BLOCK, redundancy=c1.C1_REDUNDANCY,
                                      stem=str(tmp_path / c1.C1_STEM)):
        open(name, "wb").write(b)
    info = par2hip.par2_info(str(tmp_path / "c1_input.par2"))
    assert (info["block_size"], info["blocks"], info["files"]) == (65536, 64, 1)
    assert (info["recovery_slices"], info["min_exponent"], info["max_exponent"]) == (4, 0, 3)
    main_body = struct.pack("<QI", 65536, 1) + par2fmt.load([str(tmp_path / "c1_input.par2")]).file_ids[0]
    assert info["set_id"] == hashlib.md5(main_body).digest()


def test_damaged_packets_and_garbage_are_skipped(tmp_path):
    idx = open(os.path.join(QP, "testdata.par2"), "rb").read()
    # garbage before and between packets, and a damaged FileDesc (MD5 mismatch)
    first = struct.unpack_from("<Q", idx, 8)[0]
    bad = bytearray(idx)
    pos = idx.find(b"PAR 2.0\0FileDesc") - 48
    bad[pos + 100] ^= 0xFF
    blob = b"junk" * 7 + bytes(bad[:first]) + b"\x00\x01\x02" + bytes(bad[first:])
    (tmp_path / "testdata.par2").write_bytes(blob)
    info = par2hip.par2_info(str(tmp_path / "testdata.par2"))
    assert info["block_size"] == 5376 and info["files"] == 9  # the damaged description is dropped
    with pytest.raises(par2hip.InvalidFormat):
        par2hip.par2_info(str(tmp_path / "missing.par2"))
    (tmp_path / "empty.par2").write_bytes(b"")
    with pytest.raises(par2hip.InvalidFormat):
        par2hip.par2_info(str(tmp_path / "empty.par2"))


def test_volume_must_start_with_a_packet(tmp_path):
    """A volume is recognised by the packet header at its offset 0
    (parser.rs:606-623): leading junk hides its recovery slices."""
    for f in os.listdir(QP):
        shutil.copy(os.path.join(QP, f), tmp_path / f)
    vol = tmp_path / "testdata.vol31+29.par2"
    vol.write_bytes(b"junk" + vol.read_bytes())
    info = par2hip.par2_info(str(tmp_path / "testdata.par2"))
    assert (info["recovery_slices"], info["max_exponent"]) == (31, 30)
