"""Minimal PAR2 2.0 packet reader used by the tests to open the golden fixture.

Follows the packet layout the reference parses in src/parser.rs:181-547
(64-byte header: magic "PAR2\\0PKT", u64 length, MD5 over set-id+type+body,
set id, type) and the canonical file order of src/parser.rs:556-596
(the Main packet's recovery-set File ID list).  Test infrastructure only.
"""
from __future__ import annotations

import glob
import hashlib
import os
import struct
from dataclasses import dataclass, field

MAGIC = b"PAR2\0PKT"
T_MAIN = b"PAR 2.0\0Main\0\0\0\0"
T_DESC = b"PAR 2.0\0FileDesc"
T_IFSC = b"PAR 2.0\0IFSC\0\0\0\0"
T_RECV = b"PAR 2.0\0RecvSlic"
T_CREATOR = b"PAR 2.0\0Creator\0"


@dataclass
class FileDesc:
    file_id: bytes
    md5: bytes
    md5_16k: bytes
    length: int
    name: str


@dataclass
class Par2Set:
    slice_size: int = 0
    file_ids: list = field(default_factory=list)      # Main packet order (canonical)
    files: dict = field(default_factory=dict)         # file_id -> FileDesc
    ifsc: dict = field(default_factory=dict)          # file_id -> [(md5, crc32)]
    recovery: dict = field(default_factory=dict)      # exponent -> bytes
    bad_packets: int = 0

    def ordered_files(self):
        return [self.files[f] for f in self.file_ids]


def read_packets(buf: bytes):
    """Yield (type, set_id, body) of every packet whose MD5 verifies; resync on bad magic."""
    i = 0
    n = len(buf)
    while i + 64 <= n:
        if buf[i:i + 8] != MAGIC:
            i += 1
            continue
        (length,) = struct.unpack_from("<Q", buf, i + 8)
        if length < 64 or length % 4 or i + length > n:
            i += 1
            continue
        md5 = buf[i + 16:i + 32]
        if hashlib.md5(buf[i + 32:i + length]).digest() != md5:
            i += 1
            yield None
            continue
        yield (buf[i + 48:i + 64], buf[i + 32:i + 48], buf[i + 64:i + length])
        i += length


def load(paths) -> Par2Set:
    bufs = []
    for p in paths:
        with open(p, "rb") as f:
            bufs.append(f.read())
    return load_buffers(bufs)


def load_buffers(bufs) -> Par2Set:
    """Same as load() on the bytes of the .par2 files."""
    s = Par2Set()
    for buf in bufs:
        for pkt in read_packets(buf):
            if pkt is None:
                s.bad_packets += 1
                continue
            typ, _sid, body = pkt
            if typ == T_MAIN:
                s.slice_size, nrec = struct.unpack_from("<QI", body, 0)
                s.file_ids = [body[12 + 16 * k:28 + 16 * k] for k in range(nrec)]
            elif typ == T_DESC:
                fid, md5, md5k = body[0:16], body[16:32], body[32:48]
                (length,) = struct.unpack_from("<Q", body, 48)
                name = body[56:].rstrip(b"\0").decode()
                s.files[fid] = FileDesc(fid, md5, md5k, length, name)
            elif typ == T_IFSC:
                fid = body[:16]
                ent = []
                for k in range(16, len(body), 20):
                    ent.append((body[k:k + 16], struct.unpack_from("<I", body, k + 16)[0]))
                s.ifsc[fid] = ent
            elif typ == T_RECV:
                (e,) = struct.unpack_from("<I", body, 0)
                s.recovery[e] = body[4:]
    return s


def load_dir(d: str, stem: str = "testdata") -> Par2Set:
    return load(sorted(glob.glob(os.path.join(d, stem + "*.par2"))))
