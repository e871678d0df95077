"""The host MD5 behind the creator's whole-file, 16 KiB and packet digests
(par2-rs_amd/csrc/md5_host.hpp and md5_host.cpp, built with g++) against Python's hashlib
on every length 0..300, on split updates and on a 5 MiB string; plus the RFC
1321 appendix A.5 suite through the same program's first lines."""
import hashlib
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lcg_bytes(n: int) -> bytes:
    """The byte stream md5_host_check.cpp hashes: the top byte of a 32-bit LCG."""
    out = bytearray(n)
    x = 0x12345678
    for i in range(n):
        x = (x * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = x >> 24
    return bytes(out)


def test_md5_host_matches_hashlib(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path / "md5_host_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "par2-rs_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "md5_host_check.cpp"),
                    os.path.join(ROOT, "par2-rs_amd", "csrc", "md5_host.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, check=True).stdout.split()
    data = lcg_bytes(5 << 20)
    want = [hashlib.md5(data[:n]).hexdigest() for n in range(301)]
    want += [hashlib.md5(data[:1000]).hexdigest()] * 5
    want.append(hashlib.md5(data).hexdigest())
    assert out == want
    # RFC 1321 A.5 values for the empty string and "abc" as a sanity anchor
    assert hashlib.md5(b"").hexdigest() == "d41d8cd98f00b204e9800998ecf8427e" == out[0]
