"""The host-only PAR2 set loader (par2-rs_amd/csrc/par2set.cpp: parser, volume
discovery, canonical order, checked block layout, block-size rule) built on its
own with AddressSanitizer and UndefinedBehaviorSanitizer and run over the
reference's QuickPar fixture, crafted overflow indexes, damaged and truncated
copies, garbage and a bounded random fuzz (tests/cpp/par2set_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "par2-rs_amd", "csrc")


def test_par2set_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path / "par2set_check"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I", CSRC, os.path.join(CSRC, "par2set.cpp"),
                    os.path.join(CSRC, "md5_host.cpp"), os.path.join(ROOT, "tests", "cpp", "par2set_check.cpp"),
                    "-o", str(exe)], check=True)
    # verify_asan_link_order=0: a library preloaded by the environment may come first
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "quickpar"), str(tmp_path), "3000"],
                         capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr[-4000:]
