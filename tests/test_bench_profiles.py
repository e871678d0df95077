"""bench.py reads its roofline traffic and unit utilisation from the committed
counter summaries under profiles/: every workload the bench reports a
roofline for has its contraction's row there, the figures parse, and every
summary names the command (committed files only) that made it."""
from __future__ import annotations

import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

CONTRACTIONS = [("plain_encode_1GiB", "gf_muladd_mfma"), ("plain_repair_1GiB", "gf_muladd_lds"),
                ("batched_256_sets", "gf_batch"), ("small_r_R4_1GiB", "gf_muladd_mfma"),
                ("small_r_R4_1GiB", "gf_muladd_lds")]


def test_kernel_tags():
    assert bench.kernel_tag("void par2hip::k_gf_mfma<false, false, 4, 2, 4, 4>(unsigned short const* const*)") == \
        "gf_muladd_mfma"
    assert bench.kernel_tag("par2hip::k_gf_mfma<true, true>(int)") == "gf_muladd_mfma_i8"
    assert bench.kernel_tag("void par2hip::k_gf_mfma_ring<2, 2, 4, 8>(unsigned short const* const*)") == \
        "gf_muladd_mfma"
    assert bench.kernel_tag("void par2hip::k_gf_lds<8, 2, 16>(x)") == "gf_muladd_lds"
    assert bench.kernel_tag("par2hip::k_gf_lds_batch(x)") == "gf_batch"
    assert bench.kernel_tag("__amd_rocclr_copyBuffer") == "copy"


def test_every_roofline_has_its_traffic_row():
    for workload, tag in CONTRACTIONS:
        traffic, src = bench.profile_traffic(workload, tag)
        assert traffic is not None and traffic > 1e8, (workload, tag, src)
        assert "profiles/hbm_traffic_pmc.md" in src


def test_missing_row_is_reported_not_guessed():
    traffic, why = bench.profile_traffic("plain_encode_1GiB", "no_such_kernel")
    assert traffic is None and "no no_such_kernel row" in why


def test_unit_rows():
    u = bench.profile_units("plain_encode_1GiB", "gf_muladd_mfma")
    assert u is not None and 0 < u["mfma_util"] <= 1 and 1.0 < u["clock_ghz"] < 3.0
    b = bench.profile_units("batched_256_sets", "gf_batch")
    assert b is not None and 0 < b["lds_util"] <= 1


def test_summaries_name_their_commands():
    names = [os.path.join(bench.PROFILES, f"{w}_kernel_trace.md") for w in bench.TRACE_RUNS]
    names += [bench.TRAFFIC_SUMMARY, bench.UNITS_SUMMARY]
    for path in names:
        assert os.path.exists(path), path
        text = open(path).read()
        cmds = re.findall(r"^rocprofv3 .*$", text, re.M)
        assert cmds, path
        for c in cmds:
            prog = c.split(" -- ", 1)[1].split()
            assert prog[0] == "python3" and prog[1] == "bench.py", c


def test_md5_issue_floor():
    """The MD5 floor counts the blocks MD5 pads a message to (RFC 1321: the
    bytes, 0x80 and the 8-byte length, rounded up to 64) at 276 issue slots
    x 4 clocks each at 2.4 GHz, one round per SIMD's worth of 64-chain waves."""
    for s in (1, 55, 56, 63, 64, 119, 120, 512 << 10, 1 << 20):
        blocks = -(-(s + 9) // 64)
        assert bench.md5_issue_floor_ms(64, s, 1) == blocks * 276 * 4 / 2.4e9 * 1e3, s
    one = bench.md5_issue_floor_ms(64 * 128, 4096, 128)
    assert bench.md5_issue_floor_ms(64 * 128 + 1, 4096, 128) == 2 * one
    assert abs(bench.md5_issue_floor_ms(1024, 1 << 20, 128) - 16385 * 1104 / 2.4e6) < 1e-9
