"""Benchmark of the PAR2 Reed-Solomon hot path on one MI355X.

Default workload (plain GPU encode): one 1 GiB set of 1024 input slices of
1 MiB (uniform random bytes, seeded), 102 recovery slices with exponents
0..101 over GF(2^16) poly 0x1100B, plus the IFSC MD5 and CRC32 of every input
slice.  Inputs are resident in HBM before timing starts; one step is one
par2hip_encode_dev call (coefficients, the contraction as the library plans
it beside the MD5 chains, both hashes) and its completion.  Other workloads
(``--workload``): repair (2000 x 512 KiB, 100 erasures, verify hashes of all
slices + inverse + rebuild), micro (each contraction variant and hash alone),
batched (256 sets per planned, graph-replayed batch), served (Poisson
requests through the serving loop against the plain path at equal p99),
create and repair_files (whole files to and from .par2 sets), encode_host
(the default set handed over in host memory, both transfers timed), copy
(the calibration copy for the HBM traffic counters).

Prints ONE JSON line (rank 0): metric/value/unit, ms_per_step, p50/p99 step
latency, the roofline object of the step's contraction (HIP-event durations
measured on the kernel's own stream inside the timed region; HBM bytes, or
matrix-core operations for the bit-sliced form) and the CPU baseline (the
oracle's restatement of the reference's AVX2 strategy and hashing, timed on
this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "par2-rs_amd"))

import par2hip  # noqa: E402
from workloads import draw_requests, draw_sets, random_bytes  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch come from the committed counter summary, never from
# constants here: rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of each
# workload (PMC_RUNS), each counter scaled by what it reads for a 1 GiB
# device-to-device copy, summarised by `bench.py --pmc-summary`.  A workload
# whose contraction is no longer the kernel the summary measured reports
# traffic null and says why.
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_SUMMARY = os.path.join(PROFILES, "hbm_traffic_pmc.md")
PMC_OUT = "gpurun_out/pmc"
PMC_RUNS = {
    "copy_1GiB": "--workload copy --steps 3",
    "plain_encode_1GiB": "--steps 3 --warmup 1 --no-cpu --legs",
    "plain_repair_1GiB": "--workload repair --steps 3 --warmup 1 --no-cpu",
    "batched_256_sets": "--workload batched --steps 1 --warmup 1 --no-cpu",
    "small_r_R4_1GiB": "--workload small_r --small-r 4 --steps 3 --warmup 1 --autotune 0",
}
# Unit utilisation of the contractions (SQ counters of the same workloads,
# committed with their commands in profiles/gf_units_utilisation_pmc.md).
UNITS_SUMMARY = os.path.join(PROFILES, "gf_units_utilisation_pmc.md")
FP4_DENSE_PEAK_TFLOPS = 10100.0
INT8_DENSE_PEAK_TOPS = 5000.0
# The MD5 chains that bound the encode and repair steps: MI355X's peak shader
# clock, and the instructions a chain wave must issue per 64-byte block in the
# helper-wave form those steps run (its helper wave precomputes each step's
# message word plus constant): 4 per MD5 step (round function, add3, rotate,
# add), the four state additions and 16 LDS reads of the 64 precomputed sums.
SHADER_CLOCK_GHZ = 2.4
MD5_SLOTS_PER_BLOCK = 276
METRIC = "encode/repair GB/s source data + p50/p99 latency vs plain GPU path, bit-exact"


def md5_issue_floor_ms(n: int, s: int, hash_simds: int) -> float:
    """Lower bound on hashing n full s-byte slices with one serial MD5 chain
    per lane: a wave of 64 chains issues one instruction per 4 clocks, so a
    chain of B blocks (the slice's and the final padding block) takes B x 276
    x 4 clocks at the peak shader clock; waves beyond one per SIMD of the hash
    CUs run in further rounds."""
    blocks = (s + 8) // 64 + 1
    rounds = -(-((n + 63) // 64) // max(hash_simds, 1))
    return rounds * blocks * MD5_SLOTS_PER_BLOCK * 4 / (SHADER_CLOCK_GHZ * 1e9) * 1e3


def md5_floor_fields(n: int, s: int, md5_ms: float, ctx) -> dict:
    """The step's MD5 time beside its issue floor."""
    floor = md5_issue_floor_ms(n, s, 4 * int(ctx.device_info().get("hash_cus", 0)))
    return {"md5_ms": round(md5_ms, 4), "md5_issue_floor_ms": round(floor, 4),
            "md5_floor_frac": round(floor / md5_ms, 4) if md5_ms > 0 else None}


def kernel_tag(symbol: str) -> str:
    """The library's timing tag of a device kernel symbol."""
    name = symbol.split("(")[0].replace("par2hip::", "").strip()
    if name.startswith("void "):
        name = name[5:]
    base = name.split("<")[0]
    if base == "k_gf_mfma":
        return "gf_muladd_mfma_i8" if name.startswith("k_gf_mfma<true") else "gf_muladd_mfma"
    if base == "k_gf_mfma_ring":
        return "gf_muladd_mfma"
    return {"k_gf_lds": "gf_muladd_lds", "k_gf_lds_batch": "gf_batch", "k_md5_slices": "md5_slices",
            "k_crc32_slices": "crc32_slices", "k_md5_list": "md5_list", "k_crc32_list": "crc32_list",
            "__amd_rocclr_copyBuffer": "copy"}.get(base, base)


def summary_rows(path: str, ncells: int) -> list[list[str]]:
    """Data rows of the markdown tables in `path` with `ncells` cells."""
    rows = []
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return rows
    for ln in lines:
        ln = ln.strip()
        if not ln.startswith("|") or set(ln) <= set("|-: "):
            continue
        cells = [c.strip().strip("`") for c in ln.strip("|").split("|")]
        if len(cells) == ncells:
            rows.append(cells)
    return rows


def profile_traffic(workload: str, tag: str):
    """HBM bytes per launch of the `tag` kernel in `workload` (the largest
    such kernel: the main contraction, not a fold) from the committed counter
    summary, and where the figure came from (or why there is none)."""
    best, sym = None, None
    for w, kernel, t, _n, _f, _wr, total in summary_rows(TRAFFIC_SUMMARY, 7):
        if w == workload and t == tag and (best is None or float(total) > best):
            best, sym = float(total), kernel
    if best is None:
        have = sorted({t for w, _k, t, *_ in summary_rows(TRAFFIC_SUMMARY, 7) if w == workload})
        return None, f"no {tag} row for {workload} in profiles/hbm_traffic_pmc.md (it holds {have})"
    return best, f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of {sym}, profiles/hbm_traffic_pmc.md"


def profile_units(workload: str, tag: str):
    """Unit utilisation of the `tag` kernel in `workload` from the committed
    SQ-counter summary (None when it has no such row)."""
    keys = ["clock_ghz", "mfma_util", "lds_util", "valu_busy", "lds_bank_conflict", "busy_cu_frac"]
    for row in summary_rows(UNITS_SUMMARY, 10):
        w, kernel, t, _ms = row[:4]
        if w == workload and t == tag:
            vals = {}
            for k, v in zip(keys, row[4:]):
                v = v.rstrip("%")
                vals[k] = None if v in ("-", "") else round(float(v) / (1 if k == "clock_ghz" else 100), 4)
            if vals.get("mfma_util") is not None and vals.get("busy_cu_frac"):
                vals["mfma_util_busy_cus"] = round(vals["mfma_util"] / vals["busy_cu_frac"], 4)
            vals["kernel"] = kernel
            vals["source"] = "rocprofv3 --pmc SQ counters, profiles/gf_units_utilisation_pmc.md"
            return vals
    return None


def pmc_commands() -> list[str]:
    """The rocprofv3 counter runs behind the two summaries, one counter set
    per run (FETCH_SIZE and WRITE_SIZE use 3 and 2 of the 4 TCC counters)."""
    cmds = []
    for w, a in PMC_RUNS.items():
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            cmds.append(f"rocprofv3 --pmc {c} -d {PMC_OUT}/{w}_{c} -o run --output-format csv -- "
                        f"python3 bench.py {a}")
        if w != "copy_1GiB":
            cmds.append(f"rocprofv3 --kernel-trace --pmc {' '.join(UNIT_COUNTERS)} -d {PMC_OUT}/{w}_units -o run "
                        f"--output-format csv -- python3 bench.py {a}")
    return cmds


# rocprofv3 --kernel-trace --stats runs: profile name -> bench.py arguments
TRACE_OUT = "gpurun_out/trace"
TRACE_RUNS = {
    "plain_encode_1GiB_R102_default_bench": "",
    "plain_repair_2000x512KiB": "--workload repair --no-cpu",
    "batched_256_sets": "--workload batched --steps 5 --warmup 3 --no-cpu",
    "served_poisson": "--workload served --no-cpu",
    "small_r_1GiB": "--workload small_r --steps 10",
    "gf_variants_1GiB_R102": "--workload micro --steps 5 --warmup 1",
    "inverse_m1000_5000_12000": "--workload inverse --steps 3 --warmup 1",
    "create_1x1GiB_R103": "--workload create --create-files 1 --create-file-mib 1024 --create-block 1048576 "
                          "--steps 3 --warmup 1",
    "create_16x64MiB": "--workload create --steps 3 --warmup 1",
    "repair_files_16x64MiB": "--workload repair_files --steps 3 --warmup 1",
}


def trace_commands() -> list[str]:
    return [f"rocprofv3 --kernel-trace --stats -d {TRACE_OUT}/{w} -o run --output-format csv -- "
            f"python3 bench.py {a}".rstrip() for w, a in TRACE_RUNS.items()]


def trace_summary() -> None:
    """Writes profiles/<name>_kernel_trace.md for every run of trace_commands()
    whose output is under gpurun_out/trace: the kernels by total time, and the
    bench's own line (its HIP-event kernel averages come from the same run)."""
    import csv
    import glob

    for w, a in TRACE_RUNS.items():
        stats = glob.glob(os.path.join(TRACE_OUT, w, "**", "*kernel_stats.csv"), recursive=True)
        if not stats:
            continue
        rows = list(csv.DictReader(open(stats[0])))
        rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
        cmd = f"rocprofv3 --kernel-trace --stats -d {TRACE_OUT}/{w} -o run --output-format csv -- " \
              f"python3 bench.py {a}".rstrip()
        lines = [f"# Kernel trace: {w}", "", "Command (repo root, one MI355X), written by "
                 "`python3 bench.py --trace-summary`:", "", "```", cmd, "```", "",
                 "`run_kernel_stats.csv`, kernels by total time (all of the run: warm-up, timed steps and the",
                 "checks):", "", "| kernel | calls | avg ms | min ms | max ms | % of kernel time |",
                 "|---|---|---|---|---|---|"]
        for r in rows[:24]:
            name = r["Name"].split("(")[0]
            lines.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                         f"{float(r['MinNs']) / 1e6:.4f} | {float(r['MaxNs']) / 1e6:.4f} | "
                         f"{float(r['Percentage']):.2f} |")
        log_path = os.path.join(TRACE_OUT, w + ".log")
        line = None
        if os.path.exists(log_path):
            for ln in open(log_path):
                if ln.startswith("{"):
                    line = json.loads(ln)
        # each roofline kernel of the bench line against its launches in the
        # trace: the main launches (at least half the longest one's duration)
        # of the symbol with that tag that spends the most time
        # (the top-level workload only: the legs of the default line have
        # runs of their own, where their kernels are the largest of their tag)
        roofs = []
        if line and isinstance(line.get("roofline"), dict) and line["roofline"].get("kernel"):
            roofs.append(("step", line["roofline"]["kernel"], line["roofline"].get("kernel_ms")))
        traces = glob.glob(os.path.join(TRACE_OUT, w, "**", "*kernel_trace.csv"), recursive=True)
        if roofs and traces:
            durs = {}
            for r in csv.DictReader(open(traces[0])):
                durs.setdefault(r["Kernel_Name"], []).append(
                    (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
            lines += ["", "Roofline kernels of the bench line against the trace (main launches: at least half the",
                      "longest launch of that symbol; the trace's launches run alone or beside the same work as in",
                      "the bench's timed steps):", "",
                      "| leg | tag | symbol | main launches | trace avg ms | bench HIP-event avg ms |",
                      "|---|---|---|---|---|---|"]
            for leg, tag, bench_ms in roofs:
                cands = [(sum(d), k, d) for k, d in durs.items() if kernel_tag(k) == tag]
                if not cands:
                    continue
                _, sym, d = max(cands)
                top = max(d)
                main = [x for x in d if x >= 0.5 * top]
                lines.append(f"| {leg} | {tag} | `{sym.split('(')[0]}` | {len(main)} | "
                             f"{sum(main) / len(main):.4f} | {bench_ms} |")
        if line:
            lines += ["", "The bench's line from the same run (HIP events on each kernel's own stream):", ""]

            def brief(o, depth=0):
                keep = ("metric", "value", "unit", "ms_per_step", "p50_ms", "p99_ms", "bit_exact_sample",
                        "kernel_ms", "roofline", "config", "batch_eager_ms", "plain_path", "cpu_baseline", "exact")
                out = {}
                for k, v in o.items():
                    if k in keep:
                        out[k] = v
                    elif isinstance(v, dict) and depth == 0 and k in ("micro", "small_r", "inverse"):
                        out[k] = v  # these workloads' whole result is one table of per-case timings
                    elif isinstance(v, dict) and depth == 0 and ("value" in v or "kernel_ms" in v):
                        out[k] = brief(v, 1)
                return out

            lines += ["```", json.dumps(brief(line), indent=1)[:6000], "```"]
        open(os.path.join(PROFILES, f"{w}_kernel_trace.md"), "w").write("\n".join(lines) + "\n")


UNIT_COUNTERS = ["SQ_VALU_MFMA_BUSY_CYCLES", "SQ_LDS_IDX_ACTIVE", "SQ_ACTIVE_INST_VALU", "SQ_LDS_BANK_CONFLICT",
                 "SQ_BUSY_CU_CYCLES", "GRBM_GUI_ACTIVE"]
XCCS = 8          # MI355X: 8 XCDs; GRBM_GUI_ACTIVE is summed over them
CUS, SIMDS = 256, 1024


def read_counters(d: str):
    """Per-dispatch counter sums and kernel names from a rocprofv3 CSV run
    directory: {dispatch: (kernel, {counter: value}, duration_ns or None)}."""
    import csv
    import glob

    out = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for row in csv.DictReader(open(f)):
            key = (row.get("Agent_Id"), row["Dispatch_Id"])
            k, cnt, dur = out.setdefault(key, (row["Kernel_Name"], {}, None))
            cnt[row["Counter_Name"]] = cnt.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    durs = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        for row in csv.DictReader(open(f)):
            durs[(row.get("Agent_Id"), row["Dispatch_Id"])] = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
    return {k: (v[0], v[1], durs.get(k)) for k, v in out.items()}


def pmc_summary() -> None:
    """Writes profiles/hbm_traffic_pmc.md and profiles/gf_units_utilisation_pmc.md
    from the counter runs of pmc_commands() (averages per dispatch by kernel)."""
    def main_launches(vals):
        """The launches of one kernel symbol that do its main work: at least
        half the largest one's figure (a fold or a staging copy of the same
        kernel is far smaller)."""
        top = max(vals)
        return [v for v in vals if v >= 0.5 * top]

    def per_kernel(d, counter):
        acc = {}
        for kernel, cnt, _ in read_counters(d).values():
            if counter in cnt:
                acc.setdefault(kernel, []).append(cnt[counter])
        out = {}
        for k, vals in acc.items():
            main = main_launches(vals)
            out[k] = (len(main), sum(main) / len(main))
        return out

    cf = per_kernel(f"{PMC_OUT}/copy_1GiB_FETCH_SIZE", "FETCH_SIZE")
    cw = per_kernel(f"{PMC_OUT}/copy_1GiB_WRITE_SIZE", "WRITE_SIZE")
    copy = next(k for k in cf if kernel_tag(k) == "copy")
    f_scale = (1 << 30) / (cf[copy][1] * 1024)
    w_scale = (1 << 30) / (cw[copy][1] * 1024)
    lines = ["# HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", "",
             "Each workload's counters in runs of their own (repo root, one MI355X), written by",
             "`python3 bench.py --pmc-summary` from these commands (`python3 bench.py --pmc-commands`):", "", "```"]
    lines += [c for c in pmc_commands() if "--kernel-trace" not in c]
    lines += ["```", "",
              f"Calibration: a 1 GiB device-to-device copy (`{copy}`, 1,048,576 KiB read and written)",
              f"reads FETCH_SIZE = {cf[copy][1]:,.1f} KiB and WRITE_SIZE = {cw[copy][1]:,.1f} KiB per dispatch, so",
              f"fetched bytes = FETCH_SIZE x 1024 x {f_scale:.4f} and written bytes = WRITE_SIZE x 1024 x "
              f"{w_scale:.4f}.  Counters serialise the dispatches, so every kernel ran alone (CU masks kept).",
              "Averages per dispatch over each kernel's main launches (those with at least half the largest",
              "launch's count: a fold or a staging copy through the same kernel is far smaller), for every",
              "kernel that moves at least 1 MB; bench.py reads its contraction's",
              "`roofline.traffic` from the `HBM bytes` column (the largest row of its kernel tag).", "",
              "| workload | kernel | tag | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes |",
              "|---|---|---|---|---|---|---|"]
    for w in PMC_RUNS:
        if w == "copy_1GiB":
            continue
        fk = per_kernel(f"{PMC_OUT}/{w}_FETCH_SIZE", "FETCH_SIZE")
        wk = per_kernel(f"{PMC_OUT}/{w}_WRITE_SIZE", "WRITE_SIZE")
        rows = []
        for k in fk:
            if k not in wk:
                continue
            total = (fk[k][1] * f_scale + wk[k][1] * w_scale) * 1024
            if total >= 1e6:
                rows.append((total, k, fk[k][0], fk[k][1], wk[k][1]))
        for total, k, n, f, wr in sorted(rows, reverse=True):
            lines.append(f"| {w} | `{k.split('(')[0]}` | {kernel_tag(k)} | {n} | {f:,.1f} | {wr:,.1f} | {total:.4e} |")
    open(TRAFFIC_SUMMARY, "w").write("\n".join(lines) + "\n")

    ulines = ["# Unit utilisation of the contractions (rocprofv3 SQ counters)", "",
              "One counter run per workload (repo root, one MI355X), written by `python3 bench.py --pmc-summary`:",
              "", "```"]
    ulines += [c for c in pmc_commands() if "--kernel-trace" in c]
    ulines += ["```", "",
               f"Per dispatch, counters summed over the device; G = GRBM_GUI_ACTIVE / {XCCS} (one XCC's busy",
               f"clocks).  MfmaUtil = SQ_VALU_MFMA_BUSY_CYCLES / (G x {SIMDS} SIMDs); LdsUtil, VALUBusy and",
               f"LDS bank conflict = SQ_LDS_IDX_ACTIVE, SQ_ACTIVE_INST_VALU, SQ_LDS_BANK_CONFLICT / (G x {CUS}",
               f"CUs); busy CUs = SQ_BUSY_CU_CYCLES / (G x {CUS}); clock = G / kernel duration.  Averages",
               "over each kernel's main launches (at least half the longest one's duration); counters",
               "serialise dispatches (CU masks kept).", "",
               "| workload | kernel | tag | ms | clock GHz | MfmaUtil | LdsUtil | VALUBusy | LDS bank conflict "
               "| busy CUs |", "|---|---|---|---|---|---|---|---|---|---|"]
    for w in PMC_RUNS:
        if w == "copy_1GiB":
            continue
        acc = {}
        disp = [x for x in read_counters(f"{PMC_OUT}/{w}_units").values()
                if all(c in x[1] for c in UNIT_COUNTERS) and x[2] and x[1]["GRBM_GUI_ACTIVE"] > 0]
        longest = {}
        for kernel, _cnt, dur in disp:
            longest[kernel] = max(longest.get(kernel, 0), dur)
        for kernel, cnt, dur in disp:
            if dur < 0.5 * longest[kernel]:
                continue
            g = cnt["GRBM_GUI_ACTIVE"] / XCCS
            v = [dur / 1e6, g / dur, cnt["SQ_VALU_MFMA_BUSY_CYCLES"] / (g * SIMDS),
                 cnt["SQ_LDS_IDX_ACTIVE"] / (g * CUS), cnt["SQ_ACTIVE_INST_VALU"] / (g * CUS),
                 cnt["SQ_LDS_BANK_CONFLICT"] / (g * CUS), cnt["SQ_BUSY_CU_CYCLES"] / (g * CUS)]
            a = acc.setdefault(kernel, [0, [0.0] * len(v)])
            a[0] += 1
            a[1] = [x + y for x, y in zip(a[1], v)]
        rows = sorted(((a[1][0] / a[0], k, [x / a[0] for x in a[1]]) for k, a in acc.items()), reverse=True)
        for ms, k, v in rows:
            if kernel_tag(k) in ("gf_muladd_lds", "gf_muladd_mfma", "gf_muladd_mfma_i8", "gf_batch", "md5_slices",
                                 "crc32_slices", "md5_list", "crc32_list") and ms >= 0.05:
                ulines.append(f"| {w} | `{k.split('(')[0]}` | {kernel_tag(k)} | {v[0]:.3f} | {v[1]:.2f} | "
                              f"{100 * v[2]:.1f}% | {100 * v[3]:.1f}% | {100 * v[4]:.1f}% | {100 * v[5]:.1f}% | "
                              f"{100 * v[6]:.1f}% |")
    open(UNITS_SUMMARY, "w").write("\n".join(ulines) + "\n")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpus():
    """Host cores this process may use: its CPU affinity, capped by a cgroup
    CPU quota when one is set (a container's share of a larger host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    how = f"affinity {n}"
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            q = max(1, int(int(quota) // int(period)))
            if q < n:
                n, how = q, how + f", cgroup quota {q}"
    except (OSError, ValueError):
        pass
    return n, how


_CPU_BUILD = None


def cpu_port_build():
    """Load the oracle's restatement compiled for this host (-O3
    -march=native, the reference's benchmark flags) when the host has a C
    compiler; otherwise the shipped -O3 x86-64-v3 build.  The native build
    must agree with the shipped one on a small encode before it is used."""
    global _CPU_BUILD
    if _CPU_BUILD is not None:
        return _CPU_BUILD
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    probe = random_bytes(13, 64 * 4096).reshape(64, 4096).view(np.uint16)
    want = oracle.encode_avx2(probe, np.arange(7), threads=2)
    out = os.path.join(tempfile.mkdtemp(prefix="par2cpu_"), "liboracle_native.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"OUT={out}"], check=True,
                       capture_output=True, timeout=120)
        shipped = oracle.LIB_PATH
        oracle.use_library(out)
        if not np.array_equal(oracle.encode_avx2(probe, np.arange(7), threads=2), want):
            oracle.use_library(shipped)
            raise RuntimeError("native build disagrees with the shipped one")
        _CPU_BUILD = "gcc -O3 -march=native (built on this host)"
    except Exception as e:  # no compiler, or a build that does not agree
        log("cpu baseline: native build unavailable:", e)
        _CPU_BUILD = "gcc -O3 -march=x86-64-v3 (shipped build)"
    return _CPU_BUILD


def cpu_baseline_encode(n, s_sample, r, threads):
    """Reference algorithm restated (AVX2 nibble-shuffle muladd with the
    StreamingEncoder batching, plus per-slice MD5+CRC32), on n slices of
    s_sample bytes; returns source GB/s and the wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    data = random_bytes(11, n * s_sample).reshape(n, s_sample)
    oracle.encode_avx2(data[:, :64].view(np.uint16), np.arange(r), threads=threads)  # warm
    t0 = time.perf_counter()
    oracle.encode_avx2(data.view(np.uint16), np.arange(r), threads=threads)
    oracle.hash_slices_mt(data, n, s_sample, threads=threads)
    dt = time.perf_counter() - t0
    return n * s_sample / dt / 1e9, dt


def cpu_baseline_repair(n, s_sample, m, threads):
    """Repair on the CPU with the reference's algorithm restated: CRC32+MD5
    verify of every slice, then the rebuild contraction of m slices from n
    sources (same GF work as m recovery slices over n inputs)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    data = random_bytes(12, n * s_sample).reshape(n, s_sample)
    t0 = time.perf_counter()
    oracle.hash_slices_mt(data, n, s_sample, threads=threads)
    oracle.encode_avx2(data.view(np.uint16), np.arange(m), threads=threads)
    dt = time.perf_counter() - t0
    return n * s_sample / dt / 1e9, dt


def bench_encode(args, ctx):
    n, s, r = 1024, 1 << 20, 102
    data = random_bytes(1234, n * s)
    d_sl = ctx.alloc(n * s)
    d_rec = ctx.alloc(r * s)
    d_md5 = ctx.alloc(n * 16)
    d_crc = ctx.alloc(n * 4)
    d_sl.upload(data)
    exps = np.arange(r, dtype=np.uint32)
    if args.kernel:
        ctx.set_gf_kernel(args.kernel)
    for _ in range(args.warmup):
        ctx.encode_dev(d_sl, n, s, exps, d_rec, d_md5, d_crc)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_read()
    lat = []
    ctx.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.encode_dev(d_sl, n, s, exps, d_rec, d_md5, d_crc)
        ctx.sync()
        lat.append(time.perf_counter() - t0)
    ctx.sync()
    total = time.perf_counter() - t_start
    stats = ctx.profile_read()
    ctx.profile(False)
    # spot-check the result against the oracle on a column window (bit-exact)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    rec = d_rec.download().reshape(r, s).view(np.uint16)
    w0 = 4096
    cols = np.ascontiguousarray(data.reshape(n, s).view(np.uint16)[:, w0:w0 + 32])
    exact = bool(np.array_equal(rec[:, w0:w0 + 32], oracle.encode(cols, exps)))
    src_bytes = n * s
    algo_bytes = n * s + r * s  # read every input once, write every recovery slice once
    gf_name = next((k for k in ("gf_muladd_mfma", "gf_muladd_mfma_i8") if k in stats), "gf_muladd_lds")
    launches, gf_ms = stats.get(gf_name, (0, 0.0))
    gf_avg = gf_ms / max(launches, 1)
    md5_avg = stats.get("md5_slices", (1, 0.0))[1] / max(stats.get("md5_slices", (1, 0))[0], 1)
    crc_avg = stats.get("crc32_slices", (1, 0.0))[1] / max(stats.get("crc32_slices", (1, 0))[0], 1)
    achieved = algo_bytes / (gf_avg * 1e-3) / 1e9 if gf_avg > 0 else 0.0
    traffic, traffic_src = profile_traffic("plain_encode_1GiB", gf_name)
    if args.traffic is not None:
        traffic, traffic_src = args.traffic, "--traffic"
    units = profile_units("plain_encode_1GiB", gf_name)
    muladds = n * r * (s // 2)
    hbm = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5)}
    if gf_name.startswith("gf_muladd_mfma"):
        # the bit-sliced contraction is one (16R x 16N) x (16N x W) GF(2) GEMM on
        # the matrix cores: 256 binary multiply-adds, 512 operations, per GF
        # multiply-add; its roof is the dense FP4 (int8) rate
        i8 = gf_name.endswith("_i8")
        ops = 512.0 * muladds / (gf_avg * 1e-3) / 1e12 if gf_avg > 0 else 0.0
        peak = INT8_DENSE_PEAK_TOPS if i8 else FP4_DENSE_PEAK_TFLOPS
        roof = {"bound": "mfma", "achieved": round(ops, 1), "peak": peak, "unit": "TOP/s" if i8 else "TFLOP/s",
                "frac": round(ops / peak, 4), "hbm_achieved_GBps": round(achieved, 2),
                "hbm_frac": round(achieved / HBM_PEAK_GBS, 5)}
    else:
        roof = dict(hbm)
    return {
        "workload": "plain_encode",
        "config": {"workload": "plain_encode_1GiB", "n_slices": n, "slice_bytes": s, "recovery_slices": r,
                   "exponents": "0..101", "poly": "0x1100B", "hashes": "md5+crc32 per slice",
                   "resident": "hbm"},
        "src_bytes": src_bytes, "total_s": total, "lat": lat, "exact": exact,
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in stats.items()},
        "roofline": {**roof, "traffic": traffic,
                     "traffic_x_algorithmic": round(traffic / algo_bytes, 3) if traffic else None,
                     "traffic_source": traffic_src, "kernel": gf_name, "kernel_ms": round(gf_avg, 4), "algorithmic_bytes": algo_bytes,
                     "gf_muladds": muladds, **md5_floor_fields(n, s, md5_avg, ctx),
                     "crc32_ms": round(crc_avg, 4), "units": units},
        "cpu": (n, s, r),
    }


def bench_encode_host(args, ctx):
    """Plain encode from host memory: the same 1 GiB set (1024 x 1 MiB, R=102)
    in pageable host memory through par2hip_encode, which stages it through the
    context's pinned buffer, encodes with the IFSC hashes and copies the
    recovery slices and digests back into the caller's (reused) arrays.  Timed
    per call including both transfers: the rate a caller holding host buffers
    sees, not the device-resident kernel metric."""
    import ctypes as C
    import hashlib

    n, s, r = 1024, 1 << 20, 102
    data = random_bytes(1234, n * s).reshape(n, s)
    exps = np.arange(r, dtype=np.uint32)
    rec = np.ones((r, s), np.uint8)  # touched once: no page faults inside the timed calls
    md5 = np.ones((n, 16), np.uint8)
    crc = np.ones(n, np.uint32)
    lib = par2hip.lib()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)

    def call():
        rc = lib.par2hip_encode(ctx.handle, ptr(data), n, s, ptr(exps), r, ptr(rec), ptr(md5), ptr(crc))
        if rc != 0:
            raise RuntimeError(f"par2hip_encode failed ({rc})")

    for _ in range(args.warmup):
        call()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        call()
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    w0 = 98765
    cols = np.ascontiguousarray(data.view(np.uint16)[:, w0:w0 + 32])
    exact = bool(np.array_equal(rec.view(np.uint16)[:, w0:w0 + 32], oracle.encode(cols, exps)))
    exact = exact and all(md5[i].tobytes() == hashlib.md5(data[i].tobytes()).digest() for i in (0, 511, n - 1))
    return {
        "workload": "encode_host",
        "config": {"workload": "plain_encode_1GiB_from_host", "n_slices": n, "slice_bytes": s,
                   "recovery_slices": r, "exponents": "0..101", "hashes": "md5+crc32 per slice",
                   "resident": "host (pageable); both transfers inside the timed call"},
        "src_bytes": n * s, "total_s": total, "lat": lat, "exact": exact, "kernel_ms": {},
        "cpu": (n, s, r),
    }


def bench_micro(args, ctx):
    """Each kernel of the encode step alone on the plain-encode data (device-resident)."""
    n, s, r = 1024, 1 << 20, 102
    data = random_bytes(1234, n * s)
    d_sl = ctx.alloc(n * s)
    d_rec = ctx.alloc(r * s)
    d_md5 = ctx.alloc(n * 16)
    d_crc = ctx.alloc(n * 4)
    d_sl.upload(data)
    exps = np.arange(r, dtype=np.uint32)
    out = {}

    def timed(name, fn, reps):
        fn()
        ctx.sync()
        ctx.profile(True)
        ctx.profile_read()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        wall = (time.perf_counter() - t0) / reps * 1e3
        st = ctx.profile_read()
        ctx.profile(False)
        out[name] = {"wall_ms": round(wall, 4), **{k: round(v[1] / max(v[0], 1), 4) for k, v in st.items()}}
        log(name, out[name])

    tuned = ctx.set_autotune(False)
    for k, nm in ((1, "gf_lds"), (2, "gf_mfma"), (3, "gf_mfma_i8")):
        ctx.set_gf_kernel(k)
        timed(nm, lambda: ctx.encode_dev(d_sl, n, s, exps, d_rec), args.steps)
    ctx.set_gf_kernel(1)
    ctx.set_autotune(True)  # tuned on the warm-up call
    timed("gf_lds_tuned", lambda: ctx.encode_dev(d_sl, n, s, exps, d_rec), args.steps)
    ctx.set_autotune(False)
    ctx.set_gf_kernel(0)
    timed("md5_only", lambda: ctx.hash_slices_dev(d_sl, n, s, d_md5, None), max(2, args.steps // 4))
    # the repair-shaped contraction: 2000 sources x 100 outputs x 256 Ki words
    n3, s3, r3 = 2000, 512 << 10, 100
    d_sl3 = ctx.alloc(n3 * s3)
    d_rec3 = ctx.alloc(r3 * s3)
    d_sl3.upload(random_bytes(99, n3 * s3))
    exps3 = np.arange(r3, dtype=np.uint32)
    for k, nm in ((1, "gf_lds_2000x100"), (2, "gf_mfma_2000x100"), (3, "gf_mfma_i8_2000x100")):
        ctx.set_gf_kernel(k)
        timed(nm, lambda: ctx.encode_dev(d_sl3, n3, s3, exps3, d_rec3), args.steps)
    ctx.set_gf_kernel(1)
    ctx.set_autotune(True)
    timed("gf_lds_2000x100_tuned", lambda: ctx.encode_dev(d_sl3, n3, s3, exps3, d_rec3), args.steps)
    ctx.set_autotune(tuned)
    ctx.set_gf_kernel(0)
    d_sl3.free()
    d_rec3.free()
    timed("crc32_only", lambda: ctx.hash_slices_dev(d_sl, n, s, None, d_crc), args.steps)
    return out


def bench_small_r(args, ctx):
    """Few recovery slices over the plain-encode input (1024 x 1 MiB in HBM):
    the contraction alone for R = 1, 2, 4, 8 with the LDS-table and the FP4
    matrix-core forms.  With so few outputs each launch is bound by reading
    the 1 GiB once, so the roofline is HBM: (N*S + R*S) / kernel time."""
    n, s = 1024, 1 << 20
    data = random_bytes(1234, n * s)
    d_sl = ctx.alloc(n * s)
    d_sl.upload(data)
    d_rec = ctx.alloc(8 * s)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    out = {}
    prev = ctx.set_gf_kernel(0)
    for r in args.small_r:
        exps = np.arange(r, dtype=np.uint32)
        cols = np.ascontiguousarray(data.reshape(n, s).view(np.uint16)[:, 1000:1032])
        want = oracle.encode(cols, exps)
        for k, name in ((1, "lds"), (2, "mfma_fp4")):
            ctx.set_gf_kernel(k)
            ctx.encode_dev(d_sl, n, s, exps, d_rec)
            ctx.sync()
            exact = bool(np.array_equal(d_rec.download(r * s).reshape(r, s).view(np.uint16)[:, 1000:1032], want))
            ctx.profile(True)
            ctx.profile_read()
            for _ in range(args.steps):
                ctx.encode_dev(d_sl, n, s, exps, d_rec)
            ctx.sync()
            st = ctx.profile_read()
            ctx.profile(False)
            gf = next(v for key, v in st.items() if key.startswith("gf_muladd"))
            ms = gf[1] / max(gf[0], 1)
            algo = n * s + r * s
            out[f"R{r}_{name}"] = {"kernel_ms": round(ms, 4), "GBps": round(algo / (ms * 1e-3) / 1e9, 1),
                                   "hbm_frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "exact": exact}
            if r == 4:  # the counter-profiled case
                tag = next(key for key in st if key.startswith("gf_muladd"))
                traffic, src = profile_traffic("small_r_R4_1GiB", tag)
                out[f"R{r}_{name}"].update(traffic=traffic, traffic_source=src,
                                           traffic_x_algorithmic=round(traffic / algo, 3) if traffic else None)
            log("small R", r, name, out[f"R{r}_{name}"])
    ctx.set_gf_kernel(prev)
    return out


def bench_repair(args, ctx):
    """Plain repair: 2000 x 512 KiB slices (~1 GiB), 100 erased at seeded
    positions, 100 recovery slices present.  One step = par2hip_repair_set_dev:
    CRC32 + MD5 verify of every slice against its IFSC entry, Gauss-Jordan of
    the 100 x 100 erasure matrix, fold, and the rebuild contraction."""
    n, s, m = 2000, 512 << 10, 100
    data = random_bytes(4321, n * s).reshape(n, s)
    exps = np.arange(m, dtype=np.uint32)
    d_sl = ctx.alloc(n * s)
    d_rec = ctx.alloc(m * s)
    d_out = ctx.alloc(m * s)
    d_md5 = ctx.alloc(n * 16)
    d_crc = ctx.alloc(n * 4)
    d_sl.upload(data)
    ctx.encode_dev(d_sl, n, s, exps, d_rec, d_md5, d_crc)  # IFSC entries + recovery of the intact set
    ctx.sync()
    md5 = d_md5.download().reshape(n, 16)
    crc = d_crc.download().view(np.uint32)
    rng = np.random.default_rng(77)
    missing = np.sort(rng.choice(n, m, replace=False)).astype(np.uint32)
    damaged = data.copy()
    damaged[missing] = 0
    d_sl.upload(damaged)
    if args.kernel:
        ctx.set_gf_kernel(args.kernel)

    def step():
        return ctx.repair_set_dev(d_sl, n, s, md5, crc, d_rec, exps, d_out)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.profile(True)
    ctx.profile_read()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        got = step()
        ctx.sync()
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    stats = ctx.profile_read()
    ctx.profile(False)
    out = d_out.download().reshape(m, s)
    exact = bool(np.array_equal(got, missing) and np.array_equal(out, data[missing]))
    gf_name = next((k for k in ("gf_muladd_mfma", "gf_muladd_mfma_i8") if k in stats), "gf_muladd_lds")
    launches, gf_ms = stats.get(gf_name, (0, 0.0))
    gf_avg = gf_ms / max(launches, 1)
    algo_bytes = n * s + m * s + m * s
    achieved = algo_bytes / (gf_avg * 1e-3) / 1e9 if gf_avg > 0 else 0.0
    traffic, traffic_src = profile_traffic("plain_repair_1GiB", gf_name)
    if args.traffic is not None:
        traffic, traffic_src = args.traffic, "--traffic"
    return {
        "workload": "plain_repair",
        "config": {"workload": "plain_repair_1GiB", "n_slices": n, "slice_bytes": s, "erasures": m,
                   "recovery_slices": m, "resident": "hbm", "verify": "crc32+md5 of every slice vs IFSC"},
        "src_bytes": n * s, "total_s": total, "lat": lat, "exact": exact,
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in stats.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_x_algorithmic": round(traffic / algo_bytes, 3) if traffic else None,
                     "traffic_source": traffic_src, "kernel": gf_name, "kernel_ms": round(gf_avg, 4),
                     "algorithmic_bytes": algo_bytes, "gf_muladds": m * n * (s // 2),
                     **md5_floor_fields(n, s, stats.get("md5_slices", (1, 0.0))[1] /
                                        max(stats.get("md5_slices", (1, 0))[0], 1), ctx),
                     "units": profile_units("plain_repair_1GiB", gf_name)},
        "cpu": (n, s, m),
    }


def bench_batched(args, ctx):
    """256 PAR2 sets of uneven shapes (half create, half repair) in one batched
    run vs the plain path (one set at a time), both with IFSC hashes.  A step
    replays the batch planned once and captured as a HIP graph; the eager
    form (planning on every call) is reported beside it."""
    sets, total = draw_sets(args.seed)
    pool = random_bytes(5, 1 << 30)  # host pool; each set takes a window of it
    rng = np.random.default_rng(6)
    jobs = []
    log("batched workload:", len(sets), "sets,", round(total / 2**30, 2), "GiB of source")
    for x in sets:
        n, s = x["n"], x["s"]
        nb = n * s
        off = int(rng.integers(0, len(pool) - min(nb, len(pool)) + 1))
        d_sl = ctx.alloc(nb)
        if nb <= len(pool):
            d_sl.upload(pool[off:off + nb])
        else:
            for o in range(0, nb, len(pool)):
                d_sl.upload(pool[: min(len(pool), nb - o)], offset=o)
        job = {"kind": x["kind"], "slices": d_sl, "n": n, "slice_size": s,
               "md5": ctx.alloc(n * 16), "crc": ctx.alloc(n * 4)}
        if x["kind"] == "encode":
            job.update(exps=np.arange(x["r"], dtype=np.uint32), recovery=ctx.alloc(x["r"] * s))
        else:
            m = x["m"]
            d_rec = ctx.alloc(m * s)
            ctx.encode_dev(d_sl, n, s, np.arange(m, dtype=np.uint32), d_rec)
            missing = np.sort(rng.choice(n, m, replace=False)).astype(np.uint32)
            job.update(exps=np.arange(m, dtype=np.uint32), recovery=d_rec, missing=missing, out=ctx.alloc(m * s))
        jobs.append(job)
    ctx.sync()
    # the erased slices of repair jobs are left in place: the rebuild ignores them

    def plain():
        for j in jobs:
            if j["kind"] == "encode":
                ctx.encode_dev(j["slices"], j["n"], j["slice_size"], j["exps"], j["recovery"], j["md5"], j["crc"])
            else:
                ctx.hash_slices_dev(j["slices"], j["n"], j["slice_size"], j["md5"], j["crc"])
                ctx.reconstruct_dev(j["slices"], j["n"], j["slice_size"], j["missing"], j["recovery"], j["exps"],
                                    j["out"])
            ctx.sync()

    def batched():
        ctx.batch_run_dev(jobs)
        ctx.sync()

    def timed(fn, reps):
        fn()
        lat = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            lat.append(time.perf_counter() - t0)
        return lat

    def outputs_equal(ref):
        return all(np.array_equal(a, (j["recovery"].download() if j["kind"] == "encode" else j["out"].download()))
                   and np.array_equal(b, j["md5"].download()) and np.array_equal(c, j["crc"].download())
                   for j, (a, b, c) in zip(jobs, ref))

    def clear():
        for j in jobs:
            (j["recovery"] if j["kind"] == "encode" else j["out"]).upload(np.zeros(16, np.uint8))

    tuning = {}

    def graph_replay():
        plan = ctx.batch_plan(jobs, graph=True)
        tuning.update(plan.tuning())

        def replay():
            plan.launch()
            ctx.sync()

        clear()
        replay()
        ok = outputs_equal(ref)
        lat = timed(replay, args.steps)
        plan.close()
        return ok, lat

    was_tuned = ctx.set_autotune(False)
    # bit-exactness: batched outputs equal the plain path's for every job
    plain()
    ref = [(j["recovery"].download() if j["kind"] == "encode" else j["out"].download(), j["md5"].download(),
            j["crc"].download()) for j in jobs]
    clear()
    batched()
    exact = outputs_equal(ref)
    ctx.profile(True)
    ctx.profile_read()
    lat_e = timed(batched, args.steps)
    stats = ctx.profile_read()
    ctx.profile(False)
    # the same batch planned once and replayed as one HIP graph per step, with
    # the fixed tile shapes, then planned with autotuning: the single-set
    # tuner first fills the tuned table on the sets themselves (one of the
    # policies the plan times), then the plan times every policy on the batch
    # and keeps the fastest
    exact_fixed, lat_fixed = graph_replay()
    ctx.set_autotune(True)
    plain()
    exact_plain_tuned = outputs_equal(ref)
    exact_tuned, lat_b = graph_replay()
    lat_tuned = lat_b
    lat_p = timed(plain, max(1, args.steps // 2))
    ctx.set_autotune(was_tuned)
    exact_graph = exact_fixed and exact_tuned and exact_plain_tuned
    for j in jobs:  # the sets' device memory goes back before the next workload
        for k in ("slices", "md5", "crc", "recovery", "out"):
            if k in j:
                j[k].free()
    # the batch contraction reads every set's sources once and writes its
    # outputs once (encode: n*s in, r*s out; repair: the n-m present and m
    # recovery slices in, m rebuilt out)
    algo = sum(x["n"] * x["s"] + (x["r"] if x["kind"] == "encode" else x["m"]) * x["s"] for x in sets)
    muladds = sum((x["r"] if x["kind"] == "encode" else x["m"]) * x["n"] * (x["s"] // 2) for x in sets)
    gf_ms = stats.get("gf_batch", (1, 0.0))[1] / max(stats.get("gf_batch", (1, 0))[0], 1)
    achieved = algo / (gf_ms * 1e-3) / 1e9 if gf_ms > 0 else 0.0
    traffic, traffic_src = profile_traffic("batched_256_sets", "gf_batch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_x_algorithmic": round(traffic / algo, 3) if traffic else None,
                "traffic_source": traffic_src, "kernel": "gf_batch", "kernel_ms": round(gf_ms, 4),
                "algorithmic_bytes": algo, "gf_muladds": muladds,
                "gf_muladds_per_s": float(f"{muladds / (gf_ms * 1e-3):.4g}") if gf_ms > 0 else None,
                "units": profile_units("batched_256_sets", "gf_batch")}
    return {
        "workload": "batched",
        "roofline": roofline,
        "config": {"workload": "batched_256_sets", "sets": len(sets), "source_GiB": round(total / 2**30, 3),
                   "encode_sets": sum(x["kind"] == "encode" for x in sets),
                   "repair_sets": sum(x["kind"] == "repair" for x in sets), "seed": args.seed,
                   "hashes": "md5+crc32 of every input slice", "resident": "hbm"},
        "src_bytes": total, "total_s": float(np.sum(lat_b)), "lat": lat_b, "exact": bool(exact and exact_graph),
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in stats.items()},
        "eager_ms": float(np.mean(lat_e) * 1e3),
        "plain_ms": float(np.mean(lat_p) * 1e3), "plain_gbs": total / float(np.mean(lat_p)) / 1e9,
        "tile_shapes": {"tuned_ms": round(float(np.mean(lat_tuned) * 1e3), 3),
                        "fixed_ms": round(float(np.mean(lat_fixed) * 1e3), 3),
                        "tuned_speedup": round(float(np.mean(lat_fixed) / np.mean(lat_tuned)), 4),
                        "policy_chosen": ["default", "tuned_table", "one_group", "small_v"][tuning["policy"]],
                        "policy_ms": dict(zip(["default", "tuned_table", "one_group", "small_v"], tuning["ms"]))},
    }


def bench_served(args, ctx):
    """Served workload (Poisson arrivals, continuous batching through the
    serving loop) against the plain per-request path (one request at a time,
    single-set entry points) on the same request stream."""
    import threading

    rng = np.random.default_rng(args.seed + 100)
    template = draw_requests(args.seed, 3000)
    pool_bytes = 16 << 30
    d_pool = ctx.alloc(pool_bytes)
    block = random_bytes(21, 1 << 30)
    for o in range(0, pool_bytes, len(block)):
        d_pool.upload(block, offset=o)
    del block
    ctx.sync()
    # every request reads its slices (and, for repairs, its recovery rows) as
    # windows of the device pool; erasure positions are drawn per request
    for q in template:
        nb = q["n"] * q["s"]
        q["off"] = int(rng.integers(0, (pool_bytes - nb) // 256)) * 256
        q["exps"] = np.arange(q["k"], dtype=np.uint32)
        if q["kind"] == "repair":
            q["rec_off"] = int(rng.integers(0, (pool_bytes - q["k"] * q["s"]) // 256)) * 256
            q["missing"] = np.sort(rng.choice(q["n"], q["k"], replace=False)).astype(np.uint32)
    max_out = max(q["k"] * q["s"] for q in template)
    max_n = max(q["n"] for q in template)
    d_out, d_md5, d_crc = ctx.alloc(max_out), ctx.alloc(max_n * 16), ctx.alloc(max_n * 4)

    def view(off):
        return d_pool.view(off)

    def job_of(q):
        j = {"kind": q["kind"], "slices": view(q["off"]), "n": q["n"], "slice_size": q["s"], "exps": q["exps"]}
        if q["kind"] == "repair":
            j.update(recovery=view(q["rec_off"]), missing=q["missing"])
        return j

    def plain_one(q):
        n, s = q["n"], q["s"]
        if q["kind"] == "encode":
            ctx.encode_dev(view(q["off"]), n, s, q["exps"], d_out, d_md5, d_crc)
        else:
            ctx.hash_slices_dev(view(q["off"]), n, s, d_md5, d_crc)
            ctx.reconstruct_dev(view(q["off"]), n, s, q["missing"], view(q["rec_off"]), q["exps"], d_out)
        ctx.sync()

    def run_plain(reqs, rate):
        """Real-time arrivals at `rate`/s, served one at a time in order."""
        lat, t0 = [], time.perf_counter()
        for q in reqs:
            arr = t0 + q["t"] / rate
            now = time.perf_counter()
            if now < arr:
                time.sleep(arr - now)
            plain_one(q)
            lat.append(time.perf_counter() - arr)
        span = time.perf_counter() - t0
        return np.array(lat) * 1e3, span

    srv = par2hip.Server(0, hbm_budget=args.hbm_budget, lanes=args.lanes, max_batch=args.max_batch)

    def run_served(reqs, rate):
        tickets = [None] * len(reqs)
        base = srv.now_ms()
        t0 = time.perf_counter()

        def submitter():
            for i, q in enumerate(reqs):
                arr = t0 + q["t"] / rate
                now = time.perf_counter()
                if now < arr:
                    time.sleep(arr - now)
                tickets[i] = srv.submit(job_of(q), hashes=True, arrival_ms=base + q["t"] / rate * 1e3)

        th = threading.Thread(target=submitter)
        th.start()
        lat, sizes, last = [], [], 0.0
        for i in range(len(reqs)):
            while tickets[i] is None:
                time.sleep(0.0002)
            res = srv.wait(tickets[i])
            lat.append(res["t_done_ms"] - res["t_arrival_ms"])
            sizes.append(res["batch_size"])
            last = max(last, res["t_done_ms"])
            srv.release(tickets[i])
        th.join()
        span = (last - base) / 1e3
        return np.array(lat), span, float(np.mean(sizes))

    # bit-exactness: served outputs equal the plain path's for a sample
    sample = template[:24]
    exact = True
    for q in sample:
        plain_one(q)
        rows = q["k"] * q["s"]
        want = (d_out.download(rows), d_md5.download(q["n"] * 16), d_crc.download(q["n"] * 4))
        t = srv.submit(job_of(q), hashes=True)
        res = srv.wait(t)
        got = (par2hip.d2h(res["d_out"], rows, ctx), par2hip.d2h(res["d_md5"], q["n"] * 16, ctx),
               par2hip.d2h(res["d_crc"], q["n"] * 4, ctx))
        exact &= all(np.array_equal(a, b) for a, b in zip(want, got))
        srv.release(t)

    # plain capacity: requests back to back
    cap_reqs = template[:150]
    t0 = time.perf_counter()
    for q in cap_reqs:
        plain_one(q)
    t_service = (time.perf_counter() - t0) / len(cap_reqs)
    rate_p = 0.7 / t_service
    n_p = args.served_requests
    lat_p, span_p = run_plain(template[:n_p], rate_p)
    bytes_p = sum(q["n"] * q["s"] for q in template[:n_p])
    p99_p = float(np.percentile(lat_p, 99))
    log(f"plain: service {t_service * 1e3:.2f} ms/request, rate {rate_p:.1f}/s, p50 "
        f"{np.percentile(lat_p, 50):.1f} ms p99 {p99_p:.1f} ms, {bytes_p / span_p / 1e9:.2f} GB/s")
    runs = []
    for k in args.served_loads:
        rate = k * rate_p
        n_s = min(len(template), int(args.served_requests * max(1.0, k / 2)))
        lat_s, span_s, mean_batch = run_served(template[:n_s], rate)
        bytes_s = sum(q["n"] * q["s"] for q in template[:n_s])
        r = {"load_x_plain": k, "rate_per_s": round(rate, 1), "requests": n_s,
             "p50_ms": round(float(np.percentile(lat_s, 50)), 2), "p99_ms": round(float(np.percentile(lat_s, 99)), 2),
             "GBps": round(bytes_s / span_s / 1e9, 3), "mean_batch": round(mean_batch, 2)}
        log("served", r)
        runs.append(r)
        if r["p99_ms"] > p99_p:
            break
    ok = [r for r in runs if r["p99_ms"] <= p99_p]
    best = max(ok, key=lambda r: r["GBps"]) if ok else runs[0]
    st = srv.stats()
    srv.close()
    d_pool.free()
    mean_req = float(np.mean([q["n"] * q["s"] for q in template]))
    return {
        "workload": "served",
        "config": {"workload": "served_poisson", "request_size": "lognormal median 64 MiB sigma 1.0 cap 8 GiB",
                   "mean_request_MiB": round(mean_req / 2**20, 1), "slice_bytes": "U{256 KiB..1 MiB} step 4 B",
                   "repairs": "30%, erasures ceil(U[1,10]% N)", "creates": "redundancy U[1,20]%",
                   "hbm_budget_GB": round(args.hbm_budget / 1e9, 1), "lanes": args.lanes,
                   "max_batch": args.max_batch, "seed": args.seed, "hashes": "md5+crc32 of every input slice",
                   "resident": "hbm (device input pool)"},
        "value": best["GBps"], "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "exact": bool(exact),
        "served": runs, "server_stats": st,
        "plain_path": {"rate_per_s": round(rate_p, 1), "service_ms": round(t_service * 1e3, 3),
                       "p50_ms": round(float(np.percentile(lat_p, 50)), 2), "p99_ms": round(p99_p, 2),
                       "GBps": round(bytes_p / span_p / 1e9, 3),
                       "speedup_at_equal_p99": round(best["GBps"] / (bytes_p / span_p / 1e9), 3)},
    }


def bench_inverse(args, ctx):
    """The device inverse of m x m erasure matrices with gapped exponents
    (par2hip_gf_invert: the blocked Gauss-Jordan above 256, one workgroup
    below), one call per size, its kernel times from HIP events."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    glog, gexp = oracle.tables()
    out = {}
    for m in args.inverse_sizes:
        rng = np.random.default_rng(m)
        c = oracle.par2_constants(32768)
        lc = glog[c[rng.choice(32768, m, replace=False)]].astype(np.uint64)
        # gapped exponents as a repair with lost volumes has them; a generalised
        # Vandermonde matrix with arbitrary exponents can be singular, so they
        # are drawn from a range that leaves few gaps (consecutive at the top
        # size, where 3m exceeds the exponent range)
        top = 3 * m + 100
        exps = (np.sort(rng.choice(top, m, replace=False)) if top <= 65535 else np.arange(m)).astype(np.uint64)
        A = np.empty((m, m), np.uint16)
        for i0 in range(0, m, 1024):  # row blocks: no m x m 64-bit temporary
            A[i0:i0 + 1024] = gexp[(exps[i0:i0 + 1024, None] * lc[None, :]) % 65535]
        ctx.profile(True)
        ctx.profile_read()
        t0 = time.perf_counter()
        try:
            inv = ctx.gf_invert(A)
        except par2hip.RepairFailed:
            out[str(m)] = {"singular": True}
            ctx.profile(False)
            continue
        wall = time.perf_counter() - t0
        st = ctx.profile_read()
        ctx.profile(False)
        # row 0 of A * A^-1 must be e_0 (table product; the tests use the carry-less one)
        la = glog[A[0]].astype(np.int64)
        row = np.zeros(m, np.uint16)
        for j0 in range(0, m, 1024):
            li = glog[inv[:, j0:j0 + 1024]].astype(np.int64)
            prod = np.where((A[0][:, None] != 0) & (inv[:, j0:j0 + 1024] != 0), gexp[(la[:, None] + li) % 65535], 0)
            row[j0:j0 + 1024] = np.bitwise_xor.reduce(prod.astype(np.uint16), axis=0)
        e0 = np.zeros(m, np.uint16)
        e0[0] = 1
        out[str(m)] = {"wall_ms": round(wall * 1e3, 2), "row0_identity": bool(np.array_equal(row, e0)),
                       **{k: round(v[1], 3) for k, v in st.items()}}
        log("inverse", m, out[str(m)])
        del A, inv
    return out


def bench_create(args, ctx):
    """End-to-end PAR2 creation (Par2Creator::create, creator.rs:261-531)
    from host memory to written .par2 files, exponential volumes.  Default:
    16 files of 64 MiB (1 GiB), the reference's automatic block size, 10%
    redundancy.  --create-files 1 --create-file-mib 1024 --create-block
    1048576 is the plain-encode set as the creator makes it: 1024 blocks of
    1 MiB, and -r 10 gives R = 103 by the creator's rounding (creator.rs:
    320-335), one more than the 102 of the kernel workload.  One step = one
    par2hip_create_par2 call (file digests on host threads overlapped with
    H2D, IFSC hashes, recovery slices and recovery packet digests on the
    device, file writes)."""
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    import par2fmt

    nfiles, fbytes = args.create_files, args.create_file_mib << 20
    block, pct = args.create_block, args.create_redundancy
    pool = random_bytes(31, nfiles * fbytes)
    files = [(f"part{i:02d}.bin", pool[i * fbytes:(i + 1) * fbytes].tobytes()) for i in range(nfiles)]
    out_dir = tempfile.mkdtemp(prefix="par2create_")
    stem = os.path.join(out_dir, "set")
    # exactness: the C1 configuration against the oracle's creator, byte for byte
    c1 = random_bytes(41, 4 << 20).tobytes()
    got = ctx.create_par2([("c1.bin", c1)], os.path.join(out_dir, "c1"), block_size=65536, redundancy=5.0)
    want = oracle.par2_create([("c1.bin", c1)], block_size=65536, redundancy=5.0, stem=os.path.join(out_dir, "c1"))
    exact = [p for p, _ in got] == [n for n, _ in want] and all(open(p, "rb").read() == b for (p, _), (_, b) in
                                                                 zip(got, want))
    written = ctx.create_par2(files, stem, block_size=block, redundancy=pct)  # warm-up
    lat = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        written = ctx.create_par2(files, stem, block_size=block, redundancy=pct)
        lat.append(time.perf_counter() - t0)
    # structure and a column window of the recovery slices of the 1 GiB set
    st = par2fmt.load([p for p, _ in written])
    bs = st.slice_size
    blocks_per_file = -(-fbytes // bs)
    order = [st.files[f].name for f in st.file_ids]
    exact &= st.bad_packets == 0 and sorted(order) == sorted(n for n, _ in files)
    byname = dict(files)
    import hashlib
    import zlib
    for fid in st.file_ids[:2]:
        d = st.files[fid]
        exact &= d.md5 == hashlib.md5(byname[d.name]).digest()
        blk = byname[d.name][:bs]
        exact &= st.ifsc[fid][0] == (hashlib.md5(blk).digest(), zlib.crc32(blk))
    w0, nw = 1024, 32
    cols = np.stack([np.frombuffer(byname[name][k * bs:k * bs + bs].ljust(bs, b"\0"), np.uint16)[w0:w0 + nw]
                     for name in order for k in range(blocks_per_file)])
    r = len(st.recovery)
    ref = oracle.encode(np.ascontiguousarray(cols), np.arange(r))
    exact &= all(np.array_equal(np.frombuffer(st.recovery[e], np.uint16)[w0:w0 + nw], ref[e]) for e in range(r))
    total_out = sum(sz for _, sz in written)
    n_blocks = nfiles * blocks_per_file
    for p, _ in written + got:
        os.remove(p)
    os.rmdir(out_dir)
    src = nfiles * fbytes
    res = {
        "workload": "create",
        "config": {"workload": f"create_{nfiles}x{fbytes >> 20}MiB", "files": nfiles, "file_bytes": fbytes,
                   "block_size": bs, "block_rule": "given" if block else "automatic (creator.rs:29-67)",
                   "input_slices": n_blocks, "recovery_slices": r, "redundancy_percent": pct,
                   "volumes": "exponential", "output_bytes": total_out, "source": "host memory -> .par2 files"},
        "src_bytes": src, "total_s": float(np.sum(lat)), "lat": lat, "exact": bool(exact), "kernel_ms": {},
    }
    if not args.no_cpu:
        # the reference's method on the host: AVX2 encode with the
        # StreamingEncoder batching, IFSC hashing and the whole-file MD5s,
        # all multithreaded; packet assembly and writes not included
        threads, how = host_cpus()
        build = cpu_port_build()
        data = np.concatenate([np.frombuffer(b, np.uint8) for _, b in files])
        nb = n_blocks
        padded = np.zeros(nb * bs, np.uint8)
        for i in range(nfiles):
            padded[i * blocks_per_file * bs:i * blocks_per_file * bs + fbytes] = data[i * fbytes:(i + 1) * fbytes]
        from concurrent.futures import ThreadPoolExecutor
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            digests = ex.map(lambda f: hashlib.md5(f[1]).digest(), files)
            oracle.encode_avx2(padded.reshape(nb, bs).view(np.uint16), np.arange(r), threads=threads)
            oracle.hash_slices_mt(padded.reshape(nb, bs), nb, bs, threads=threads)
            list(digests)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(src / dt / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
                               "cores_from": how, "build": build,
                               "sample": f"the same {src / 2**30:.3g} GiB set: AVX2 encode of {nb} x {bs} B into {r} recovery "
                                         f"slices + IFSC MD5/CRC32 + whole-file MD5s, {dt:.2f} s wall"}
    return res


def bench_repair_files(args, ctx):
    """End-to-end repair from files (Par2Repairer: parse the index and its
    volumes, verify every protected file, rebuild, commit through .par2tmp
    files; src/parser.rs:181-547, src/verify.rs:87-330, src/repair.rs:445-1125).
    The set (16 files of 64 MiB by default, automatic block size, 10%) is
    created once; before each step (untimed) 100 blocks spread over half the
    files are overwritten on disk.  One step = one par2hip_repair_files call:
    loading, device verification of all files, the rebuild, the temp files
    written, verified and renamed."""
    import tempfile

    nfiles, fbytes = args.create_files, args.create_file_mib << 20
    pool = random_bytes(31, nfiles * fbytes)
    files = [(f"part{i:02d}.bin", pool[i * fbytes:(i + 1) * fbytes].tobytes()) for i in range(nfiles)]
    out_dir = tempfile.mkdtemp(prefix="par2repair_")
    for name, b in files:
        with open(os.path.join(out_dir, name), "wb") as f:
            f.write(b)
    written = ctx.create_par2(files, os.path.join(out_dir, "set"), block_size=args.create_block,
                              redundancy=args.create_redundancy)
    index = written[0][0]
    info = par2hip.par2_info(index)
    bs = info["block_size"]
    per_file = -(-fbytes // bs)
    m = min(100, info["recovery_slices"])
    rng = np.random.default_rng(7)
    damaged_files = list(range(0, nfiles, 2)) or [0]

    def damage():
        picks = set()
        while len(picks) < m:
            picks.add((int(rng.choice(damaged_files)), int(rng.integers(0, per_file))))
        for fi, k in picks:
            with open(os.path.join(out_dir, files[fi][0]), "r+b") as f:
                f.seek(k * bs)
                n = min(bs, fbytes - k * bs)
                f.write(bytes(x ^ 0xA5 for x in files[fi][1][k * bs:k * bs + min(n, 64)]))
        return picks

    damage()
    ctx.repair_files(index)  # warm-up
    lat, exact = [], True
    for _ in range(args.steps):
        picks = damage()
        t0 = time.perf_counter()
        rep = ctx.repair_files(index)
        lat.append(time.perf_counter() - t0)
        exact &= rep["repaired"] == 1 and rep["blocks_damaged"] == len(picks)
        for fi in {fi for fi, _ in picks}:
            with open(os.path.join(out_dir, files[fi][0]), "rb") as f:
                exact &= f.read() == files[fi][1]
    for name in os.listdir(out_dir):
        os.remove(os.path.join(out_dir, name))
    os.rmdir(out_dir)
    src = nfiles * fbytes
    res = {
        "workload": "repair_files",
        "config": {"workload": f"repair_files_{nfiles}x{fbytes >> 20}MiB", "files": nfiles, "file_bytes": fbytes,
                   "block_size": bs, "blocks": info["blocks"], "recovery_slices": info["recovery_slices"],
                   "damaged_blocks": m, "damaged_files": len(damaged_files),
                   "source": ".par2 files + damaged files -> repaired files"},
        "src_bytes": src, "total_s": float(np.sum(lat)), "lat": lat, "exact": bool(exact), "kernel_ms": {},
    }
    if not args.no_cpu:
        # the reference's method on the host for the same set: MD5/CRC32 of
        # every block and the rebuild contraction of m blocks (file reads,
        # parsing and writes not included)
        threads, how = host_cpus()
        build = cpu_port_build()
        nb = nfiles * per_file
        value, dt = cpu_baseline_repair(nb, bs, m, threads)
        res["cpu_baseline"] = {"value": round(value, 4), "unit": "GB/s", "cores": threads, "kind": "port",
                               "cores_from": how, "build": build,
                               "sample": f"{nb} blocks x {bs} B: MD5/CRC32 verify of every block + rebuild "
                                         f"contraction of {m} blocks (AVX2 nibble-shuffle), {dt:.2f} s wall"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["encode", "repair", "micro", "batched", "served", "copy", "create",
                                           "repair_files", "encode_host", "inverse", "small_r"],
                    default="encode")
    ap.add_argument("--hbm-budget", type=int, default=200 * 10**9, help="served: server HBM budget in bytes")
    ap.add_argument("--lanes", type=int, default=2, help="served: concurrent batches")
    ap.add_argument("--max-batch", type=int, default=512, help="served: requests per batch at most")
    ap.add_argument("--served-requests", type=int, default=400, help="served: requests per plain run")
    ap.add_argument("--served-loads", type=float, nargs="+", default=[6, 8, 10, 11, 12, 13, 14],
                    help="served: offered loads as multiples of the plain run's")
    ap.add_argument("--create-files", type=int, default=16, help="create / repair_files: input files")
    ap.add_argument("--create-file-mib", type=int, default=64, help="create / repair_files: MiB per input file")
    ap.add_argument("--create-block", type=int, default=0, help="create / repair_files: block size in bytes (0: automatic)")
    ap.add_argument("--create-redundancy", type=float, default=10.0, help="create / repair_files: redundancy percent")
    ap.add_argument("--small-r", type=int, nargs="+", default=[1, 2, 4, 8], help="small_r: recovery counts")
    ap.add_argument("--inverse-sizes", type=int, nargs="+", default=[1000, 5000, 12000],
                    help="inverse: erasure counts m (up to 32768)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 LDS tables, 2 MFMA FP4, 3 MFMA int8")
    ap.add_argument("--autotune", type=int, default=None,
                    help="1: tile shapes tuned on the first (warm-up) call of each shape class; 0: fixed heuristic "
                         "(default: 0 for encode/repair, whose steps the MD5 chains bound, so the contraction's "
                         "shape and its measured HBM traffic stay fixed; 1 otherwise)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--legs", nargs="*", default=["repair", "batched", "served"],
                    choices=["repair", "batched", "served"],
                    help="encode: other workloads measured in the same run and reported as sub-objects "
                         "(--legs with no value: none)")
    ap.add_argument("--leg-steps", type=int, default=5, help="steps of the batched leg")
    ap.add_argument("--pmc-commands", action="store_true",
                    help="print the rocprofv3 counter runs behind profiles/hbm_traffic_pmc.md and "
                         "profiles/gf_units_utilisation_pmc.md, one per line, and exit")
    ap.add_argument("--pmc-summary", action="store_true",
                    help="write those two summaries from the runs' output under gpurun_out/pmc and exit")
    ap.add_argument("--trace-commands", action="store_true",
                    help="print the rocprofv3 --kernel-trace --stats runs behind profiles/*_kernel_trace.md")
    ap.add_argument("--trace-summary", action="store_true",
                    help="write profiles/<name>_kernel_trace.md from those runs' output under gpurun_out/trace")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per GF launch measured with rocprofv3 --pmc (FETCH_SIZE+WRITE_SIZE)")
    args = ap.parse_args()
    if args.pmc_commands:
        print("\n".join(pmc_commands()))
        return
    if args.pmc_summary:
        pmc_summary()
        return
    if args.trace_commands:
        print("\n".join(trace_commands()))
        return
    if args.trace_summary:
        trace_summary()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; a multi-GPU run would split every slice's word "
              "range across GPUs with no collective, which this bench does not launch", file=sys.stderr)
        sys.exit(2)
    if par2hip.available() <= 0:
        print("bench.py: no gfx950 device visible to libpar2hip.so", file=sys.stderr)
        sys.exit(1)
    if args.autotune is None:
        # encode/repair: the MD5 chains bound the step, the contraction keeps
        # the heuristic's shape (and its measured traffic); served: every
        # request is a new shape class, so tuning would be paid per request
        args.autotune = 0 if args.workload in ("encode", "repair", "served", "encode_host") else 1
    ctx = par2hip.Context(0)
    ctx.set_autotune(bool(args.autotune))
    log("workload", args.workload, "steps", args.steps, "warmup", args.warmup, "device", ctx.device_info())
    if args.workload == "copy":
        # calibration for the HBM byte counters: device-to-device copies of a
        # known size (1 GiB read + 1 GiB written per copy)
        nb = 1 << 30
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        a.upload(random_bytes(3, nb))
        ctx.copy_dev(b, a, nb)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.copy_dev(b, a, nb)
        ctx.sync()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"copy": {"bytes_read": nb, "bytes_written": nb, "ms": round(dt * 1e3, 4),
                                   "GBps_rw": round(2 * nb / dt / 1e9, 1)}}), flush=True)
        ctx.close()
        return
    if args.workload == "small_r":
        print(json.dumps({"small_r": bench_small_r(args, ctx)}), flush=True)
        ctx.close()
        return
    if args.workload == "inverse":
        print(json.dumps({"inverse": bench_inverse(args, ctx)}), flush=True)
        ctx.close()
        return
    if args.workload == "micro":
        print(json.dumps({"micro": bench_micro(args, ctx)}), flush=True)
        ctx.close()
        return
    if args.workload == "served":
        out = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "u16", "data": "synthetic (seeded uniform random bytes)", **served_line(bench_served(args, ctx)),
               "device": ctx.device_info()}
        print(json.dumps(out), flush=True)
        ctx.close()
        return
    if args.workload == "batched":
        res = bench_batched(args, ctx)
    elif args.workload == "create":
        res = bench_create(args, ctx)
    elif args.workload == "repair_files":
        res = bench_repair_files(args, ctx)
    elif args.workload == "encode_host":
        res = bench_encode_host(args, ctx)
    else:
        res = bench_encode(args, ctx) if args.workload == "encode" else bench_repair(args, ctx)
    out = {"metric": METRIC, **step_line(res, args), "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u16", "data": "synthetic (seeded uniform random bytes)",
           "device": ctx.device_info()}
    out = {k: out[k] for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                               "higher_is_better", "scaling", "vs_baseline", "dtype", "data")} | out
    if not args.no_cpu:
        cpu = cpu_leg(args, res)
        if cpu:
            out["cpu_baseline"] = cpu
    if args.workload == "encode" and args.legs:
        # the metric's other halves in the same run: plain repair, the batched
        # sets against the plain path, and the served stream against the plain
        # per-request path at equal p99
        legs = {}
        for leg in args.legs:
            log("leg", leg)
            if leg == "repair":
                r = bench_repair(args, ctx)
                legs["repair"] = step_line(r, args)
                if not args.no_cpu:
                    legs["repair"]["cpu_baseline"] = cpu_leg(args, r)
            elif leg == "batched":
                steps = args.steps
                args.steps = max(3, min(args.steps, args.leg_steps))
                legs["batched"] = step_line(bench_batched(args, ctx), args)
                args.steps = steps
            elif leg == "served":
                legs["served"] = served_line(bench_served(args, ctx))
        out.update(legs)
    print(json.dumps(out), flush=True)
    ctx.close()


def step_line(res, args):
    """The JSON fields of a stepped workload: throughput, step latency, the
    exactness check, kernel times, roofline and the plain-path comparison."""
    lat = np.array(res["lat"]) * 1e3
    steps = len(res["lat"])
    ms = res["total_s"] / steps * 1e3
    value = res["src_bytes"] / (ms * 1e-3) / 1e9
    out = {"value": round(value, 3), "unit": "GB/s", "steps": steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 4), "config": res["config"],
           "p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
           "bit_exact_sample": res["exact"],
           "kernel_ms": {k: round(v, 4) for k, v in res["kernel_ms"].items()}}
    if "roofline" in res:
        out["roofline"] = res["roofline"]
    if "eager_ms" in res:
        out["batch_eager_ms"] = round(res["eager_ms"], 3)
        out["step"] = "one replay of the batch captured as a HIP graph (planned once)"
    if "tile_shapes" in res:
        out["tile_shapes"] = res["tile_shapes"]
    if "plain_ms" in res:
        out["plain_path"] = {"ms_per_batch": round(res["plain_ms"], 3), "value": round(res["plain_gbs"], 3),
                             "unit": "GB/s", "speedup_batched": round(value / res["plain_gbs"], 3)}
    if "cpu_baseline" in res:
        out["cpu_baseline"] = res["cpu_baseline"]
    return out


def served_line(res):
    return {"value": res["value"], "unit": "GB/s", "steps": len(res["served"]), "warmup": 0, "ms_per_step": None,
            "config": res["config"], "p50_ms": res["p50_ms"], "p99_ms": res["p99_ms"],
            "bit_exact_sample": res["exact"], "served": res["served"], "server_stats": res["server_stats"],
            "plain_path": res["plain_path"]}


def cpu_leg(args, res):
    """The reference's CPU method (the oracle's restatement of its AVX2
    strategy and hashing) on this host's cores for the same workload."""
    if "cpu" not in res:
        return None
    threads, how = host_cpus()
    build = cpu_port_build()
    n, ss, r = res["cpu"]
    log("cpu baseline", n, ss, r, "threads", threads, build)
    if res["workload"] in ("plain_encode", "encode_host"):
        gbs, dt = cpu_baseline_encode(n, ss, r, threads)
        what = (f"{n} slices x {ss} B, R={r}: AVX2 nibble-shuffle encode (StreamingEncoder batching) "
                f"+ MD5/CRC32 per slice, one pass, {dt:.2f} s wall")
    else:
        gbs, dt = cpu_baseline_repair(n, ss, r, threads)
        what = (f"{n} slices x {ss} B: MD5/CRC32 verify of every slice + rebuild contraction of {r} slices "
                f"(AVX2 nibble-shuffle), one pass, {dt:.2f} s wall")
    return {"value": round(gbs, 4), "unit": "GB/s", "cores": threads, "kind": "port", "cores_from": how,
            "build": build, "sample": what}


if __name__ == "__main__":
    main()
