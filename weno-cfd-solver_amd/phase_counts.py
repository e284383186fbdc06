"""Static VALU instruction counts per phase of the fused sweep's march.

python weno-cfd-solver_amd/phase_counts.py [--loop]

Builds wenox_sweep.hip for gfx950 (fast variant) with -DWX_PHASE_MARKS=1,
which puts a scheduling barrier and an assembly comment at each phase
boundary of the march loop (window slide and loads, pass 1, pass 2, the
FP32-split detection, Gauss-point fluxes, Riemann solvers, sensor and
blend), and counts the instructions of the first march loop (the y sweep)
of the characteristic + HLLC/HLLE-blend kernels with one wavefront per
workgroup: FP64 and FP32-split.  Counts are static (one face row of the
loop body; branches such as the positivity limiter's correction count
whether taken or not).  The barriers keep instructions on their side of a
boundary in the machine schedule; arithmetic the compiler places before the
pass-1 marker (loads and their first uses) counts under "window".
"""
import collections
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "wenox_sweep.hip")
KERNELS = {
    "fp64": "_ZN5wenox5fasts10k_sweep_xyIdLi0EdLi0ELb0ELb1ELi1ELb0EEEvNS0_9SweepArgsES2_jjj",
    "fp32split": "_ZN5wenox5fasts10k_sweep_xyIdLi2EfLi1ELb0ELb1ELi1ELb0EEEvNS0_9SweepArgsES2_jjj",
}
PHASES = ("window", "pass1", "pass2", "detect", "fluxes", "riemann", "sensor_blend")


def build_asm(out_dir, marks=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    asm = os.path.join(out_dir, "sweep.s")
    cmd = [hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-S",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
           "-DWX_FAST=1", f"-DWX_PHASE_MARKS={int(marks)}", "-ffp-contract=on", SRC, "-o", asm]
    subprocess.run(cmd, check=True, capture_output=True)
    return asm


def count(lines, symbol):
    i = next(k for k, ln in enumerate(lines) if ln.startswith(symbol + ":"))
    cur, out = None, collections.defaultdict(collections.Counter)
    for ln in lines[i:]:
        t = ln.strip()
        m = re.match(r";@@phase (\w+)", t)
        if m:
            if m.group(1) == "end":
                break
            cur = m.group(1)
            continue
        if cur is None or not t or t[0] in ";." or t.endswith(":"):
            continue
        op = t.split()[0]
        if op.startswith("v_"):
            c = out[cur]
            c["valu"] += 1
            if "f64" in op:
                c["f64"] += 1
            elif "f32" in op:
                c["f32"] += 1
            if "cvt" in op:
                c["cvt"] += 1
            if re.search(r"rcp|rsq|sqrt", op):
                c["transcendental"] += 1
        elif op.startswith("ds_"):
            out[cur]["lds"] += 1
    return out


# --loop: the library build's own march loops (no phase markers), the largest
# loop of each kernel (a backward branch's span)
LOOP_KERNELS = {
    "fp64 blocked x sweep": "_ZN5wenox5fasts7k_sweepIdLi0EdLi0ELi0ELb0ELb1ELi1ELi2EEEvNS0_9SweepArgsE",
    "fp64 blocked y sweep": "_ZN5wenox5fasts7k_sweepIdLi0EdLi0ELi1ELb0ELb1ELi1ELi3EEEvNS0_9SweepArgsE",
    "fp64 row-major x+y sweep": KERNELS["fp64"],
    "fp32split row-major x+y sweep": KERNELS["fp32split"],
}


def loop_count(lines, symbol):
    i = next(k for k, ln in enumerate(lines) if ln.startswith(symbol + ":"))
    j = next(k for k in range(i + 1, len(lines)) if lines[k].startswith(".Lfunc_end"))
    body = lines[i:j]
    labels = {m.group(1): k for k, ln in enumerate(body) for m in [re.match(r"^(\.LBB\w+):", ln)] if m}
    best = None
    for k, ln in enumerate(body):
        m = re.match(r"\s+s_(?:cbranch_\w+|branch)\s+(\.LBB\w+)", ln)
        if m and m.group(1) in labels and labels[m.group(1)] < k:
            span = (labels[m.group(1)], k)
            if best is None or span[1] - span[0] > best[1] - best[0]:
                best = span
    c = collections.Counter()
    for ln in body[best[0]:best[1] + 1]:
        t = ln.strip()
        if not t or t[0] in ";." or t.endswith(":"):
            continue
        op = t.split()[0]
        if op.startswith("v_"):
            c["valu"] += 1
            if re.search(r"rcp|rsq|sqrt", op):
                c["transcendental"] += 1
            if op.startswith("v_mov"):
                c["mov"] += 1
        elif op.startswith("ds_"):
            c["lds"] += 1
    return c


def main():
    if "--loop" in sys.argv:
        with tempfile.TemporaryDirectory() as d:
            lines = open(build_asm(d, marks=False)).read().split("\n")
        print("| kernel | VALU per face row (of which v_mov / transcendental) | LDS instructions |")
        print("|---|---|---|")
        for name, sym in LOOP_KERNELS.items():
            c = loop_count(lines, sym)
            print(f"| {name} | {c['valu']} ({c['mov']} / {c['transcendental']}) | {c['lds']} |")
        return 0
    with tempfile.TemporaryDirectory() as d:
        lines = open(build_asm(d)).read().split("\n")
    res = {k: count(lines, sym) for k, sym in KERNELS.items()}
    print("| phase | " + " | ".join(f"{k} VALU (of which cvt / transcendental)" for k in res) + " |")
    print("|---|" + "---|" * len(res))
    for ph in PHASES:
        cells = []
        for k in res:
            c = res[k].get(ph, collections.Counter())
            cells.append(f"{c['valu']} ({c['cvt']} / {c['transcendental']})" if c["valu"] else "—")
        print(f"| {ph} | " + " | ".join(cells) + " |")
    print("| total | " + " | ".join(str(sum(c["valu"] for c in r.values())) for r in res.values()) + " |")


if __name__ == "__main__":
    sys.exit(main())
