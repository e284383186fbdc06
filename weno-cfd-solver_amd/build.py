"""Build libwenox.so (gfx950) in place: python weno-cfd-solver_amd/build.py

The kernel translation units are compiled once per arithmetic variant
("exact": reference operation order, no FMA contraction; "fast": FMA
contraction -- within source expressions only for the fused sweeps, so that
every tiling rounds alike -- and shared reciprocals; the layer-by-layer
kernels also "twin":
the NumPy twin's semantics), the tuning objective once, and linked with the
C-ABI host code into
wenox/libwenox.so next to the Python package that loads it.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "wenox", "libwenox.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 math (v_pk_*_f32) stays off: some of its forms have
          # produced wrong results on this target without any error
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]

UNITS = [
    ("wenox_kernels.hip", "kernels_exact.o", ["-DWX_FAST=0", "-ffp-contract=off"]),
    ("wenox_kernels.hip", "kernels_fast.o", ["-DWX_FAST=1", "-ffp-contract=fast"]),
    ("wenox_kernels.hip", "kernels_twin.o", ["-DWX_FAST=0", "-DWX_TWIN=1", "-ffp-contract=off"]),
    ("wenox_sweep.hip", "sweep_exact.o", ["-DWX_FAST=0", "-ffp-contract=off"]),
    # contraction within source expressions only ("on"): with "fast" the
    # backend's FMA formation depends on how many uses a product has after
    # inlining, which differs between template instances (wavefronts per
    # workgroup), and the FP32 arithmetic of the split precision then rounded
    # differently for different tilings
    ("wenox_sweep.hip", "sweep_fast.o", ["-DWX_FAST=1", "-ffp-contract=on"]),
    ("wenox_objective.hip", "objective.o", ["-ffp-contract=off"]),
    ("wenox_sums.hip", "sums.o", ["-ffp-contract=off"]),
    ("wenox_abi.cpp", "abi.o", ["-x", "hip", "-ffp-contract=off"]),
    ("wenox_batch.cpp", "batch.o", ["-x", "hip", "-ffp-contract=off"]),
]


def _compile(unit, tmp, verbose):
    src, obj, extra = unit
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(tmp, obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines()
            if ln.strip() and "packed-fp32-ops" not in ln and "unknown target feature" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n" + "\n".join(msgs))
    return msgs


def build(verbose=False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    tmp = tempfile.mkdtemp(prefix="wenox_build_")
    try:
        with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
            for msgs in ex.map(lambda u: _compile(u, tmp, verbose), UNITS):
                for m in msgs:
                    print(m, file=sys.stderr)
        objs = [os.path.join(tmp, u[1]) for u in UNITS]
        staged = OUT + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", staged] + objs
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
        os.replace(staged, OUT)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
