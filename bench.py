"""Throughput of the device-resident WENO5/HLLC SSP-RK3 step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--workload quad4096|dmr|kh8192|batch512] [--engine fused|unfused]
                [--variant fast|exact] [--precision fp64|fp32split] [--no-cpu] [--no-extra]

A step is one full SSP-RK3 time step over the whole grid -- CFL dt, 3 x (x and
y sweeps, RK combine with the ghost cells), time advance -- replayed from a
hipGraph with the state resident in HBM; both step graphs (one step, eight
steps) are captured by the first, untimed run() call.

Default workload (the one the throughput target is quoted on): the
four-quadrant Riemann problem on [0,1]^2 at 4096 x 4096 cells, FP64,
characteristic WENO5 + HLLC (sensor window (1e4, 1.1e4): phi = 1), SSP-RK3 at
CFL 0.4, mirror ("transmissive") ghost cells, synthetic initial data built on
the host.  Before timing, the sweep tiling is picked by the solver's
autotuner.  Prints one JSON line.

`value` is the FP64 fast-variant throughput.  Beside it the line carries: the
exact variant (the reference's operation order, bitwise equal to the oracle)
on the same grid; FP32-split on the same grid with its relative L1 and
L-infinity density distance from the FP64 run; the scheme variants
{characteristic, component-wise} x {HLLC, HLLC/HLLE blend}; the double Mach
reflection 1920 x 480, Kelvin-Helmholtz 8192^2 and the 512-instance batched
workload (FP64 and FP32-split each); the drop-in l_local_py call; the jp_cri
tuning objective; the roofline blocks (per-kernel device times from HIP
events on the solver's stream, HBM traffic and VALU counts from the counter
profiles under profiles/); and the CPU oracle timed on the host's cores
(cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "weno-cfd-solver_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12
BYTES_PER_CELL_UPDATE = 256.0   # 4 conserved values x 8 B: 64 B stage 1 + 2 x 96 B stages 2, 3
VALU_WAVE_INSTR_PEAK = 256 * 4 * 2.4e9 / 4   # CUs x SIMDs x clock / 4 cycles per wave64 FP64 op

WORKLOADS = {
    "dmr": dict(name="double_mach_reflection", nx=1920, ny=480, cons=False, force_hlle=False,
                jp=(1e4, 1.1e4), bc="dmr", cfl=0.1),
    # the reference scheme (characteristic WENO5 + HLLC) turns this problem
    # non-finite between steps 130 and 140 at CFL 0.4 (the exact variant, bitwise
    # the oracle's arithmetic, at the same step; FP32-split, whose rounding
    # perturbs the growing mode, before step 120): timed runs restart from the
    # initial state at most every `window` steps (see run_gpu)
    "quad4096": dict(name="four_quadrant_riemann", nx=4096, ny=4096, cons=False, force_hlle=False,
                     jp=(1e4, 1.1e4), bc="mirror", cfl=0.4, window=100),
    "kh8192": dict(name="kelvin_helmholtz", nx=8192, ny=8192, cons=False, force_hlle=False,
                   jp=(1e4, 1.1e4), bc="periodic", cfl=0.4),
    # 512 independent four-quadrant instances, sizes from {64,96,128,192,256,384}^2
    "batch512": dict(name="batched_four_quadrant", nx=0, ny=0, cons=False, force_hlle=False,
                     jp=(1e4, 1.1e4), bc="mirror", cfl=0.4, count=512,
                     sizes=(64, 96, 128, 192, 256, 384)),
}


def describe(w):
    if w["name"] == "batched_four_quadrant":
        return (f"{w['count']} independent four-quadrant instances, sizes uniform from "
                f"{{{','.join(str(n) for n in w['sizes'])}}}^2 (seed 20261203), characteristic WENO5 + HLLC, "
                f"per-instance dt at CFL {w['cfl']}, bc {w['bc']}")
    return (f"{w['name']} {w['nx']}x{w['ny']}, characteristic WENO5 + HLLC (jp {w['jp'][0]:g},{w['jp'][1]:g}), "
            f"SSP-RK3, CFL {w['cfl']}, bc {w['bc']}")


def make_state(w):
    from wenox import problems
    if w["name"] == "double_mach_reflection":
        return problems.double_mach(w["nx"], w["ny"])
    if w["name"] == "four_quadrant_riemann":
        return problems.four_quadrant(w["nx"])
    return problems.kelvin_helmholtz(w["nx"])


def run_gpu(w, q, dx, steps, warmup, engine, variant, precision, profile=True, tune=True, chunks=None,
            layout=None):
    """layout: "row" or "blocked" state layout of the fused FP64 fast step;
    None: the autotuner picks it (with `chunks` given: row-major)."""
    import numpy as np
    from wenox import Solver
    s = Solver(w["nx"], w["ny"], dx, gamma=1.4, cfl=w["cfl"], cons_switch=w["cons"],
               force_hlle=w["force_hlle"], jp_cri=w["jp"], bc=w["bc"], engine=engine,
               variant=variant, precision=precision)
    try:
        s.upload(q)
        tuned = None
        if engine == "fused":
            if chunks:
                s.set_tiling(*chunks)
                s.set_blocked(layout == "blocked")
            elif tune:
                tuned = s.autotune(layouts=None if layout is None else (layout == "blocked",))
        s.run(nstep=warmup)
        # A workload with a `window` is timed in segments of at most that many
        # steps, each from the initial state (re-uploaded between segments,
        # outside the timed segments): its scheme turns the state non-finite
        # after about that many steps.
        window = w.get("window")
        el, k, T = 0.0, 0, 0.0
        while k < steps:
            seg = steps - k if not window else min(window, steps - k)
            if window:
                s.upload(q)
            t0 = time.perf_counter()
            T, kk = s.run(nstep=seg)
            el += time.perf_counter() - t0
            k += kk
            assert kk == seg, (kk, seg)
        state = s.download()[3:-3, 3:-3, :4]
        out = dict(seconds=el, steps=k, value=w["nx"] * w["ny"] * k / el, dev_ms_per_step=s.ms_per_step,
                   state=state, finite=bool(np.isfinite(state).all()), t=T, variant=variant,
                   segments=-(-steps // window) if window else 1,
                   # steps from the initial state to `state`
                   state_steps=seg if window else warmup + steps)
        if precision == "fp32split":
            out["split_jump"] = s.split_jump
        if engine == "fused":
            out["chunks"] = list(s.tiling)
            if tuned:
                out["autotune_ms"] = {
                    "sweep_x": {f"w{w}c{c}": round(v, 5) for (w, c), v in tuned["sweep_x"].items()},
                    "sweep_y": {f"w{w}c{c}": round(v, 5) for (w, c), v in tuned["sweep_y"].items()},
                    "pairs": {"x{0}w{2}_y{1}w{3}".format(*t): round(v, 5) for t, v in tuned["pairs"].items()},
                    "combine_cols": {str(c): round(v, 5) for c, v in tuned["combine"].items()},
                    "pipeline_stage": {str(c): round(v, 5) for c, v in tuned.get("pipeline", {}).items()},
                    "layout_step_ms": {k: round(v, 5) for k, v in tuned["step_ms"].items()}}
        if engine == "fused":
            out["layout"] = "blocked" if s.blocked else "row"
            out["combine_cols"] = s.combine_cols
            out["pipeline"] = s.pipeline
        if profile and engine == "fused":
            out["kernels_ms"] = s.profile(reps=10)
        return out
    finally:
        s.close()


# the batched workload's configured state ranges: rho~U[0.1,2], p~U[0.02,2], u,v~U[-1.5,1.5]
CONFIG_RANGES = dict(rho=(0.1, 2.0), p=(0.02, 2.0), vel=1.5)


def run_batch(w, steps, warmup, precision, chunk=16, ranges=None, tune=True):
    """The batched engine (per-instance dt), with the step structure and the
    sweep chunk its autotuner picks: the persistent step (one launch per
    eight steps, work items from a device queue) or one launch per phase over
    all instances; value = sum of instance cells x steps / wall.
    ranges: the instances' state ranges (default: the narrowed ranges of
    problems.random_riemann_instance)."""
    import numpy as np
    from wenox import Batch, problems
    rng = np.random.default_rng(20261203)
    sizes = [(int(n), int(n)) for n in rng.choice(w["sizes"], size=w["count"])]
    inst = [problems.random_riemann_instance(rng, nx, ny, **(ranges or {})) for nx, ny in sizes]
    cells = sum(nx * ny for nx, ny in sizes)
    with Batch(sizes, [d for _, d in inst], cfl=w["cfl"], cons_switch=w["cons"], force_hlle=w["force_hlle"],
               jp_cri=w["jp"], bc=w["bc"], variant="fast", precision=precision, chunk=chunk) as b:
        for i, (q, _) in enumerate(inst):
            b.upload(i, q)
        if tune:
            tuned = b.autotune()
        else:
            # the given chunk and the engine's default step structure
            tuned = {"best": b.chunks[0], "persistent": b.persistent}
        b.run(nstep=warmup)
        t0 = time.perf_counter()
        b.run(nstep=steps)
        el = time.perf_counter() - t0
        T, k = b.status()
        bad = [i for i in range(w["count"]) if not np.isfinite(b.download(i)[3:-3, 3:-3, :4]).all()]
    return dict(value=cells * steps / el, seconds=el, cells=cells, steps_each=int(np.min(k)) - warmup,
                ms_per_step=el / steps * 1e3, non_finite=len(bad), chunk=tuned["best"],
                autotune_stage_ms={str(c): round(v, 5) for c, v in tuned.items() if isinstance(c, int)},
                step_structure="persistent" if tuned.get("persistent") else "launch per phase",
                autotune_step_ms={m: {str(c): v for c, v in d.items()} for m, d in tuned.get("step_ms", {}).items()})


def _cpu_steps(q, dx, w, budget_s):
    """Whole SSP-RK3 steps of the oracle RHS (the reference's stage
    combination in numpy) until ~budget_s elapsed (at least one)."""
    import numpy as np

    import oracle
    nx, ny = w["nx"], w["ny"]
    sl = np.s_[3:nx + 3, 3:ny + 3, :4]
    q = q.copy()
    t0 = time.perf_counter()
    steps = 0
    dt = 1e-5
    while True:
        l0 = oracle.l_local_py(q, dx, 1.4, w["force_hlle"], w["jp"], w["cons"])
        q1 = q.copy(); q1[sl] = q[sl] + dt * l0
        l1 = oracle.l_local_py(q1, dx, 1.4, w["force_hlle"], w["jp"], w["cons"])
        q2 = q.copy(); q2[sl] = 0.75 * q[sl] + 0.25 * q1[sl] + 0.25 * dt * l1
        l2 = oracle.l_local_py(q2, dx, 1.4, w["force_hlle"], w["jp"], w["cons"])
        q[sl] = (1 / 3) * q[sl] + (2 / 3) * q2[sl] + (2 / 3) * dt * l2
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return steps, time.perf_counter() - t0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_cores():
    """CPUs this process may run on: its affinity set, capped by a cgroup
    CPU quota when one is set.  Returns (cores, affinity, quota or None)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    return (min(aff, quota) if quota else aff), aff, quota


def cpu_baseline(q, dx, w, budget_s=10.0, slab=256):
    """The oracle (C restatement of the reference Rust RHS, OpenMP over the
    outer loops of its passes) driving the reference's SSP-RK3 stage
    combination, timed on the host: on every core this process may use (the
    reported figure) and on one thread as the reference runs (beside it).
    The sample is whole steps on a slab of `slab` x-rows through the middle
    of the workload's state, with its 3-cell halo and the full y extent."""
    import oracle
    oracle.lib()
    nx, ny = w["nx"], w["ny"]
    rows = min(slab, nx)
    i0 = 3 + (nx - rows) // 2
    sub = q[i0 - 3:i0 + rows + 3].copy()
    ws = dict(w, nx=rows)
    cores, aff, quota = _host_cores()
    res = {}
    try:
        for t in sorted({cores, 1}, reverse=True):
            oracle.set_threads(t)
            res[t] = _cpu_steps(sub, dx, ws, budget_s)
    finally:
        oracle.set_threads(1)
    steps, el = res[cores]
    s1, e1 = res[1]
    return dict(value=rows * ny * steps / el, unit="cell-updates/s", cores=cores, kind="port",
                sample=f"{steps} SSP-RK3 step(s) (3 oracle l_local calls + numpy stage combines each) of the "
                       f"{rows} x {ny} slab at x-rows {i0 - 3}..{i0 + rows - 4} of the {nx}x{ny} {w['name']} "
                       f"state on {cores} thread(s), {el:.1f} s",
                one_thread=dict(value=rows * ny * s1 / e1, steps=s1, seconds=round(e1, 2)),
                host=_cpu_model(), cpus_affinity=aff, cpu_quota=quota)


def dropin_rate(q, dx, w, calls=5):
    """The drop-in path: weno_ext.l_local_py on the host array (pack, upload,
    unfused kernels, download per call), fast and exact variants; cell
    RHS evaluations per second including the PCIe copies."""
    import weno_ext
    out = {}
    for fast in (True, False):
        weno_ext.set_variant(fast)
        weno_ext.l_local_py(q, dx, 1.4, w["force_hlle"], tuple(w["jp"]), w["cons"])
        t0 = time.perf_counter()
        for _ in range(calls):
            weno_ext.l_local_py(q, dx, 1.4, w["force_hlle"], tuple(w["jp"]), w["cons"])
        el = (time.perf_counter() - t0) / calls
        out["fast" if fast else "exact"] = {"ms_per_call": el * 1e3,
                                            "cell_rhs_per_s": w["nx"] * w["ny"] / el}
    weno_ext.set_variant(False)
    out["note"] = ("weno_ext.l_local_py with host arrays: one RHS evaluation per call including the "
                   "host->device and device->host copies (PCIe-inclusive), not a time step")
    return out


def objective_rate(w, q, dx, steps, reps=5):
    """The jp_cri tuning objective (metrics.compute_2d_objective) of the state
    after `steps` steps: on the GPU from the solver's resident state (wall
    time per evaluation, allocation and the thinning loop's host checks
    included), and the oracle (numpy + scipy.ndimage, one host thread) on the
    downloaded fields."""
    import numpy as np
    from wenox import Solver, tuning
    with Solver(w["nx"], w["ny"], dx, gamma=1.4, cfl=w["cfl"], cons_switch=w["cons"],
                force_hlle=w["force_hlle"], jp_cri=w["jp"], bc=w["bc"], engine="fused", variant="fast") as s:
        s.upload(q)
        s.run(nstep=steps)
        J, _ = tuning.solver_objective(s)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            tuning.solver_objective(s)
            times.append(time.perf_counter() - t0)
        st = s.download()
    out = {"workload": f"{describe(w)} state after {steps} steps", "J": J,
           "gpu_ms": float(np.median(times)) * 1e3}
    try:
        from oracle import objective as ob
    except ImportError:
        return out
    rho, u, v = st[3:-3, 3:-3, 0], st[3:-3, 3:-3, 5], st[3:-3, 3:-3, 6]
    t0 = time.perf_counter()
    Jc, _ = ob.compute_2d_objective(rho, u, v, None)
    out.update(cpu_ms=(time.perf_counter() - t0) * 1e3, cpu_kind="port", cpu_cores=1,
               rel_diff_J=abs(J - Jc) / abs(Jc) if Jc else abs(J - Jc))
    return out


def kernel_bytes(w, split, fold=False, blocked=False):
    """Algorithmic HBM bytes of each launch of a step (state and D arrays
    touched once each; halo re-reads counted 0), averaged over the three
    stages.  fold: the fast FP64 step, whose x sweep also reads U^n in stages
    2 and 3 and writes its part of the stage combination, so that the combine
    reads only that and Dy^T.  blocked: the blocked fast FP64 step, whose y
    sweep reads that part and writes the new state (no combine launch)."""
    cells = w["nx"] * w["ny"]
    padded = (w["nx"] + 6) * (w["ny"] + 6)
    d = 16 if split else 32            # D planes: 4 x FP32 or 4 x FP64 per cell
    st = 32                            # state: 4 x FP64, or FP32 hi + lo
    if blocked:
        return {"sweep_x": st * padded + (2 * st / 3) * cells + d * cells,
                "sweep_y": st * padded + d * cells + st * cells, "cfl": st * cells}
    if fold:
        return {"sweep_x": st * padded + (2 * st / 3) * cells + d * cells, "sweep_y": st * padded + d * cells,
                "combine": 2 * d * cells + 2 * st * cells, "cfl": st * cells}
    return {"sweep_x": st * padded + d * cells, "sweep_y": st * padded + d * cells,
            # stage 1 reads U^n, stages 2 and 3 also U_s; both orientations written
            "combine": 2 * d * cells + (st + 2 * st / 3) * cells + 2 * st * cells,
            "cfl": st * cells}


def roofline(res, w, split=False):
    """HBM roofline at the metric's 256 algorithmic B per cell-update for the
    step (CFL dt + 3 x (x and y sweeps, combine, ghost fill); FP64: the last
    combine also forms the next step's CFL maxima and the time advance rides
    on the next dt launch), timed as the sum of the per-kernel HIP-event
    averages on the solver's stream; then each kernel against its own
    algorithmic bytes.  `traffic` (HBM bytes per step from FETCH_SIZE /
    WRITE_SIZE counters) is filled where a counter profile of this workload
    exists."""
    km = res.get("kernels_ms")
    cells = w["nx"] * w["ny"]
    step_ms = km["step_sum"] if km else res["dev_ms_per_step"]
    achieved = cells * BYTES_PER_CELL_UPDATE / (step_ms * 1e-3)
    tiling = tuple(res.get("chunks") or ())
    fold = not split and res.get("variant") == "fast"
    blocked = res.get("layout") == "blocked"
    # the counter tables hold the fused engine's fast-variant kernels, per layout
    profiled = res.get("variant") == "fast" and bool(tiling)
    layout = res.get("layout") or "row"
    tr = TRAFFIC.get((w["name"], w["nx"], w["ny"], split, layout)) if profiled else None
    out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": achieved / HBM_PEAK, "traffic": tr["bytes_per_step"] if tr else None,
           "kernel": "one SSP-RK3 step (its kernels, back to back)",
           "algorithmic_bytes_per_launch": cells * BYTES_PER_CELL_UPDATE,
           "avg_launch_ms": step_ms}
    if tr:
        # the profiled tiling; another tiling changes only the sweeps' halo re-reads
        out["traffic_source"] = tr["source"]
        out["traffic_tiling"] = list(tr["tiling"])
    if km:
        kb = kernel_bytes(w, split, fold, blocked)
        per = {}
        for k, launches in (("sweep_x", 3), ("sweep_y", 3), ("combine", 3), ("cfl", 1)):
            # FP64: the CFL maxima come out of the last combine's registers;
            # the step's own CFL launch reads no state
            if k not in kb or (k == "cfl" and not split):
                continue
            per[k] = {"avg_ms": round(km[k], 5), "launches_per_step": launches,
                      "algorithmic_bytes": kb[k], "achieved_GBps": round(kb[k] / (km[k] * 1e-3) / 1e9, 1)}
        out["kernels"] = per
        # the step launches the two sweeps together when their wave counts agree
        out["sweeps_in_step_ms"] = round(km["sweeps"], 5)
        out["ghost_fill_ms"] = round(km["ghost_fill"], 5)
        # the step launches the stage's two sweeps as one kernel (k_sweep_xy)
        # when their tilings share a wave count; that launch dominates the step
        sw_bytes = kb["sweep_x"] + kb["sweep_y"]
        one = len(tiling) == 4 and tiling[2] == tiling[3]
        name = ("k_sweep x (part E of the stage) then k_sweep y (the new state): two launches per stage"
                if blocked else "k_sweep_xy (x and y sweeps of a stage, one launch)" if one
                else "k_sweep x then y (two launches per stage)")
        out["dominant_kernel"] = {"name": name,
                                  "avg_ms": round(km["sweeps"], 5), "launches_per_step": 3,
                                  "algorithmic_bytes": sw_bytes,
                                  "achieved_GBps": round(sw_bytes / (km["sweeps"] * 1e-3) / 1e9, 1),
                                  "share_of_step": round(3 * km["sweeps"] / step_ms, 4)}
        v = VALU.get((w["name"], w["nx"], w["ny"], split, layout, "sweeps"))
        if v and tuple(v["tiling"]) == tiling and profiled:
            rate = v["wave_instr_per_launch"] / (km["sweeps"] * 1e-3)
            out["dominant_kernel"]["valu"] = {"wave_instr_per_launch": v["wave_instr_per_launch"],
                                              "achieved": rate, "peak": VALU_WAVE_INSTR_PEAK,
                                              "unit": "wave64 VALU instr/s", "frac": rate / VALU_WAVE_INSTR_PEAK,
                                              "source": v["source"]}
    return out


# HBM bytes per step and VALU instructions per sweep launch of the fused
# engine's fast variant from rocprofv3 counter runs (FETCH_SIZE x 2 +
# WRITE_SIZE x 1 KiB: the factors a plain 512 MiB device copy reads back;
# SQ_INSTS_VALU per dispatch), keyed by workload, precision and sweep tiling
# (chunk_x, chunk_y, waves_x, waves_y): the halo re-reads depend on the
# tiling, so other tilings report null.
_DMR_SRC64 = "profiles/dmr1920x480_fp64.md"
_DMR_SRC32 = "profiles/dmr1920x480_fp32split.md"
_QUAD_SRC64 = "profiles/quad4096_fp64.md"
_QUAD_SRC64B = "profiles/quad4096_fp64_blocked.md"
_QUAD_SRC32 = "profiles/quad4096_fp32split.md"
_KH_SRC64 = "profiles/kh8192_fp64.md"
_KH_SRC32 = "profiles/kh8192_fp32split.md"
_KH_SRC64B = "profiles/kh8192_fp64_blocked.md"
_DMR_SRC64B = "profiles/dmr1920x480_fp64_blocked.md"
TRAFFIC = {
    ("double_mach_reflection", 1920, 480, False, "row"):
        {"bytes_per_step": 1.0534e9, "tiling": (10, 9, 4, 4), "source": _DMR_SRC64},
    ("double_mach_reflection", 1920, 480, False, "blocked"):
        {"bytes_per_step": 0.7338e9, "tiling": (10, 9, 4, 4), "source": _DMR_SRC64B},
    ("double_mach_reflection", 1920, 480, True, "row"):
        {"bytes_per_step": 1.1756e9, "tiling": (10, 8, 2, 1), "source": _DMR_SRC32},
    ("four_quadrant_riemann", 4096, 4096, False, "row"):
        {"bytes_per_step": 16.7864e9, "tiling": (48, 48, 1, 1), "source": _QUAD_SRC64},
    ("four_quadrant_riemann", 4096, 4096, False, "blocked"):
        {"bytes_per_step": 11.8661e9, "tiling": (48, 48, 1, 1), "source": _QUAD_SRC64B},
    ("four_quadrant_riemann", 4096, 4096, True, "row"):
        {"bytes_per_step": 16.0526e9, "tiling": (48, 48, 1, 1), "source": _QUAD_SRC32},
    ("kelvin_helmholtz", 8192, 8192, False, "row"):
        {"bytes_per_step": 66.7687e9, "tiling": (57, 57, 1, 1), "source": _KH_SRC64},
    ("kelvin_helmholtz", 8192, 8192, False, "blocked"):
        {"bytes_per_step": 47.4133e9, "tiling": (57, 57, 1, 1), "source": _KH_SRC64B},
    ("kelvin_helmholtz", 8192, 8192, True, "row"):
        {"bytes_per_step": 62.7354e9, "tiling": (57, 57, 1, 1), "source": _KH_SRC32},
}
# VALU wave-instructions of a stage's sweeps as the step launches them (one
# k_sweep_xy, or the x and y k_sweep launches together when their wave counts
# differ or the layout is blocked; FP32-split: including the in-launch FP64
# recompute of flagged face rows)
VALU = {
    ("double_mach_reflection", 1920, 480, False, "row", "sweeps"):
        {"wave_instr_per_launch": 4.984e7, "tiling": (10, 9, 4, 4), "source": _DMR_SRC64},
    ("double_mach_reflection", 1920, 480, False, "blocked", "sweeps"):
        {"wave_instr_per_launch": 2.532e7 + 2.453e7, "tiling": (10, 9, 4, 4), "source": _DMR_SRC64B},
    ("double_mach_reflection", 1920, 480, True, "row", "sweeps"):
        {"wave_instr_per_launch": 3.600e7 + 3.959e7, "tiling": (10, 8, 2, 1), "source": _DMR_SRC32},
    ("four_quadrant_riemann", 4096, 4096, False, "row", "sweeps"):
        {"wave_instr_per_launch": 8.014e8, "tiling": (48, 48, 1, 1), "source": _QUAD_SRC64},
    ("four_quadrant_riemann", 4096, 4096, False, "blocked", "sweeps"):
        {"wave_instr_per_launch": 4.000e8 + 4.052e8, "tiling": (48, 48, 1, 1), "source": _QUAD_SRC64B},
    ("four_quadrant_riemann", 4096, 4096, True, "row", "sweeps"):
        {"wave_instr_per_launch": 1.057e9, "tiling": (48, 48, 1, 1), "source": _QUAD_SRC32},
    ("kelvin_helmholtz", 8192, 8192, False, "row", "sweeps"):
        {"wave_instr_per_launch": 3.295e9, "tiling": (57, 57, 1, 1), "source": _KH_SRC64},
    ("kelvin_helmholtz", 8192, 8192, False, "blocked", "sweeps"):
        {"wave_instr_per_launch": 1.625e9 + 1.686e9, "tiling": (57, 57, 1, 1), "source": _KH_SRC64B},
    ("kelvin_helmholtz", 8192, 8192, True, "row", "sweeps"):
        {"wave_instr_per_launch": 4.294e9, "tiling": (57, 57, 1, 1), "source": _KH_SRC32},
}


def copy_calibration(mib=512, reps=5):
    """A plain device-to-device copy of known size (torch's elementwise copy
    kernel): its FETCH_SIZE / WRITE_SIZE counters per dispatch, divided by the
    bytes it must move, scale the counters read on the solver's kernels into
    HBM bytes."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n * 8, "bytes_written": n * 8, "launches": reps}}),
          flush=True)


def _dist(a, b):
    """Relative L1 and L-infinity distance of a from b."""
    import numpy as np
    d = np.abs(a - b)
    return {"rel_l1": float(d.sum() / np.abs(b).sum()), "rel_linf": float(d.max() / np.abs(b).max())}


METRIC = "cell-updates/s FP64 vs FP32-split @<=1e-6 L1 err; % of HBM roofline"
SPLIT_DTYPE = "f32 bulk, f32 hi+lo state"


def split_block(w, q, dx, steps, warmup, r64):
    """FP32-split on the same state and steps as the FP64 run r64."""
    r32 = run_gpu(w, q, dx, steps, warmup, "fused", "fast", "fp32split")
    return {"value": r32["value"], "unit": "cell-updates/s", "dtype": SPLIT_DTYPE,
            "speedup_vs_fp64": r32["value"] / r64["value"],
            "ms_per_step": r32["seconds"] / r32["steps"] * 1e3,
            "density_vs_fp64": _dist(r32["state"][..., 0], r64["state"][..., 0]),
            "after_steps": r32["state_steps"], "split_jump": r32["split_jump"],
            "state_finite": r32["finite"], "chunks": r32.get("chunks"),
            "roofline": roofline(r32, w, split=True)}


def config1_block():
    """Configuration 1 as stated: the four-quadrant Riemann problem on 256 x
    256, component-wise WENO5 + HLLC, SSP-RK3 at CFL 0.4, mirror ghost cells,
    to t_end = 0.25 (the reference driver's time loop, on the GPU): FP64 fast
    and FP32-split rates over the whole run, FP32-split's density distance
    from FP64 at t_end.  (Against the oracle: tests/test_gpu_fullsize.py.)"""
    from wenox import Solver, problems
    n = 256
    q0, dx = problems.four_quadrant(n)
    out, dens = {"workload": "four_quadrant_riemann 256x256, component-wise WENO5 + HLLC, SSP-RK3, CFL 0.4, "
                             "bc mirror, t_end 0.25"}, {}
    for prec in ("fp64", "fp32split"):
        with Solver(n, n, dx, cfl=0.4, cons_switch=True, jp_cri=(1e4, 1.1e4), bc="mirror", engine="fused",
                    variant="fast", precision=prec) as s:
            s.upload(q0)
            s.run(nstep=9)          # both step graphs captured
            s.upload(q0)
            t0 = time.perf_counter()
            T, k = s.run(t_final=0.25)
            el = time.perf_counter() - t0
            dens[prec] = s.download()[3:-3, 3:-3, 0]
        out[prec] = {"value": n * n * k / el, "steps": k, "t": T, "seconds": el}
    out["fp32split"]["density_vs_fp64"] = _dist(dens["fp32split"], dens["fp64"])
    out["fp32split"]["speedup_vs_fp64"] = out["fp32split"]["value"] / out["fp64"]["value"]
    return out


def workload_block(key, steps, warmup):
    """FP64 (fast, autotuned; exact on the same tiling) and FP32-split on one
    more workload."""
    w = WORKLOADS[key]
    q, dx = make_state(w)
    r64 = run_gpu(w, q, dx, steps, warmup, "fused", "fast", "fp64")
    rex = run_gpu(w, q, dx, steps, warmup, "fused", "exact", "fp64", profile=False, chunks=r64["chunks"],
                  layout="row")
    out = {"workload": describe(w), "steps": steps, "warmup": warmup,
           "value_fp64": r64["value"], "ms_per_step_fp64": r64["seconds"] / r64["steps"] * 1e3,
           "value_fp64_exact": rex["value"], "chunks": r64["chunks"], "layout": r64["layout"],
           "autotune_layout_step_ms": (r64.get("autotune_ms") or {}).get("layout_step_ms"),
           "state_finite": r64["finite"] and rex["finite"],
           "roofline_fp64": roofline(r64, w)}
    out["fp32_split"] = split_block(w, q, dx, steps, warmup, r64)
    return out


# HBM bytes and VALU wave-instructions per step of the persistent batched
# step from its counter profile (FETCH_SIZE x 2 + WRITE_SIZE x 1 KiB;
# SQ_INSTS_VALU per launch / 8 steps), keyed by step structure, with the
# chunk the profile was made on (another chunk changes only the sweeps'
# halo re-reads)
_BATCH_SRC = "profiles/batch512_fp64_persistent.md"
BATCH_TRAFFIC = {
    "persistent": {"bytes_per_step": 18.2424e9, "wave_instr_per_step": 3.6647e9, "chunk": 64, "source": _BATCH_SRC},
}


def batch_roofline(r):
    """The batch's HBM roofline at 256 algorithmic B per cell-update over
    the wall-clock rate; traffic and VALU issue from the counter profile of
    the same step structure (made at the chunk it names)."""
    out = {"bound": "hbm", "achieved": r["value"] * BYTES_PER_CELL_UPDATE / 1e9, "peak": HBM_PEAK / 1e9,
           "unit": "GB/s", "frac": r["value"] * BYTES_PER_CELL_UPDATE / HBM_PEAK, "traffic": None,
           "algorithmic_bytes_per_step": r["cells"] * BYTES_PER_CELL_UPDATE}
    tr = BATCH_TRAFFIC.get("persistent" if r["step_structure"] == "persistent" else "phases")
    if tr:
        rate = tr["wave_instr_per_step"] / (r["ms_per_step"] * 1e-3)
        out.update(traffic=tr["bytes_per_step"], traffic_source=tr["source"], traffic_chunk=tr["chunk"],
                   valu={"wave_instr_per_step": tr["wave_instr_per_step"], "achieved": rate,
                         "peak": VALU_WAVE_INSTR_PEAK, "unit": "wave64 VALU instr/s",
                         "frac": rate / VALU_WAVE_INSTR_PEAK, "chunk": tr["chunk"]})
    return out


def batch_block(w, steps, warmup, chunk=None):
    """chunk: fixed rows per sweep workgroup (no autotune), for counter runs."""
    kw = dict(chunk=chunk, tune=False) if chunk else {}
    r = run_batch(w, steps, warmup, "fp64", **kw)
    r32 = run_batch(w, steps, warmup, "fp32split", **kw)
    # the configured (wider) state ranges, for the instability evidence: the
    # reference scheme turns most of those instances non-finite
    rc = run_batch(w, steps, warmup, "fp64", ranges=CONFIG_RANGES, **kw)
    return {"workload": describe(w), "cells": r["cells"], "steps": steps, "warmup": warmup,
            "value_fp64": r["value"], "ms_per_step_fp64": r["ms_per_step"], "chunk": r["chunk"],
            "step_structure": r["step_structure"], "autotune_step_ms": r["autotune_step_ms"],
            "autotune_stage_ms": r["autotune_stage_ms"],
            # the reference scheme itself turns some instances non-finite (the oracle
            # reproduces it); they keep stepping, so the count does not change the cost
            "instances_non_finite": r["non_finite"],
            "roofline_fp64": batch_roofline(r),
            "state_ranges": "rho~U[1,2], p~U[1,2], u,v~U[-0.5,0.5] per quadrant",
            "configured_ranges_fp64": {"state_ranges": "rho~U[0.1,2], p~U[0.02,2], u,v~U[-1.5,1.5]",
                                       "value": rc["value"], "instances_non_finite": rc["non_finite"]},
            "fp32_split": {"value": r32["value"], "speedup_vs_fp64": r32["value"] / r["value"],
                           "chunk": r32["chunk"], "instances_non_finite": r32["non_finite"]}}


def main():
    import numpy as np
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="quad4096", choices=sorted(WORKLOADS))
    ap.add_argument("--engine", default="fused", choices=["fused", "unfused"])
    ap.add_argument("--variant", default="fast", choices=["fast", "exact"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32split"],
                    help="precision of the main timed run (the headline value is FP64)")
    ap.add_argument("--chunks", type=int, nargs=4, metavar=("CX", "CY", "WX", "WY"),
                    help="sweep tiling: rows per wavefront and wavefronts per workgroup of the x / y "
                         "sweeps (skips the autotuner; batch512: CX rows per sweep workgroup)")
    ap.add_argument("--layout", default="auto", choices=["auto", "row", "blocked"],
                    help="state layout of the FP64 fast step (auto: the autotuner's choice; with --chunks: row)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 512 MiB device copy (counter calibration) and exit")
    ap.add_argument("--objective", action="store_true",
                    help="measure only the jp_cri tuning objective of the workload's state after "
                         "warmup + steps steps and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU decomposition is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    w = WORKLOADS[args.workload]
    if args.objective:
        q, dx = make_state(w)
        obj = objective_rate(w, q, dx, args.warmup + args.steps)
        if args.no_cpu:
            obj = {k: v for k, v in obj.items() if not k.startswith(("cpu", "rel_diff"))}
        print(json.dumps({"tuning_objective": obj}), flush=True)
        return
    if args.workload == "batch512":
        # --chunks C ...: the batch's sweep chunk fixed at C (no autotune)
        b = batch_block(w, args.steps, args.warmup, chunk=args.chunks[0] if args.chunks else None)
        res = {"metric": METRIC, "value": b["value_fp64"], "unit": "cell-updates/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": b["ms_per_step_fp64"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic", "config": {"workload": describe(w), "cells": b["cells"], "chunk": b["chunk"]},
               "instances_non_finite": b["instances_non_finite"], "roofline": b["roofline_fp64"],
               "fp32_split": b["fp32_split"]}
        print(json.dumps(res), flush=True)
        return
    q, dx = make_state(w)
    fast64 = args.engine == "fused" and args.variant == "fast" and args.precision == "fp64"
    layout = None if args.layout == "auto" else args.layout
    if not fast64:
        layout = "row"
    elif layout is None and args.chunks:
        layout = "row"
    r64 = run_gpu(w, q, dx, args.steps, args.warmup, args.engine, args.variant, args.precision,
                  chunks=args.chunks, layout=layout)
    res = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["seconds"] / r64["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else SPLIT_DTYPE,
        "data": "synthetic",
        "config": {"workload": describe(w), "engine": args.engine, "variant": args.variant,
                   "chunks": r64.get("chunks"), "layout": r64.get("layout"),
                   "combine_cols": r64.get("combine_cols"), "pipeline": r64.get("pipeline"),
                   "autotune_ms": r64.get("autotune_ms")},
        "state_finite": r64["finite"],
        "timed_segments": r64["segments"],
        "dev_ms_per_step": r64["dev_ms_per_step"],
        "roofline": roofline(r64, w, split=args.precision != "fp64"),
    }
    if not args.no_extra and args.engine == "fused" and args.precision == "fp64":
        # the parity variant on the same grid and tiling
        rex = run_gpu(w, q, dx, args.steps, args.warmup, "fused", "exact", "fp64", chunks=r64.get("chunks"),
                      layout="row")
        res["exact_variant"] = {"value": rex["value"], "ms_per_step": rex["seconds"] / rex["steps"] * 1e3,
                                "speed_vs_fast": rex["value"] / r64["value"],
                                "density_fast_vs_exact": _dist(r64["state"][..., 0], rex["state"][..., 0]),
                                "state_finite": rex["finite"], "roofline": roofline(rex, w)}
        del rex
        res["fp32_split"] = split_block(w, q, dx, args.steps, args.warmup, r64)
        if args.workload == "quad4096":
            # {characteristic, component-wise} x {HLLC (phi = 1), HLLC/HLLE blend jp (0.5, 2)},
            # FP64 fast on the tiling tuned above
            variants = {}
            for cons in (False, True):
                for jp, tag in (((1e4, 1.1e4), "hllc"), ((0.5, 2.0), "blend_jp0.5_2")):
                    wv = dict(w, cons=cons, jp=jp)
                    rv = run_gpu(wv, q, dx, 16, 8, "fused", "fast", "fp64", profile=False,
                                 chunks=r64.get("chunks"), layout=r64.get("layout"))
                    variants[("componentwise_" if cons else "characteristic_") + tag] = {
                        "value": rv["value"], "ms_per_step": rv["seconds"] / rv["steps"] * 1e3,
                        "state_finite": rv["finite"]}
            res["scheme_variants_fp64"] = variants
        r64.pop("state", None)
        res["config1"] = config1_block()
        for key, steps, warmup in (("dmr", 200, 20), ("kh8192", 100, 10)):
            if key != args.workload:
                res[key] = workload_block(key, steps, warmup)
        res["batch512"] = batch_block(WORKLOADS["batch512"], 200, 10)
        wd = WORKLOADS["dmr"]
        qd, dd = make_state(wd)
        res["dropin_l_local"] = dropin_rate(qd, dd, wd)
        res["tuning_objective"] = objective_rate(wd, qd, dd, 220)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(q, dx, w)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
