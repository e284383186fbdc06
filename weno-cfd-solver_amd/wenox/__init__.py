"""wenox: MI355X-native (gfx950) 2D WENO5 / HLLC Euler hot path.

``Solver`` keeps the padded conserved state in HBM and advances it with the
reference's SSP-RK3 + ghost fill + CFL time loop in hipGraph replays of one or
eight steps (``Rie_prob_2d_rs.py:169-207,311-345``; ``double_mach_reflection.py:
97-134,229-260``).  The drop-in module ``weno_ext`` (sibling package) keeps
the reference's per-call surface.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import EXACT, FAST, TWIN, PanicException, WenoxError  # noqa: F401

BC_MIRROR, BC_PERIODIC, BC_DMR_REF, BC_DMR_STD, BC_COPY = 0, 1, 2, 3, 4
ENGINE_UNFUSED, ENGINE_FUSED = 0, 1
PREC_FP64, PREC_FP32SPLIT = 0, 1

_BC_NAMES = {"mirror": BC_MIRROR, "transmissive": BC_MIRROR, "zero_gradient": BC_MIRROR,
             "periodic": BC_PERIODIC, "dmr_ref": BC_DMR_REF, "dmr": BC_DMR_STD, "copy": BC_COPY}


# "exact": the Rust crate's operation order (bitwise the oracle); "fast":
# FMA contraction and shared reciprocals (rounding-level differences);
# "twin": the pure-NumPy WENO_HLLC_2d.py's semantics (exact sqrt(3)/12, its
# HLLC and HLLE), unfused FP64 engine, component-wise reconstruction.
_VARIANTS = {"exact": EXACT, "fast": FAST, "twin": TWIN}


def device_count() -> int:
    return _lib.load().wenox_device_count()


class Solver:
    """Device-resident SSP-RK3 solver of one 2D problem.

    Parameters follow the reference drivers' dictionaries: ``nx``, ``ny``,
    ``dx`` (grid_dict), ``cfl``, ``force_hlle``, ``jp_cri``, ``cons_switch``
    (control_dict "weno type"), ``gamma`` (phys_dict).
    """

    def __init__(self, nx, ny, dx, gamma=1.4, cfl=0.4, cons_switch=False, force_hlle=False,
                 jp_cri=(0.2, 1.0), bc="mirror", dmr_bound=0, dmr_x0=1.0 / 6.0,
                 engine="unfused", variant="exact", precision="fp64", use_graph=True):
        d = _lib.SolverDesc()
        d.nx, d.ny, d.dx, d.gamma, d.cfl = int(nx), int(ny), float(dx), float(gamma), float(cfl)
        d.cons_switch = int(bool(cons_switch))
        d.force_hlle = int(bool(force_hlle))
        d.jp_lo, d.jp_hi = float(jp_cri[0]), float(jp_cri[1])
        d.bc = _BC_NAMES[bc] if isinstance(bc, str) else int(bc)
        d.dmr_bound = int(dmr_bound)
        d.dmr_x0 = float(dmr_x0)
        d.engine = {"unfused": ENGINE_UNFUSED, "fused": ENGINE_FUSED}[engine] if isinstance(engine, str) else int(engine)
        d.variant = _VARIANTS[variant] if isinstance(variant, str) else int(variant)
        d.precision = {"fp64": PREC_FP64, "fp32split": PREC_FP32SPLIT}[precision] if isinstance(precision, str) else int(precision)
        d.use_graph = int(bool(use_graph))
        self.desc = d
        self.nx, self.ny = d.nx, d.ny
        self._h = ctypes.c_void_p()
        L = _lib.load()
        _lib.check(L.wenox_solver_create(ctypes.byref(d), ctypes.byref(self._h)))
        self.t = 0.0
        self.steps = 0

    def close(self):
        if self._h:
            _lib.load().wenox_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, q: np.ndarray, t0: float = 0.0):
        """q: (nx+6, ny+6, c>=4) float64; with c >= 8 the first CFL step uses
        channels 5-7 (u, v, p) as the reference does after init."""
        q = np.asarray(q, dtype=np.float64)
        v = _lib.view(q)
        _lib.check(_lib.load().wenox_solver_upload(self._h, ctypes.byref(v), float(t0)))
        self.t = float(t0)

    def download(self) -> np.ndarray:
        """(nx+6, ny+6, 8): conserved channels and the primitives (rho_safe, u, v, p)."""
        out = np.empty((self.nx + 6, self.ny + 6, 8))
        _lib.check(_lib.load().wenox_solver_download(self._h, out.ctypes.data))
        return out

    def run(self, nstep=None, t_final=None):
        """Advance by at most ``nstep`` steps and, if ``t_final`` is given, while
        T < t_final - 1e-12 (Rie_prob_2d_rs.py:335).  Returns (T, steps taken)."""
        if nstep is None and t_final is None:
            raise ValueError("Either 't_final' or 'nstep' must be provided.")
        ms = int(nstep) if nstep is not None else int(1e8)
        tf = float(t_final) if t_final is not None else -1.0
        t = ctypes.c_double()
        k = ctypes.c_int64()
        _lib.check(_lib.load().wenox_solver_run(self._h, ms, tf, ctypes.byref(t), ctypes.byref(k)))
        self.t = t.value
        self.steps += k.value
        return t.value, k.value

    def rhs(self) -> np.ndarray:
        """L(U) of the current (ghost-filled) state, (nx, ny, 4)."""
        out = np.empty((self.nx, self.ny, 4))
        _lib.check(_lib.load().wenox_solver_rhs(self._h, out.ctypes.data))
        return out

    def conserved_sums(self, magnitudes: bool = False) -> np.ndarray:
        """Compensated FP64 sums of rho, rho u, rho v, E over the interior
        cells of the current state (times dx**2: the totals); with
        magnitudes=True also the sums of their absolute values (8 values)."""
        n = 8 if magnitudes else 4
        out = np.empty(n)
        _lib.check(_lib.load().wenox_solver_conserved_sums(self._h, out.ctypes.data, n))
        return out

    @property
    def ms_per_step(self) -> float:
        return _lib.load().wenox_solver_last_ms_per_step(self._h)

    KERNELS = ("cfl", "sweep_x", "sweep_y", "combine", "ghost_fill", "step_sum", "sweeps", "stage")

    def profile(self, reps: int = 10) -> dict:
        """Per-kernel device time (ms) of the fused engine, HIP events on the
        solver's stream, each kernel launched `reps` times alone.  "sweeps" is
        the stage's x and y sweep as the step launches them (one kernel when
        their tilings share a wave count); "step_sum" adds it up per step.
        "cfl" is the step's CFL launch: with FP64 the one-wave dt reduction
        (the last RK combine of the previous step forms the maxima), with
        FP32-split the reduction over the state.  "ghost_fill" is 0 where the
        step's RK combine writes the ghost cells itself (mirror, copy and
        the double-Mach-reflection fills on grids whose edge tiles are at
        least 3 cells wide).  "stage" is one whole RK stage as the step
        launches it (sweeps, combine and the slab pipeline), and "step_sum"
        is then 3 x (stage + ghost_fill) + cfl.  With the blocked layout the
        sweeps and the stage are averaged over the three stages (they differ:
        the x sweep of stage 1 reads no U^n, the y sweep of stage 3 also
        reduces the next step's wave speeds)."""
        out = np.zeros(8)
        _lib.check(_lib.load().wenox_solver_profile(self._h, int(reps), out.ctypes.data, 8))
        return dict(zip(self.KERNELS, out.tolist()))

    @property
    def chunks(self):
        cx, cy = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().wenox_solver_plan(self._h, ctypes.byref(cx), ctypes.byref(cy)))
        return cx.value, cy.value

    def set_chunks(self, chunk_x: int, chunk_y: int):
        _lib.check(_lib.load().wenox_solver_set_chunks(self._h, int(chunk_x), int(chunk_y)))

    @property
    def tiling(self):
        """(chunk_x, chunk_y, waves_x, waves_y): rows marched per wavefront and
        wavefronts per workgroup of the x and y sweeps."""
        v = [ctypes.c_int() for _ in range(4)]
        _lib.check(_lib.load().wenox_solver_tiling(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    @property
    def combine_cols(self) -> int:
        """y extent (32 or 64) of the RK-combine tiles."""
        v = ctypes.c_int()
        _lib.check(_lib.load().wenox_solver_combine_cols(self._h, ctypes.byref(v)))
        return v.value

    def set_combine_cols(self, cols: int):
        _lib.check(_lib.load().wenox_solver_set_combine_cols(self._h, int(cols)))

    @property
    def pipeline(self) -> int:
        """Slabs of the fused FP64 fast stage whose RK combines overlap the
        next slab's sweeps (1: none)."""
        v = ctypes.c_int()
        _lib.check(_lib.load().wenox_solver_pipeline(self._h, ctypes.byref(v)))
        return v.value

    def set_pipeline(self, parts: int):
        _lib.check(_lib.load().wenox_solver_set_pipeline(self._h, int(parts)))

    @property
    def split_jump(self) -> float:
        """FP32-split: relative jump of density or energy across a face's
        stencils above which the face takes FP64 states."""
        v = ctypes.c_double()
        _lib.check(_lib.load().wenox_solver_split_jump(self._h, ctypes.byref(v)))
        return v.value

    def set_split_jump(self, tau: float):
        """Set the FP32-split jump threshold (> 0; inf disables the test)."""
        _lib.check(_lib.load().wenox_solver_set_split_jump(self._h, float(tau)))

    @property
    def blocked(self) -> bool:
        """Blocked state layout of the fused FP64 fast step (its default):
        both sweeps read one blocked copy of the state and the y sweep
        finishes each RK stage (no transposed copy, no combine launch); the
        same states bit for bit as the row-major step (set_blocked(False)),
        whose RK combine, combine tile width and slab pipeline apply only
        there."""
        v = ctypes.c_int()
        _lib.check(_lib.load().wenox_solver_layout(self._h, ctypes.byref(v)))
        return bool(v.value)

    def set_blocked(self, on: bool = True):
        _lib.check(_lib.load().wenox_solver_set_layout(self._h, int(bool(on))))

    def set_tiling(self, chunk_x: int, chunk_y: int, waves_x: int, waves_y: int):
        _lib.check(_lib.load().wenox_solver_set_tiling(self._h, int(chunk_x), int(chunk_y), int(waves_x),
                                                        int(waves_y)))

    def set_cfl_rule(self, rho_floor: float = 1e-14, cmax_eps: float = 1e-16):
        """dt = CFL dx / (max(sqrt(g p / max(rho, rho_floor)) + |V|) + cmax_eps);
        set before upload() for the first step to follow it too."""
        _lib.check(_lib.load().wenox_solver_set_cfl_rule(self._h, float(rho_floor), float(cmax_eps)))

    def set_pfloor(self, pmin: float = 1e-8):
        """After every step raise the total energy of each cell (ghosts
        included) whose pressure is below ``pmin`` so that p = pmin, as
        WENO_HLLC_2d.con_var2primi_pfloor does (:813-837); 0 switches it off."""
        _lib.check(_lib.load().wenox_solver_set_pfloor(self._h, float(pmin)))

    # (wavefronts per workgroup, rows per wavefront).  Dense in the row count:
    # the sweep time steps with the number of workgroup rounds on the 256 CUs
    # (two resident waves per SIMD), so one row more or less per wave can
    # move a grid across a round boundary.
    # slab counts of the pipelined FP64 stage the autotuner tries
    PIPELINE_PARTS = (1, 2, 3, 4, 6, 8)
    TILINGS = (tuple((1, c) for c in (8, 12, 16, 24, 32, 40, 48, 56, 64))
               + tuple((2, c) for c in range(4, 25, 2))
               + tuple((4, c) for c in range(3, 17)))

    def autotune(self, candidates=TILINGS, reps: int = 3, finalists: int = 4, layouts=None) -> dict:
        """Pick the sweep tiling -- (wavefronts per workgroup, rows per
        wavefront) for x and y -- with the lowest measured time on this
        device and keep it.  Each direction is timed over `candidates`; the
        `finalists` fastest of each (plus the model's default) are then timed
        in pairs as the step launches them (one kernel when the wave counts
        agree).  The default is kept unless a pair beats it by 3 %.  Then
        the RK-combine tile width and the slab pipeline (row-major layout).
        `layouts`: the state layouts to tune (False row-major, True blocked;
        default both where the blocked layout applies: fused engine, fast
        variant, FP64); the one whose tuned step (profile()["step_sum"]) is
        faster is kept, the current one unless the other wins by 1 %.
        Returns the kept layout's {"sweep_x": {cand: ms}, "sweep_y": {...},
        "pairs": {...}, "best": (cx, cy, wx, wy), "combine": {...},
        "pipeline": {...}} with "blocked" and "step_ms" {layout: ms}."""
        if layouts is None:
            d = self.desc
            ok = d.engine == ENGINE_FUSED and d.variant == FAST and d.precision == PREC_FP64
            layouts = (False, True) if ok else (self.blocked,)
        start = self.blocked
        runs = {}
        for lay in layouts:
            self.set_blocked(lay)
            res = self._autotune_layout(candidates, reps, finalists)
            res["step"] = self.profile(reps)["step_sum"]
            res["state"] = (self.tiling, self.combine_cols, self.pipeline)
            runs[lay] = res
        keep = min(runs, key=lambda l: runs[l]["step"])
        if start in runs and keep != start and runs[keep]["step"] > 0.99 * runs[start]["step"]:
            keep = start
        self.set_blocked(keep)
        tiling, cols, parts = runs[keep]["state"]
        self.set_tiling(*tiling)
        self.set_combine_cols(cols)
        self.set_pipeline(parts)
        out = {k: v for k, v in runs[keep].items() if k not in ("state", "step")}
        out["blocked"] = keep
        out["step_ms"] = {("blocked" if l else "row"): r["step"] for l, r in runs.items()}
        return out

    def _autotune_layout(self, candidates, reps, finalists) -> dict:
        cx0, cy0, wx0, wy0 = self.tiling
        per = {}
        for d, key in ((0, "sweep_x"), (1, "sweep_y")):
            times = {}
            start = (wx0, cx0) if d == 0 else (wy0, cy0)
            for w, c in tuple(candidates) + ((start,) if start not in candidates else ()):
                if d == 0:
                    self.set_tiling(c, cy0, w, wy0)
                else:
                    self.set_tiling(cx0, c, wx0, w)
                times[(w, c)] = self.profile(reps)[key]
            per[key] = times
        self.set_tiling(cx0, cy0, wx0, wy0)
        fx = sorted(per["sweep_x"], key=per["sweep_x"].get)[:finalists]
        fy = sorted(per["sweep_y"], key=per["sweep_y"].get)[:finalists]
        pairs = {}
        for (wx, cx) in fx + [(wx0, cx0)]:
            for (wy, cy) in fy + [(wy0, cy0)]:
                t = (cx, cy, wx, wy)
                if t in pairs:
                    continue
                self.set_tiling(*t)
                pairs[t] = self.profile(reps)["sweeps"]
        default = (cx0, cy0, wx0, wy0)
        best = min(pairs, key=pairs.get)
        if pairs[best] > 0.97 * pairs[default]:
            best = default
        self.set_tiling(*best)
        comb, pipe = {}, {}
        if not self.blocked:
            # RK-combine tile width (the blocked step has no combine launch)
            c0 = self.combine_cols
            for cols in (32, 64):
                self.set_combine_cols(cols)
                comb[cols] = self.profile(reps)["combine"]
            cb = min(comb, key=comb.get)
            self.set_combine_cols(cb if comb[cb] < 0.97 * comb[c0] else c0)
            # slab pipeline of the FP64 stage (its combines overlap the next slab's sweeps)
            if self.desc.precision == PREC_FP64 and self.desc.variant == FAST:
                for parts in self.PIPELINE_PARTS:
                    self.set_pipeline(parts)
                    pipe[parts] = self.profile(reps)["stage"]
                bp = min(pipe, key=pipe.get)
                self.set_pipeline(bp if pipe[bp] < 0.97 * pipe[1] else 1)
        return {"sweep_x": per["sweep_x"], "sweep_y": per["sweep_y"], "pairs": pairs, "best": best,
                "combine": comb, "pipeline": pipe}


class Batch:
    """Many independent problems of uneven sizes advanced together on the
    device, each with its own dx, time and CFL dt.  The fast FP64 batch can
    run persistently (set_persistent): one launch per step, or per eight
    steps, whose workgroups take the instances' work items from a device
    queue, each waiting only for its own instance's earlier phases.
    Otherwise every phase of a step is one launch over all instances.  A step
    (or eight persistent steps) is one hipGraph replay."""

    def __init__(self, sizes, dx, gamma=1.4, cfl=0.4, cons_switch=False, force_hlle=False,
                 jp_cri=(0.2, 1.0), bc="mirror", variant="fast", precision="fp64", chunk=16,
                 use_graph=True):
        d = _lib.SolverDesc()
        d.gamma, d.cfl = float(gamma), float(cfl)
        d.cons_switch, d.force_hlle = int(bool(cons_switch)), int(bool(force_hlle))
        d.jp_lo, d.jp_hi = float(jp_cri[0]), float(jp_cri[1])
        d.bc = _BC_NAMES[bc] if isinstance(bc, str) else int(bc)
        d.engine = ENGINE_FUSED
        d.variant = _VARIANTS[variant] if isinstance(variant, str) else int(variant)
        d.precision = {"fp64": PREC_FP64, "fp32split": PREC_FP32SPLIT}[precision] if isinstance(precision, str) else int(precision)
        d.use_graph = int(bool(use_graph))
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        self.n = len(self.sizes)
        nx = np.array([s[0] for s in self.sizes], dtype=np.int64)
        ny = np.array([s[1] for s in self.sizes], dtype=np.int64)
        dxa = np.broadcast_to(np.asarray(dx, dtype=np.float64), (self.n,)).copy()
        self._h = ctypes.c_void_p()
        _lib.check(_lib.load().wenox_batch_create(ctypes.byref(d), self.n, nx.ctypes.data, ny.ctypes.data,
                                                  dxa.ctypes.data, int(chunk), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            _lib.load().wenox_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, i, q, t0=0.0):
        q = np.asarray(q, dtype=np.float64)
        v = _lib.view(q)
        _lib.check(_lib.load().wenox_batch_upload(self._h, int(i), ctypes.byref(v), float(t0)))

    def download(self, i):
        nx, ny = self.sizes[i]
        out = np.empty((nx + 6, ny + 6, 8))
        _lib.check(_lib.load().wenox_batch_download(self._h, int(i), out.ctypes.data))
        return out

    def run(self, nstep=None, t_final=None):
        """Returns the number of instance-steps taken (summed over instances)."""
        ms = int(nstep) if nstep is not None else int(1e8)
        tf = float(t_final) if t_final is not None else -1.0
        k = ctypes.c_int64()
        _lib.check(_lib.load().wenox_batch_run(self._h, ms, tf, ctypes.byref(k)))
        return k.value

    def status(self):
        T = np.zeros(self.n)
        k = np.zeros(self.n, dtype=np.int64)
        _lib.check(_lib.load().wenox_batch_status(self._h, T.ctypes.data, k.ctypes.data))
        return T, k

    @property
    def ms_per_step(self) -> float:
        return _lib.load().wenox_batch_last_ms_per_step(self._h)

    @property
    def chunks(self):
        cx, cy = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().wenox_batch_chunks(self._h, ctypes.byref(cx), ctypes.byref(cy)))
        return cx.value, cy.value

    def set_chunks(self, chunk_x: int, chunk_y: int):
        _lib.check(_lib.load().wenox_batch_set_chunks(self._h, int(chunk_x), int(chunk_y)))

    @property
    def blocked(self) -> bool:
        """Blocked state layout of the batched FP64 fast step (its default),
        as Solver.blocked; the same states bit for bit as the row-major
        step (set_blocked(False))."""
        v = ctypes.c_int()
        _lib.check(_lib.load().wenox_batch_layout(self._h, ctypes.byref(v)))
        return bool(v.value)

    def set_blocked(self, on: bool = True):
        _lib.check(_lib.load().wenox_batch_set_layout(self._h, int(bool(on))))

    @property
    def persistent(self) -> bool:
        """Persistent step in effect (the default of the blocked fast FP64
        batch); the same states bit for bit as the launch-per-phase step."""
        v = ctypes.c_int()
        _lib.check(_lib.load().wenox_batch_persistent(self._h, ctypes.byref(v)))
        return bool(v.value)

    def set_persistent(self, on: bool = True):
        _lib.check(_lib.load().wenox_batch_set_persistent(self._h, int(bool(on))))

    @property
    def packs(self) -> tuple:
        """(pack_max, packs in effect): the persistent step lays the padded
        columns of same-shape instances side by side along a wavefront's
        lanes, in packs of at most pack_max members (0 packs when the
        persistent step does not run)."""
        m, n = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().wenox_batch_packs(self._h, ctypes.byref(m), ctypes.byref(n)))
        return m.value, n.value

    @property
    def persist_xcds(self) -> tuple:
        """(in use, probed): bit masks of the XCDs whose work queues the
        persistent step serves, and of those a probe launch reached at
        creation (fewer than eight on a partitioned device)."""
        u, p = ctypes.c_uint(), ctypes.c_uint()
        _lib.check(_lib.load().wenox_batch_persist_xcds(self._h, ctypes.byref(u), ctypes.byref(p)))
        return u.value, p.value

    def set_persist_xcds(self, mask: int):
        """Serve only the XCDs in `mask` (a non-empty subset of the probed
        ones; 0 restores the default); the same states bit for bit."""
        _lib.check(_lib.load().wenox_batch_set_persist_xcds(self._h, int(mask)))

    def set_split_jump(self, tau: float):
        """FP32-split jump threshold for every instance (Solver.set_split_jump)."""
        _lib.check(_lib.load().wenox_batch_set_split_jump(self._h, float(tau)))

    def set_pack_max(self, pack_max: int):
        """Members per pack at most (1: no packing); the same states bit for
        bit for every value."""
        _lib.check(_lib.load().wenox_batch_set_pack_max(self._h, int(pack_max)))

    def set_spin_limit(self, spins: int):
        """Polls of a persistent step's dependency wait before it gives up
        (run() then raises WenoxError); small values force that path."""
        _lib.check(_lib.load().wenox_batch_set_spin_limit(self._h, int(spins)))

    def profile_steps(self, steps: int = 8) -> float:
        """Device time (ms) per step of `steps` whole steps as run() launches
        them; the states and step controls are restored afterwards."""
        v = ctypes.c_double()
        _lib.check(_lib.load().wenox_batch_profile_steps(self._h, int(steps), ctypes.byref(v)))
        return v.value

    def profile(self, reps: int = 5) -> float:
        """Device time (ms) of one RK stage of every instance as a step
        launches it; the state is left unchanged."""
        v = ctypes.c_double()
        _lib.check(_lib.load().wenox_batch_profile(self._h, int(reps), ctypes.byref(v)))
        return v.value

    CHUNKS = (8, 12, 16, 24, 32, 48, 64, 96)

    def autotune(self, candidates=CHUNKS, reps: int = 3, steps: int = 8) -> dict:
        """Pick the rows marched per sweep workgroup (one value for the x and
        y sweeps of every instance) with the lowest measured stage time and
        keep it; the current chunks stay unless another beats them by 3 %.
        Where the persistent step applies (blocked fast FP64), each chunk is
        also timed as whole steps of both step structures (profile_steps) and
        the fastest (structure, chunk) pair is kept on the same 3 % rule.
        Returns {chunk: stage ms, "best": chunk} and, with the persistent
        step, "step_ms": {"phases" | "persistent": {chunk: ms per step}} and
        "persistent": the structure kept."""
        c0, p0 = self.chunks, self.persistent
        times = {}
        for c in candidates:
            self.set_chunks(c, c)
            times[c] = self.profile(reps)
        out = dict(times)
        modes = [False]
        if self.blocked:
            try:
                self.set_persistent(True)
                modes.append(True)
            except WenoxError:
                pass
        if len(modes) == 1:
            self.set_chunks(*c0)
            t0 = self.profile(reps)
            best = min(times, key=times.get)
            if times[best] < 0.97 * t0:
                self.set_chunks(best, best)
            out["best"] = self.chunks[0]
            return out
        step_ms = {}
        for m in modes:
            self.set_persistent(m)
            step_ms[m] = {}
            for c in candidates:
                self.set_chunks(c, c)
                step_ms[m][c] = self.profile_steps(steps)
        self.set_persistent(p0)
        self.set_chunks(*c0)
        t0 = self.profile_steps(steps)
        bm, bc = min(((m, c) for m in modes for c in candidates), key=lambda mc: step_ms[mc[0]][mc[1]])
        if step_ms[bm][bc] < 0.97 * t0:
            self.set_persistent(bm)
            self.set_chunks(bc, bc)
        out["best"] = self.chunks[0]
        out["persistent"] = self.persistent
        out["step_ms"] = {("persistent" if m else "phases"): {c: round(v, 5) for c, v in d.items()}
                          for m, d in step_ms.items()}
        return out
