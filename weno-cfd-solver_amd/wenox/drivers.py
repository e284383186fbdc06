"""Device-resident counterparts of the reference drivers.

``main_rs`` follows ``Rie_prob_2d_rs.main_rs`` (:269-376),
``main_rs_dmr`` follows ``double_mach_reflection.main_rs`` (:177-303),
``main_1d`` follows ``Rie_prob_1D_rs.main`` (:110-181) and ``main_twin``
follows the pure-NumPy ``WENO_HLLC_2d.main`` (:840-908), each with the
same dictionaries, loop control (t_final primary, nstep as a cap; CFL dt
clipped to the remaining time) and snapshot format
(``data/<YYYY-mm-dd_HH-MM-SS>/{step}.npy`` of the (nx+6, ny+6, 8) array and
``time.npy``), so ``plot_utils.interactive_plot_keyboard`` reads the output.
Every step runs on the GPU; snapshots are the only host copies: the (nx+6,
ny+6, 8) array is formed on the device and written to disk by a background
thread while the next steps run.
"""
from __future__ import annotations

import datetime
import os
import queue
import threading

import numpy as np

from . import Solver, problems


class _SnapshotWriter:
    """Writes snapshot files on a background thread (numpy releases the GIL
    while it writes), so the next GPU steps and downloads overlap the disk
    writes.  At most `depth` frames wait in memory; errors surface on put()
    or close()."""

    def __init__(self, depth=2):
        self._q = queue.Queue(maxsize=depth)
        self._err = None
        self._t = threading.Thread(target=self._work, daemon=True)
        self._t.start()

    def _work(self):
        while True:
            item = self._q.get()
            if item is None:
                return
            if self._err is None:
                try:
                    np.save(*item)
                except BaseException as e:  # reported to the stepping thread
                    self._err = e

    def _check(self):
        if self._err is not None:
            raise self._err

    def put(self, path, array):
        self._check()
        self._q.put((path, array))

    def close(self):
        self._q.put(None)
        self._t.join()
        self._check()


def _run(solver, q0, control_dict, nx, ny, snapshot_dir):
    t_final = control_dict.get("t_final", None)
    nstep_user = control_dict.get("nstep", None)
    if t_final is None and nstep_user is None:
        raise ValueError("Either 't_final' or 'nstep' must be provided.")
    max_steps = int(nstep_user) if nstep_user is not None else int(1e8)
    solver.upload(q0)
    t_list = [0.0]
    if snapshot_dir:
        # one step per call so every state can be written (Rie_prob_2d_rs.py:342-343)
        writer = _SnapshotWriter()
        try:
            writer.put(os.path.join(snapshot_dir, "0.npy"), np.array(q0, dtype=np.float64, copy=True))
            for step in range(1, max_steps + 1):
                T, k = solver.run(nstep=1, t_final=t_final)
                if k == 0:
                    break
                writer.put(os.path.join(snapshot_dir, f"{step}.npy"), solver.download())
                t_list.append(T)
        finally:
            writer.close()
        np.save(os.path.join(snapshot_dir, "time.npy"), np.array(t_list))
    else:
        solver.run(nstep=max_steps, t_final=t_final)
    return solver.download()


def _snapshot_dir(control_dict):
    if not bool(control_dict.get("file_storage", False)):
        return ""
    folder = datetime.datetime.now().strftime("%Y-%m-%d_%H-%M-%S")
    path = os.path.join("data", folder)
    os.makedirs(path, exist_ok=True)
    return path


def main_rs(grid_dict: dict, control_dict: dict, phys_dict: dict, q0=None, engine="fused",
            variant="exact"):
    """Rie_prob_2d_rs.main_rs on the device.  q0 defaults to the reference's
    own initial data (Rie_prob_2d_rs.init, :213-263); the ghost fill is the
    even mirror of apply_bc_zero_gradient (:123-152).  Returns the final
    (nx+6, ny+6, 8) state when control_dict["mode"] == "opt" (as the
    reference) and always when q0 is given."""
    gamma = phys_dict.get("gamma", 1.4)
    nx, ny = int(grid_dict["nx"]), int(grid_dict["ny"])
    dx = float(grid_dict.get("dx", 1.0))
    given = q0 is not None
    if q0 is None:
        q0 = problems.riemann_reference(nx, ny, gamma)
    with Solver(nx, ny, dx, gamma=gamma, cfl=float(control_dict.get("CFL", 0.5)),
                cons_switch=bool(control_dict.get("weno type")),
                force_hlle=bool(control_dict.get("force_hlle", False)),
                jp_cri=control_dict.get("jp_cri", (0.2, 1.0)), bc="mirror",
                engine=engine, variant=variant) as s:
        q = _run(s, q0, control_dict, nx, ny, _snapshot_dir(control_dict))
    if given or control_dict.get("mode", "") == "opt":
        return q
    return None


def main_1d(grid_info: dict, control_dict: dict, gamma: float = 1.4, file_storage: bool = False,
            engine="fused", variant="exact"):
    """Rie_prob_1D_rs.main (:110-181) on the device: the Sod tube through the
    2D kernels with ny = 3, constant-extrapolation ghost fill
    (apply_bc_corrected, :59-76), component-wise reconstruction and HLLE
    (its rk3 calls l_local_py(..., force_hlle=True, jp_cri=(1, 5), True)),
    CFL dt = CFL dx / max(sqrt(g p / max(rho, 1e-10)) + |V|) and
    control_dict["nstep"] - 1 steps.  Returns (q, time list)."""
    nx, ny = int(grid_info["nx"]), int(grid_info["ny"])
    dx = float(grid_info["dx"])
    q0 = problems.sod_reference(nx, ny, gamma)
    ctrl = {"nstep": int(control_dict["nstep"]) - 1, "file_storage": file_storage}
    with Solver(nx, ny, dx, gamma=gamma, cfl=float(control_dict["CFL"]), cons_switch=True,
                force_hlle=True, jp_cri=(1, 5), bc="copy", engine=engine, variant=variant) as s:
        s.set_cfl_rule(1e-10, 0.0)
        snapshot = _snapshot_dir(ctrl)
        q = _run(s, q0, ctrl, nx, ny, snapshot)
        t = s.t
    return q, t


def main_rs_dmr(grid_dict: dict, control_dict: dict, phys_dict: dict, engine="fused",
                variant="exact"):
    """double_mach_reflection.main_rs on the device: unitest.init_refac data
    and apply_bc_refac ghost fill, characteristic reconstruction (its
    5-argument l_local calls), CFL clipped to t_final."""
    gamma = phys_dict.get("gamma", 1.4)
    nx, ny = int(grid_dict["nx"]), int(grid_dict["ny"])
    dx = float(grid_dict.get("dx", 1.0))
    q0, bound = problems.dmr_reference(nx, ny, grid_dict.get("x", 4), grid_dict.get("y", 1), gamma)
    jp = control_dict.get("jp_cri", None) or (0.2, 1.0)
    with Solver(nx, ny, dx, gamma=gamma, cfl=float(control_dict.get("CFL", 0.5)),
                cons_switch=False, force_hlle=bool(control_dict.get("force_hlle", False)),
                jp_cri=jp, bc="dmr_ref", dmr_bound=bound, engine=engine, variant=variant) as s:
        q = _run(s, q0, control_dict, nx, ny, _snapshot_dir(control_dict))
    if control_dict.get("mode", "") == "opt":
        return q
    return None


def main_twin(Ngrid: int, N_STEP: int, fileStorage: bool = True, force_HLLE: bool = False,
              Jp_cri=(1.0, 4.0), gamma: float = 1.4):
    """WENO_HLLC_2d.main (:840-908) on the device, in the twin's semantics
    (exact sqrt(3)/12, its HLLC and HLLE; component-wise reconstruction):
    the quadrant initial data of problems.twin_quadrants, dx = 1/Ngrid,
    CFL 0.2 with dt = CFL dx / max(sqrt(g p / max(rho, 1e-10)) + |V|),
    RK3_correct with the constant-extrapolation ghost fill, the pressure floor
    1e-8 after every step, and N_STEP - 1 steps.  Snapshots (fileStorage) as
    the twin writes them; its closing interactive plot is left to the caller.
    Returns the final (Ngrid+6, Ngrid+6, 8) state."""
    n = int(Ngrid)
    q0 = problems.twin_quadrants(n, gamma)
    ctrl = {"nstep": max(int(N_STEP) - 1, 0), "file_storage": bool(fileStorage)}
    with Solver(n, n, 1 / n, gamma=gamma, cfl=0.2, cons_switch=True, force_hlle=bool(force_HLLE),
                jp_cri=tuple(Jp_cri), bc="copy", engine="unfused", variant="twin") as s:
        s.set_cfl_rule(1e-10, 0.0)
        s.set_pfloor(1e-8)
        return _run(s, q0, ctrl, n, n, _snapshot_dir(ctrl))
