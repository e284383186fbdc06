"""Write golden vectors of the reference's own Python drivers.

Run in the build container only (the reference tree is not present anywhere
else):  python tests/golden/make_driver_golden.py [/path/to/reference]

The reference's time-stepping drivers -- ghost fills, primitive sync,
SSP-RK3 stage combination, CFL time step and the time loops -- are plain
numpy around calls into the compiled module ``weno_ext``.  Nothing but a run
of those drivers pins the oracle's restatement of them (the order of the
ghost-fill writes, where the primitive sync sits, the RK operation order, the
CFL rule and the loop control).  This script imports the driver modules with
``weno_ext`` bound to the oracle (same call surface, ``lib.rs:18-153``), so
every value below is the reference driver's own arithmetic around the
oracle's kernels, and the CPU tests hold the oracle's drivers and the
package's initial data to them bitwise.

Reference functions run (file:line):
  Rie_prob_2d_rs.apply_bc_zero_gradient :123-152, apply_bc_reflecting :42-120,
  _sync_primitives :157-163, rk3_cc :169-207, init :213-263, main_rs :269-376
  (its nested _cfl_dt :311-319 and both loop forms :328-365);
  unitest.init_refac :169-205, apply_bc_refac :207-227;
  double_mach_reflection.rk3_refac :97-134, main_rs :177-303 (t_final loop
  :229-260);
  Rie_prob_1D_rs.rk3 :20-57, apply_bc_corrected :59-76, init :78-104,
  main :106-175.  That module runs a 500-cell simulation at import
  (``main(grid_dict, control_dict)`` at :208), so its function definitions
  are taken from the source with ``ast`` and executed without the module's
  top-level statements.

The time loops are run with ``file_storage`` off and ``mode`` "opt" (the
loop returns the final state); the ``dt`` sequence is recorded by wrapping
the RK function the loop calls.
"""
import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEED = 20261203


def _rand_state(rng, shape):
    q = rng.uniform(-2.0, 2.0, shape)
    q[..., 0] = rng.uniform(0.2, 2.0, shape[:2])
    return q


def _load_1d(ref, oracle):
    """Function definitions of Rie_prob_1D_rs.py without its top-level run."""
    path = os.path.join(ref, "Rie_prob_1D_rs.py")
    tree = ast.parse(open(path).read(), filename=path)
    defs = [n for n in tree.body if isinstance(n, ast.FunctionDef)]
    mod = ast.Module(body=defs, type_ignores=[])
    ns = {"np": np, "weno_ext": oracle, "interactive_plot_keyboard": lambda *a, **k: None,
          "__name__": "Rie_prob_1D_rs_defs"}
    exec(compile(mod, path, "exec"), ns)
    return ns


def main(ref):
    sys.path.insert(0, ROOT)
    import oracle
    oracle.lib()
    os.environ.setdefault("MPLBACKEND", "Agg")
    sys.modules["weno_ext"] = oracle
    sys.path.insert(0, ref)
    import Rie_prob_2d_rs as R
    import unitest as U
    import double_mach_reflection as D

    rng = np.random.default_rng(SEED)
    out = {}

    # ghost fills on random 8-channel arrays
    for k, shape in enumerate([(22, 18, 8), (16, 16, 8), (9, 30, 8)]):
        a = _rand_state(rng, shape)
        out[f"zg_in_{k}"] = a.copy()
        out[f"zg_out_{k}"] = R.apply_bc_zero_gradient(a.copy(), {})
    for k, n in enumerate([22, 15]):
        a = _rand_state(rng, (n, n, 8))
        out[f"refl_in_{k}"] = a.copy()
        out[f"refl_out_{k}"] = R.apply_bc_reflecting(a.copy(), {})

    # primitive sync
    a = _rand_state(rng, (14, 12, 8))
    out["sync_in"] = a.copy()
    out["sync_out"] = R._sync_primitives(a.copy(), 1.4)

    # reference initial data
    for n in (16, 13, 20):
        out[f"init_rie_{n}"] = R.init({"nx": n, "ny": n, "dx": 1 / n}, {}, 1.4)
    dmr_grids = [(40, 10, 1 / 100), (80, 20, 1 / 20), (120, 30, 1 / 30)]
    for k, (nx, ny, dx) in enumerate(dmr_grids):
        q, bound = U.init_refac({"nx": nx, "ny": ny, "x": 4, "y": 1, "dx": dx}, {}, 1.4)
        out[f"init_refac_q_{k}"] = q
        out[f"init_refac_bound_{k}"] = np.array(bound)
        out[f"init_refac_grid_{k}"] = np.array([nx, ny, dx])
    # the DMR fill with the wall boundary at 0, 1 and inside the row
    for bound in (0, 1, 17):
        a = _rand_state(rng, (46, 16, 8))
        out[f"bcrefac_in_{bound}"] = a.copy()
        out[f"bcrefac_out_{bound}"] = U.apply_bc_refac(a.copy(), bound, {"nx": 40, "ny": 10, "x": 4, "y": 1,
                                                                            "dx": 0.01})

    # one SSP-RK3 step of each driver
    n = 20
    g = {"nx": n, "ny": n, "dx": 1 / n}
    q0 = R.init(g, {}, 1.4)
    out["rk3cc_in"] = q0.copy()
    for cons in (False, True):
        out[f"rk3cc_out_cons{int(cons)}"] = R.rk3_cc(q0.copy(), g, 0.004, 1.4, False, cons, (0.5, 2.0))
    out["rk3cc_out_hlle"] = R.rk3_cc(q0.copy(), g, 0.004, 1.4, True, False, (0.5, 2.0))
    gd = {"nx": 40, "ny": 10, "x": 4, "y": 1, "dx": 1 / 100}
    qd, bd = U.init_refac(gd, {}, 1.4)
    out["rk3refac_in"] = qd.copy()
    out["rk3refac_bound"] = np.array(bd)
    out["rk3refac_out"] = D.rk3_refac(qd.copy(), gd, 2e-4, 1.4, False, U.apply_bc_refac, bd, (1e4, 1.1e4))

    # time loops: dt sequence and final state
    def record(mod, name):
        orig = getattr(mod, name)
        dts = []

        def wrapped(q, grid, dt, *a, **k):
            dts.append(dt)
            return orig(q, grid, dt, *a, **k)
        setattr(mod, name, wrapped)
        return dts, lambda: setattr(mod, name, orig)

    runs = {
        "loop_nstep": dict(nstep=5, CFL=0.4, cons=False, jp=(0.5, 2.0), hlle=False),
        "loop_tfinal": dict(nstep=50, t_final=0.06, CFL=0.4, cons=True, jp=(1e4, 1.1e4), hlle=False),
    }
    for key, c in runs.items():
        ctrl = {"nstep": c["nstep"], "CFL": c["CFL"], "file_storage": False, "force_hlle": c["hlle"],
                "visualize": False, "mode": "opt", "jp_cri": c["jp"], "weno type": c["cons"]}
        if "t_final" in c:
            ctrl["t_final"] = c["t_final"]
        dts, undo = record(R, "rk3_cc")
        qf = R.main_rs({"nx": 16, "ny": 16, "dx": 1 / 16}, ctrl, {"gamma": 1.4})
        undo()
        out[f"{key}_q"] = qf
        out[f"{key}_dt"] = np.array(dts)
        out[f"{key}_ctrl"] = np.array([c["nstep"], c.get("t_final", -1.0), c["CFL"], float(c["cons"]),
                                       c["jp"][0], c["jp"][1]])
    dts, undo = record(D, "rk3_refac")
    ctrl = {"nstep": 400, "CFL": 0.1, "file_storage": False, "force_hlle": False, "visualize": False,
            "mode": "opt", "t_final": 0.003, "jp_cri": (1e4, 1.1e4)}
    qf = D.main_rs({"nx": 40, "ny": 10, "dx": 1 / 100, "x": 4, "y": 1}, ctrl, {"gamma": 1.4})
    undo()
    out["loop_dmr_q"] = qf
    out["loop_dmr_dt"] = np.array(dts)

    # the 1-D driver (Rie_prob_1D_rs)
    ns = _load_1d(ref, oracle)
    g1 = {"nx": 64, "ny": 3, "dx": 1 / 64}
    ns["grid_dict"] = g1          # main() passes the module-level grid_dict to rk3 (:163)
    out["init_1d"] = ns["init"](g1, {}, 1.4)
    a = _rand_state(rng, (70, 9, 8))
    out["bc1d_in"] = a.copy()
    out["bc1d_out"] = ns["apply_bc_corrected"](a.copy(), g1)
    q1 = out["init_1d"].copy()
    out["rk3_1d_out"] = ns["rk3"](q1.copy(), g1, 0.002, 1.4, True, (1, 5))
    orig = ns["rk3"]
    rec = {"dt": [], "q": None}

    def rk3_rec(q, grid, dt, gamma, force_hlle, jp_cri):
        rec["dt"].append(dt)
        rec["q"] = orig(q, grid, dt, gamma, force_hlle, jp_cri)
        return rec["q"]
    ns["rk3"] = rk3_rec
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        ns["main"](g1, {"nstep": 7, "CFL": 0.1}, 1.4, file_storage=False)
    out["loop_1d_q"] = rec["q"]
    out["loop_1d_dt"] = np.array(rec["dt"])

    path = os.path.join(HERE, "driver_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
