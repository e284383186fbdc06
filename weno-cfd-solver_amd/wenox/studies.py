"""Accuracy studies run on the device-resident solver.

``vortex_study`` is the precision study of the isentropic vortex: periodic
[0,10]^2, vortex strength 5, freestream
(rho, u, v, p) = (1, 1, 1, 1), component-wise WENO5 with the 2-point Gauss
pass and HLLE, SSP-RK3 at CFL 0.4 to t_end.  For each grid it reports the
L1 and L-infinity errors of the density against the exact solution (cell
averages by 4x4 Gauss-Legendre quadrature), the FP32-split path's errors and
its distance from the FP64 run, and the observed order between grids.

The exact solution is the free-space vortex placed by nearest periodic image,
so it is not smooth across the lines halfway between images (the velocity
perturbation is ~2e-5 there): besides the whole-domain errors the study
reports ``L1_core``, the mean error over the cells within distance 2.5 of
the vortex centre: the seam lies 5 away, and sound (c ~ 1.2 in the frame
moving with the flow) from it does not reach that disc before t = 2.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import Solver, problems


def _run(n, t_end, precision, variant, cfl):
    q0, dx = problems.isentropic_vortex(n)
    # the twin variant (the NumPy twin's exact sqrt(3)/12 Gauss-point offset,
    # WENO_HLLC_2d.py:5) runs on the unfused engine
    engine = "unfused" if variant == "twin" else "fused"
    with Solver(n, n, dx, cfl=cfl, cons_switch=True, force_hlle=True, jp_cri=(0.2, 1.0),
                bc="periodic", engine=engine, variant=variant, precision=precision) as s:
        s.upload(q0)
        t0 = time.perf_counter()
        T, k = s.run(t_final=t_end)
        el = time.perf_counter() - t0
        q = s.download()[3:-3, 3:-3, :4]
    return q, T, k, el


def _core_mask(n, t, L=10.0, radius=2.5):
    """Cells whose centre lies within `radius` of the vortex centre at time t
    (nearest periodic image)."""
    h = L / n
    c = (np.arange(n) + 0.5) * h
    xc = (5.0 + t) % L
    d = (c - xc + 0.5 * L) % L - 0.5 * L
    return (d[:, None] ** 2 + d[None, :] ** 2) <= radius ** 2


def _orders(rows, prec):
    for a, b in zip(rows, rows[1:]):
        r = b["n"] / a["n"]
        for key in ("L1", "Linf", "L1_core"):
            b[prec][f"order_{key}"] = math.log(a[prec][key] / b[prec][key]) / math.log(r)


def vortex_study(sizes=(256, 512, 1024, 2048, 4096), t_end=2.0, cfl=0.4, variant="fast",
                 precisions=("fp64", "fp32split")):
    rows = []
    for n in sizes:
        exact = problems.vortex_cell_average(n, t_end)[..., 0]
        core = _core_mask(n, t_end)
        row = {"n": n}
        ref = None
        for prec in precisions:
            q, T, k, el = _run(n, t_end, prec, variant, cfl)
            rho = q[..., 0]
            err = np.abs(rho - exact)
            row[prec] = {"L1": float(err.mean()), "Linf": float(err.max()),
                         "L1_core": float(err[core].mean()), "Linf_core": float(err[core].max()),
                         "T": T, "steps": k,
                         "seconds": el, "cell_updates_per_s": n * n * k / el}
            if prec == "fp64":
                ref = rho
            elif ref is not None:
                row[prec]["rel_L1_vs_fp64"] = float(np.abs(rho - ref).sum() / np.abs(ref).sum())
        rows.append(row)
    for prec in precisions:
        _orders(rows, prec)
    return rows


def main(argv=None):
    """python -m wenox.studies [--sizes 256 512 ...] [--t-end 2.0] [--out study.json]"""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="isentropic vortex precision study on the GPU")
    ap.add_argument("--sizes", type=int, nargs="+", default=[256, 512, 1024, 2048, 4096])
    ap.add_argument("--t-end", type=float, default=2.0)
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--variant", default="fast", choices=["fast", "exact", "twin"],
                    help="twin: the exact sqrt(3)/12 of the NumPy twin in place of the Rust crate's 0.1443375673")
    ap.add_argument("--precisions", nargs="+", default=["fp64", "fp32split"])
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    precs = tuple(args.precisions)
    rows = []
    for n in args.sizes:
        # one grid at a time so progress shows while the larger grids run
        rows += vortex_study(sizes=(n,), t_end=args.t_end, cfl=args.cfl, variant=args.variant, precisions=precs)
        r = rows[-1]
        print(f"n={n} " + " ".join(f"{p} L1={r[p]['L1']:.3e} Linf={r[p]['Linf']:.3e} L1_core={r[p]['L1_core']:.3e} "
                                    f"steps={r[p]['steps']}" for p in precs), file=sys.stderr, flush=True)
    for prec in precs:
        _orders(rows, prec)
    out = json.dumps({"study": "isentropic vortex, component-wise WENO5 + Gauss pass + HLLE, SSP-RK3",
                      "t_end": args.t_end, "cfl": args.cfl, "variant": args.variant, "rows": rows})
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")
    print(out, flush=True)


if __name__ == "__main__":
    main()
