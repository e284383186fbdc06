"""The precision study at its configured sizes: isentropic vortex on the
periodic [0,10]^2, grids 512^2 .. 4096^2 to t = 2 (component-wise WENO5 with
the 2-point Gauss pass + HLLE, SSP-RK3 at CFL 0.4), FP64 and FP32-split.

Each grid reports the L1 and L-infinity density errors against the exact
solution (whole domain, and the disc of radius 2.5 around the vortex centre
that the non-periodic seam of the exact solution does not reach by t = 2)
and the FP32-split distance from FP64.  Asserted: FP32-split within 1e-6
relative L1 of FP64 on every grid; FP64 and FP32-split errors decrease at
every refinement; the observed L1 order of the core error from 512^2 to
1024^2 is at least 3 (the scheme is fifth order in space and third in time).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = (512, 1024, 2048, 4096)
ROWS = {}


@pytest.mark.parametrize("n", SIZES)
def test_vortex_grid(n):
    from wenox.studies import vortex_study
    print(f"[vortex] {n}^2 to t=2, FP64 and FP32-split", flush=True)
    row = vortex_study(sizes=(n,), t_end=2.0)[0]
    ROWS[n] = row
    for prec in ("fp64", "fp32split"):
        r = row[prec]
        print(f"[vortex] {n} {prec}: L1 {r['L1']:.3e} Linf {r['Linf']:.3e} L1_core {r['L1_core']:.3e} "
              f"steps {r['steps']} T {r['T']:.15g}", flush=True)
        assert abs(r["T"] - 2.0) <= 1e-12
    assert row["fp32split"]["rel_L1_vs_fp64"] <= 1e-6, row["fp32split"]["rel_L1_vs_fp64"]


def test_vortex_orders():
    if len(ROWS) < len(SIZES):
        pytest.skip("needs every grid of the sequence in this session")
    from wenox.studies import _orders
    rows = [ROWS[n] for n in SIZES]
    for prec in ("fp64", "fp32split"):
        _orders(rows, prec)
        for a, b in zip(rows, rows[1:]):
            assert b[prec]["L1"] < a[prec]["L1"] and b[prec]["L1_core"] < a[prec]["L1_core"]
        assert rows[1][prec]["order_L1_core"] >= 3.0, rows[1][prec]
    out = os.environ.get("WENOX_VORTEX_OUT")
    if out:
        with open(out, "w") as f:
            json.dump({"rows": rows}, f)
    for r in rows[1:]:
        print(r["n"], {p: {k: round(v, 3) for k, v in r[p].items() if k.startswith("order")}
                       for p in ("fp64", "fp32split")}, flush=True)
