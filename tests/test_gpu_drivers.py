"""GPU checks of the driver counterparts and the standard double Mach reflection.

* wenox.drivers.main_rs reproduces Rie_prob_2d_rs.main_rs's loop on the
  reference's own initial data and writes its snapshot files
  (data/<timestamp>/{step}.npy, time.npy) in the reference format.
* The standard DMR (Woodward-Colella boundary conditions, no reference
  counterpart) stays finite and positive and keeps the analytic post-shock
  state on the inflow and the exact shock trace on the top boundary.
"""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_main_rs_matches_oracle_and_writes_snapshots(oracle_mod, tmp_path, monkeypatch):
    from wenox import drivers, problems
    monkeypatch.chdir(tmp_path)
    grid = {"nx": 40, "ny": 40, "dx": 1 / 40}
    ctrl = {"nstep": 6, "CFL": 0.3, "file_storage": True, "force_hlle": False, "visualize": False,
            "mode": "opt", "jp_cri": (0.5, 2), "weno type": True}
    q = drivers.main_rs(grid, ctrl, {"gamma": 1.4}, engine="fused", variant="exact")
    q0 = problems.riemann_reference(40, 40)
    ref, ts = oracle_mod.run_riemann(q0.copy(), 40, 40, 1 / 40, 1.4, 0.3, False, True, (0.5, 2), nstep=6)
    np.testing.assert_array_equal(q[..., :4], ref[..., :4])
    folders = glob.glob(os.path.join("data", "*"))
    assert len(folders) == 1
    files = sorted(os.listdir(folders[0]))
    assert "time.npy" in files and "6.npy" in files and "0.npy" in files
    t = np.load(os.path.join(folders[0], "time.npy"))
    np.testing.assert_allclose(t, ts, rtol=0, atol=0)
    last = np.load(os.path.join(folders[0], "6.npy"))
    assert last.shape == (46, 46, 8)
    np.testing.assert_array_equal(last, q)


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_standard_dmr(precision):
    from wenox import Solver, problems
    nx, ny = 240, 60
    q0, dx = problems.double_mach(nx, ny)
    # CFL 0.1 as the reference's own DMR driver (double_mach_reflection.py:17);
    # at CFL 0.4 the reference scheme turns non-finite near the top boundary
    with Solver(nx, ny, dx, cfl=0.1, jp_cri=(1e4, 1.1e4), bc="dmr", engine="fused", variant="fast",
                precision=precision) as s:
        s.upload(q0)
        T, k = s.run(nstep=60)
        q = s.download()
    assert k == 60
    rho = q[..., 0]
    assert np.all(np.isfinite(q)) and rho.min() > 0.5 and rho.max() < 40.0
    # inflow ghost rows hold the post-shock state
    np.testing.assert_allclose(q[0, 10, :4], q0[0, 10, :4], rtol=1e-12)
    # top ghost row: post-shock left of the exact shock trace at time T, pre-shock right of it
    xs = 1 / 6 + (1 + 20 * T) / np.sqrt(3)
    xc = (np.arange(nx) + 0.5) * dx
    top = q[3:-3, ny + 3, 0]
    # FP32-split stores hi + lo (48 significant bits): 1.4 comes back to ~4e-15
    tol = 0.0 if precision == "fp64" else 1e-14
    assert np.all(np.abs(top[xc < xs - 2 * dx] - 8.0) <= tol * 8.0)
    assert np.all(np.abs(top[xc > xs + 2 * dx] - 1.4) <= tol * 1.4)
