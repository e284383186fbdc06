"""The 1-D driver through the 2-D kernels (Rie_prob_1D_rs.py): Sod tube with
ny = 3, constant-extrapolation ghost fill, component-wise WENO5 + HLLE.

* Bitwise equal to the oracle's restatement of the driver (exact variant).
* Against the textbook exact Sod solution (Toro ch. 4): the density L1 error
  at t ~ 0.1 falls as the grid is refined (first order at the shock).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_1d_driver_matches_oracle(oracle_mod):
    from wenox import drivers, problems
    nx, ny = 200, 3
    grid = {"nx": nx, "ny": ny, "dx": 1 / nx}
    ctrl = {"nstep": 25, "CFL": 0.1}
    q, T = drivers.main_1d(grid, ctrl, engine="fused", variant="exact")
    ref, Tr = oracle_mod.run_1d(problems.sod_reference(nx, ny), nx, ny, 1 / nx, 1.4, 0.1, 25)
    assert T == Tr
    np.testing.assert_array_equal(q[..., :4], ref[..., :4])


def test_sod_converges_to_exact_solution():
    from wenox import Solver, problems
    errs = []
    for nx in (100, 200, 400):
        q0 = problems.sod_reference(nx, 3)
        dx = 1.0 / nx
        with Solver(nx, 3, dx, cfl=0.4, cons_switch=True, force_hlle=True, jp_cri=(1, 5), bc="copy",
                    engine="fused", variant="fast") as s:
            s.set_cfl_rule(1e-10, 0.0)
            s.upload(q0)
            T, k = s.run(t_final=0.1)
            q = s.download()
        # the reference's split sits at padded row int((nx + 6) / 2): x0 = (that - 3) dx
        x0 = (int((nx + 6) / 2) - 3) * dx
        xc = (np.arange(nx) + 0.5) * dx
        rho_ex = problems.sod_exact(xc, T, x0=x0)[0]
        errs.append(np.abs(q[3:-3, 4, 0] - rho_ex).mean())
    print("Sod density L1 errors", errs)
    assert errs[1] < 0.75 * errs[0] and errs[2] < 0.75 * errs[1]
    assert errs[2] < 5e-3
