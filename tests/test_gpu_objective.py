"""GPU parity of the jp_cri tuning objective (metrics.py:1-77) and of the
tuning loop around it (opt_jp_cro.py:9-80) against the oracle.

Bars: the integer and selection work -- shock mask, skeleton, band, the 7 x 7
median -- bitwise; J and its metrics within 1e-10 relative (floating sums
reduced in a different order than numpy's pairwise sums; the line fit from
exact integer normal equations instead of an SVD least-squares solve).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-10
KEYS = ("I_curve", "E_perp", "O", "TV", "cI", "cE", "cO", "cTV")


def _shock_fields(nx, ny, seed):
    """A curved shock (circle), a slanted one and smooth velocities."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx)[:, None] + 0.5) / nx
    y = (np.arange(ny)[None, :] + 0.5) / ny
    rho = 1.0 + 2.0 * ((x - 0.4) ** 2 + (y - 0.55) ** 2 < 0.09) + 0.8 * (y > 0.3 + 0.5 * x)
    rho = rho + 0.05 * np.sin(7 * x) * np.cos(5 * y) + 0.01 * rng.random((nx, ny))
    u = 0.3 + 0.2 * np.cos(3 * y) + 0.02 * rng.standard_normal((nx, ny))
    v = -0.1 + 0.4 * np.sin(2 * x) + 0.02 * rng.standard_normal((nx, ny))
    return rho, u, v


def _check(rho, u, v, dx=1.0, dy=1.0, weights=None):
    from oracle import objective as ob
    from wenox import tuning
    S, C, band, med = tuning.objective_fields(rho, dx, dy)
    oS, oC, oband, omed = ob.objective_fields(rho, dx, dy)
    np.testing.assert_array_equal(S, oS)
    np.testing.assert_array_equal(band, oband)
    np.testing.assert_array_equal(C, oC)
    np.testing.assert_array_equal(med, omed)
    J, m = tuning.compute_2d_objective(rho, u, v, None, dx, dy, weights)
    oJ, om = ob.compute_2d_objective(rho, u, v, None, dx, dy, weights)
    for k in KEYS:
        assert m[k] == pytest.approx(om[k], rel=RTOL, abs=1e-12), k
    assert J == pytest.approx(oJ, rel=RTOL)
    return J, m


@pytest.mark.parametrize("shape", [(64, 48), (200, 120), (37, 3), (5, 5)])
def test_objective_matches_oracle(shape):
    nx, ny = shape
    print(f"[objective] shock fields {nx}x{ny}", flush=True)
    rho, u, v = _shock_fields(nx, ny, 11 + nx)
    _check(rho, u, v)
    _check(rho, u, v, dx=0.5, dy=2.0, weights={"I": 2.0, "E": 0.5, "O": 10.0, "TV": 3.0})


def test_objective_straight_shock_and_strided_views():
    nx, ny = 32, 48
    rho = np.ones((nx, ny))
    rho[:, 20:] = 3.0
    q = np.zeros((nx + 6, ny + 6, 8))
    q[3:-3, 3:-3, 0] = rho
    q[3:-3, 3:-3, 5] = 0.7
    q[3:-3, 3:-3, 6] = 0.4
    print("[objective] straight shock, strided views of an 8-channel array", flush=True)
    J, m = _check(q[3:-3, 3:-3, 0], q[3:-3, 3:-3, 5], q[3:-3, 3:-3, 6])
    assert m["O"] == 0.0 and m["I_curve"] <= 1e-12


def test_solver_state_objective(oracle_mod):
    """The objective of a solver state computed where it lives equals the one
    of its downloaded arrays bitwise, and the oracle's to 1e-10."""
    from oracle import objective as ob
    from wenox import Solver, problems, tuning
    n = 128
    q0, dx = problems.four_quadrant(n)
    print(f"[objective] four-quadrant {n}^2 after 60 steps", flush=True)
    with Solver(n, n, dx, cfl=0.4, jp_cri=(1e4, 1.1e4), engine="fused") as s:
        s.upload(q0)
        s.run(nstep=60)
        q = s.download()
        J_dev, m_dev = tuning.solver_objective(s)
    rho, u, v = q[3:-3, 3:-3, 0], q[3:-3, 3:-3, 5], q[3:-3, 3:-3, 6]
    J_host, m_host = _check(rho, u, v)
    assert J_dev == J_host
    assert m_dev == m_host
    assert m_host["I_curve"] > 0 and m_host["E_perp"] > 0 and m_host["O"] > 0


def test_evaluate_jp_pair_matches_oracle(oracle_mod):
    from oracle import objective as ob
    from wenox import problems, tuning
    nx = ny = 48
    grid = {"nx": nx, "ny": ny, "dx": 1.0 / nx}
    ctrl = {"CFL": 0.4, "t_final": 0.03, "weno type": False, "force_hlle": False}
    phys = {"gamma": 1.4}
    print("[objective] evaluate_jp_pair 48^2 t_final 0.03", flush=True)
    score = tuning.evaluate_jp_pair(0.2, 1.0, grid, ctrl, phys)
    q0 = problems.riemann_reference(nx, ny)
    ref, _ = oracle_mod.run_riemann(q0, nx, ny, 1.0 / nx, 1.4, 0.4, False, False, (0.2, 1.0), t_final=0.03)
    want, _ = ob.compute_2d_objective(ref[3:-3, 3:-3, 0], ref[3:-3, 3:-3, 5], ref[3:-3, 3:-3, 6], None)
    assert score == pytest.approx(want, rel=RTOL)
    assert tuning.evaluate_jp_pair(1.0, 0.5, grid, ctrl, phys) == tuning.PENALTY


def test_optimize_jp_runs():
    from wenox import tuning
    grid = {"nx": 32, "ny": 32, "dx": 1.0 / 32}
    ctrl = {"CFL": 0.4, "nstep": 8, "weno type": False}
    print("[objective] optimize_jp 32^2, 6 runs", flush=True)
    lo, hi, best = tuning.optimize_jp(grid, ctrl, {"gamma": 1.4}, n_calls=6, n_initial_points=3)
    assert 0.02 <= lo <= 1.0 and 0.2 <= hi <= 10.0
    assert np.isfinite(best)
