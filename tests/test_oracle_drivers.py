"""The oracle's drivers and the package's reference initial data against the
reference's own Python drivers (CPU).

``tests/golden/driver_golden.npz`` holds outputs of the reference driver
functions run with ``weno_ext`` bound to the oracle
(``tests/golden/make_driver_golden.py``): ghost fills
(``Rie_prob_2d_rs.py:42-152``, ``unitest.py:207-227``,
``Rie_prob_1D_rs.py:59-76``), primitive sync (``Rie_prob_2d_rs.py:157-163``),
one SSP-RK3 step of each driver (``Rie_prob_2d_rs.py:169-207``,
``double_mach_reflection.py:97-134``, ``Rie_prob_1D_rs.py:20-57``), the
initial data (``Rie_prob_2d_rs.py:213-263``, ``unitest.py:169-205``,
``Rie_prob_1D_rs.py:78-104``) and the time loops with their CFL time steps
(``Rie_prob_2d_rs.py:311-365``, ``double_mach_reflection.py:229-260``,
``Rie_prob_1D_rs.py:159-175``).  Every comparison is bitwise.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "driver_golden.npz")


@pytest.fixture(scope="module")
def gd():
    return np.load(GOLDEN, allow_pickle=False)


def _times(dts):
    """T after each step, accumulated as the reference loops do (T += dt)."""
    T, ts = 0.0, [0.0]
    for dt in dts.tolist():
        T += dt
        ts.append(T)
    return ts


def _eq(a, b):
    np.testing.assert_array_equal(a, b)
    assert a.shape == b.shape


def test_zero_gradient_fill(gd, oracle_mod):
    from wenox import problems
    for k in range(3):
        _eq(oracle_mod.apply_bc_zero_gradient(gd[f"zg_in_{k}"].copy()), gd[f"zg_out_{k}"])
        _eq(problems._mirror_fill(gd[f"zg_in_{k}"].copy()), gd[f"zg_out_{k}"])


def test_reflecting_fill(gd, oracle_mod):
    from wenox import problems
    for k in range(2):
        _eq(oracle_mod.apply_bc_reflecting(gd[f"refl_in_{k}"].copy()), gd[f"refl_out_{k}"])
        _eq(problems._reflecting_fill(gd[f"refl_in_{k}"].copy()), gd[f"refl_out_{k}"])


def test_sync_primitives(gd, oracle_mod):
    _eq(oracle_mod.sync_primitives(gd["sync_in"].copy(), 1.4), gd["sync_out"])


def test_reference_initial_data(gd):
    from wenox import problems
    for n in (16, 13, 20):
        _eq(problems.riemann_reference(n, n), gd[f"init_rie_{n}"])
    for k in range(3):
        nx, ny, _ = gd[f"init_refac_grid_{k}"]
        q, bound = problems.dmr_reference(int(nx), int(ny))
        _eq(q, gd[f"init_refac_q_{k}"])
        assert bound == int(gd[f"init_refac_bound_{k}"])
    _eq(problems.sod_reference(64, 3), gd["init_1d"])


@pytest.mark.parametrize("bound", [0, 1, 17])
def test_dmr_reference_fill(gd, oracle_mod, bound):
    from wenox import problems
    _eq(oracle_mod.apply_bc_refac(gd[f"bcrefac_in_{bound}"].copy(), bound, 40, 10), gd[f"bcrefac_out_{bound}"])
    _eq(problems.dmr_reference_fill(gd[f"bcrefac_in_{bound}"].copy(), bound, 40, 10), gd[f"bcrefac_out_{bound}"])


def test_copy_fill(gd, oracle_mod):
    from wenox import problems
    _eq(oracle_mod.apply_bc_copy(gd["bc1d_in"].copy(), 64, 3), gd["bc1d_out"])
    _eq(problems._copy_fill(gd["bc1d_in"].copy()), gd["bc1d_out"])


def test_rk3_cc_step(gd, oracle_mod):
    q0 = gd["rk3cc_in"]
    for cons in (0, 1):
        got = oracle_mod.rk3_cc(q0.copy(), 20, 20, 1 / 20, 0.004, 1.4, False, bool(cons), (0.5, 2.0))
        _eq(got, gd[f"rk3cc_out_cons{cons}"])
    got = oracle_mod.rk3_cc(q0.copy(), 20, 20, 1 / 20, 0.004, 1.4, True, False, (0.5, 2.0))
    _eq(got, gd["rk3cc_out_hlle"])


def test_rk3_refac_step(gd, oracle_mod):
    bound = int(gd["rk3refac_bound"])
    got = oracle_mod.rk3_refac(gd["rk3refac_in"].copy(), 40, 10, 1 / 100, 2e-4, 1.4, False, bound, (1e4, 1.1e4))
    _eq(got, gd["rk3refac_out"])


def test_rk3_1d_step(gd, oracle_mod):
    _eq(oracle_mod.rk3_1d(gd["init_1d"].copy(), 64, 3, 1 / 64, 0.002, 1.4, True, (1, 5)), gd["rk3_1d_out"])


@pytest.mark.parametrize("key", ["loop_nstep", "loop_tfinal"])
def test_riemann_time_loop(gd, oracle_mod, key):
    nstep, t_final, cfl, cons, jlo, jhi = gd[f"{key}_ctrl"]
    from wenox import problems
    q0 = problems.riemann_reference(16, 16)
    q, ts = oracle_mod.run_riemann(q0, 16, 16, 1 / 16, 1.4, cfl, False, bool(cons), (jlo, jhi),
                                   nstep=int(nstep), t_final=None if t_final < 0 else t_final)
    assert ts == _times(gd[f"{key}_dt"])
    _eq(q, gd[f"{key}_q"])
    if t_final > 0:
        # the last step is clamped to the remaining time
        assert ts[-1] == pytest.approx(t_final, abs=1e-12)


def test_dmr_time_loop(gd, oracle_mod):
    from wenox import problems
    q0, bound = problems.dmr_reference(40, 10)
    q, ts = oracle_mod.run_dmr_reference(q0, 40, 10, 1 / 100, bound, 1.4, 0.1, 0.003, 400)
    assert ts == _times(gd["loop_dmr_dt"])
    _eq(q, gd["loop_dmr_q"])


def test_1d_time_loop(gd, oracle_mod):
    from wenox import problems
    q, T = oracle_mod.run_1d(problems.sod_reference(64, 3), 64, 3, 1 / 64, 1.4, 0.1, 7)
    _eq(q, gd["loop_1d_q"])
    assert T == _times(gd["loop_1d_dt"])[-1]
