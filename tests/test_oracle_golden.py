"""Pin the oracle against golden vectors of the reference's NumPy twin.

The vectors come from ``tests/golden/make_golden.py`` (reference
``WENO_HLLC_2d.py`` with its SQRT3_12 set to the Rust constant).  On these
inputs the twin and the Rust crate perform the same IEEE operations in the same
order, so the comparison is bitwise except where noted.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "twin_golden.npz")
CASES = ["rand16x12", "rand20x20", "quad24x18", "smooth26x22"]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def _eq(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("case", CASES)
def test_flux_and_con2primi(gold, oracle_mod, case):
    u = gold[f"{case}/u"]
    _eq(oracle_mod.flux_x_py(u, 1.4), gold[f"{case}/flux_x"])
    _eq(oracle_mod.flux_y_py(u, 1.4), gold[f"{case}/flux_y"])
    p, a, rho, uu, vv, h = oracle_mod.con2primi_py(u, 1.4)
    for k, v in zip(("p", "a", "rho", "u", "v"), (p, a, rho, uu, vv)):
        _eq(v, gold[f"{case}/con2primi/{k}"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("d", ["x", "y"])
def test_componentwise_weno(gold, oracle_mod, case, d):
    u = gold[f"{case}/u"]
    fn = oracle_mod.weno_x_py if d == "x" else oracle_mod.weno_y_py
    got = fn(u, 1.4, cons=True)
    for k, v in zip(("q_l", "f_l", "q_r", "f_r"), got):
        _eq(v, gold[f"{case}/weno_{d}/{k}"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("d", ["x", "y"])
def test_hlle(gold, oracle_mod, case, d):
    ql, fl, qr, fr = (gold[f"{case}/weno_{d}/{k}"] for k in ("q_l", "f_l", "q_r", "f_r"))
    fn = oracle_mod.hlle_x_rs_py if d == "x" else oracle_mod.hlle_y_rs_py
    _eq(fn(ql, fl, qr, fr, 1.4), gold[f"{case}/hlle_{d}"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("d", ["x", "y"])
def test_sensor(gold, oracle_mod, case, d):
    pl = gold[f"{case}/sensor_{d}/p_l"]
    pr = gold[f"{case}/sensor_{d}/p_r"]
    _eq(oracle_mod.shock_sensor_grad_p(0.5 * (pl + pr), 0.05, 0.6), gold[f"{case}/sensor_{d}/phi"])


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("d", ["x", "y"])
def test_blend_hlle_branch(gold, oracle_mod, case, d):
    """phi == 0 everywhere: blend = 0*HLLC + 1*HLLE.  The twin's HLLC takes a
    sqrt of a negative number on some random faces (no max(.,0) at
    WENO_HLLC_2d.py:453,528) and returns NaN there; those faces are excluded."""
    ql, fl, qr, fr = (gold[f"{case}/weno_{d}/{k}"] for k in ("q_l", "f_l", "q_r", "f_r"))
    fn = oracle_mod.hllc_hlle_x_rs_py if d == "x" else oracle_mod.hllc_hlle_y_rs_py
    got = fn(ql, fl, qr, fr, 1.4, -2.0, -1.0, False)
    ref = gold[f"{case}/blend0_{d}"]
    ok = np.isfinite(ref)
    assert ok.mean() > 0.5
    np.testing.assert_array_equal(got[ok], ref[ok])
    # where the twin is finite the blend must equal pure HLLE
    hl = gold[f"{case}/hlle_{d}"]
    np.testing.assert_array_equal(got[ok], hl[ok])


@pytest.mark.parametrize("case", CASES)
def test_l_local_hlle_componentwise(gold, oracle_mod, case):
    u = gold[f"{case}/u"]
    dx = float(gold[f"{case}/L_hlle/dx"])
    got = oracle_mod.l_local_py(u, dx, 1.4, True, (0.5, 2.0), True)
    _eq(got, gold[f"{case}/L_hlle"])
