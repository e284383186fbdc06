"""The oracle's twin mode against the NumPy twin's own outputs.

``tests/golden/twin_mode_golden.npz`` holds what the unpatched
``WENO_HLLC_2d.py`` returns (``tests/golden/make_twin_mode_golden.py``):
component-wise WENO5 + Gauss points with the exact sqrt(3)/12, its HLLC and
HLLE, the blend, L, and two short runs of its time loop.  Inside
``oracle.twin()`` the C oracle must reproduce every array bit for bit; outside
it the Rust semantics must differ from them where the two codes differ (the
checks below would not notice a twin mode that changed nothing).
"""
import os

import numpy as np
import pytest

G = 1.4
JP = (1.0, 4.0)
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ("rand16x12", "quad24x18", "super20x16")


@pytest.fixture(scope="module")
def tg():
    return np.load(os.path.join(HERE, "golden", "twin_mode_golden.npz"))


def faces(tg, name, d):
    return [tg[f"{name}/weno_{d}/{k}"] for k in ("q_l", "f_l", "q_r", "f_r")]


def test_gauss_constant(tg):
    # the literal the device code carries for the twin's np.sqrt(3)/12
    assert float(tg["sqrt3_12"]) == np.sqrt(3.0) / 12.0 == 0.14433756729740643


@pytest.mark.parametrize("name", CASES)
def test_reconstruction(oracle_mod, tg, name):
    u = tg[f"{name}/u"]
    with oracle_mod.twin():
        for d, fn in (("x", oracle_mod.weno_x_py), ("y", oracle_mod.weno_y_py)):
            for got, want in zip(fn(u, G, True), faces(tg, name, d)):
                np.testing.assert_array_equal(got, want, err_msg=f"{name} weno_{d}")
    # the Rust constant changes the Gauss values
    assert not np.array_equal(oracle_mod.weno_x_py(u, G, True)[0], tg[f"{name}/weno_x/q_l"])


@pytest.mark.parametrize("name", CASES)
def test_riemann_and_blend(oracle_mod, tg, name):
    with oracle_mod.twin():
        for d in ("x", "y"):
            ql, fl, qr, fr = faces(tg, name, d)
            hllc = oracle_mod.hllc_x_rs_py if d == "x" else oracle_mod.hllc_y_rs_py
            hlle = oracle_mod.hlle_x_rs_py if d == "x" else oracle_mod.hlle_y_rs_py
            blend = oracle_mod.hllc_hlle_x_rs_py if d == "x" else oracle_mod.hllc_hlle_y_rs_py
            np.testing.assert_array_equal(hllc(ql, fl, qr, fr, G), tg[f"{name}/hllc_{d}"])
            np.testing.assert_array_equal(hlle(ql, fl, qr, fr, G), tg[f"{name}/hlle_{d}"])
            np.testing.assert_array_equal(blend(ql, fl, qr, fr, G, *JP, False), tg[f"{name}/blend_{d}"])


@pytest.mark.parametrize("name", CASES)
def test_l(oracle_mod, tg, name):
    u, dx = tg[f"{name}/u"], float(tg[f"{name}/dx"])
    with oracle_mod.twin():
        np.testing.assert_array_equal(oracle_mod.l_local_py(u, dx, G, False, JP, True), tg[f"{name}/L_blend"])
        np.testing.assert_array_equal(oracle_mod.l_local_py(u, dx, G, True, JP, True), tg[f"{name}/L_hlle"])


def test_riemann_random_faces(oracle_mod, tg):
    """Faces drawn over wide ranges (supersonic both ways, star regions):
    the twin's HLLC right-star coefficient differs from the Rust one, so the
    Rust semantics must miss the golden values somewhere."""
    for d in ("x", "y"):
        args = [tg[f"riemann_rand/{d}/{k}"] for k in ("q_l", "f_l", "q_r", "f_r")]
        hllc = oracle_mod.hllc_x_rs_py if d == "x" else oracle_mod.hllc_y_rs_py
        hlle = oracle_mod.hlle_x_rs_py if d == "x" else oracle_mod.hlle_y_rs_py
        assert args[0].shape[1] > 300
        with oracle_mod.twin():
            np.testing.assert_array_equal(hllc(*args, G), tg[f"riemann_rand/{d}/hllc"])
            np.testing.assert_array_equal(hlle(*args, G), tg[f"riemann_rand/{d}/hlle"])
        assert not np.array_equal(hllc(*args, G), tg[f"riemann_rand/{d}/hllc"])


def test_hlle_edge_rows(oracle_mod, tg):
    """Near-vacuum faces (raw states in the middle state) and S_L = S_R = 0
    (both masks: the twin keeps F_R) separate the twin's HLLE from the Rust
    one."""
    for d in ("x", "y"):
        args = [tg[f"riemann_edge/{d}/{k}"] for k in ("q_l", "f_l", "q_r", "f_r")]
        hlle = oracle_mod.hlle_x_rs_py if d == "x" else oracle_mod.hlle_y_rs_py
        want = tg[f"riemann_edge/{d}/hlle"]
        with oracle_mod.twin():
            np.testing.assert_array_equal(hlle(*args, G), want)
        rust = hlle(*args, G)
        assert not np.array_equal(rust[0, 1::4], want[0, 1::4])   # near-vacuum rows
        assert not np.array_equal(rust[0, 2::4], want[0, 2::4])   # S_L = S_R = 0 rows


def test_twin_initial_data(oracle_mod, tg):
    for name in ("loop24", "loop20_hlle"):
        n = int(tg[f"{name}/meta"][0])
        np.testing.assert_array_equal(oracle_mod.twin_quadrants(n), tg[f"{name}/u0"])


@pytest.mark.parametrize("name", ["loop24", "loop20_hlle"])
def test_twin_time_loop(oracle_mod, tg, name):
    n, n_step, force = (int(x) for x in tg[f"{name}/meta"])
    u, ts = oracle_mod.run_twin(tg[f"{name}/u0"].copy(), n, n_step, bool(force), JP)
    assert len(ts) == n_step
    np.testing.assert_array_equal(np.array(ts), tg[f"{name}/t"])
    np.testing.assert_array_equal(u, tg[f"{name}/u"])
    assert np.all(u[..., 7] >= 1e-8)


def test_twin_pressure_floor(oracle_mod, tg):
    """con_var2primi_pfloor on cells around and below pmin (some with density
    under the 1e-12 floor): energy raised where p < pmin, primitives written."""
    u = oracle_mod.twin_pfloor(tg["pfloor/in"].copy(), G)
    np.testing.assert_array_equal(u, tg["pfloor/out"])
    raised = tg["pfloor/out"][..., 3] != tg["pfloor/in"][..., 3]
    assert 20 <= raised.sum() < raised.size
    np.testing.assert_array_equal(u[..., 7][raised], 1e-8)


def test_product_initial_data_matches_twin(tg):
    """wenox.problems.twin_quadrants (the solver's initial data for the twin
    driver) equals the twin's own, ghost cells included."""
    from wenox import problems
    for name in ("loop24", "loop20_hlle"):
        n = int(tg[f"{name}/meta"][0])
        np.testing.assert_array_equal(problems.twin_quadrants(n), tg[f"{name}/u0"])
