"""GPU parity of the twin mode: the semantics of the pure-NumPy
WENO_HLLC_2d.py (exact sqrt(3)/12, its HLLC and HLLE, its time loop with
constant-extrapolation ghosts, CFL 0.2 and the pressure floor).

The twin build is compiled without FMA contraction and follows the twin's
operation order, so every comparison is bitwise: against the twin's own
outputs (``tests/golden/twin_mode_golden.npz``) and against the oracle's twin
mode (pinned to the same file on the CPU) on larger inputs.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 1.4
JP = (1.0, 4.0)
GOLD = os.path.join(os.path.dirname(__file__), "golden", "twin_mode_golden.npz")
CASES = ("rand16x12", "quad24x18", "super20x16")


@pytest.fixture(scope="module")
def tg():
    return np.load(GOLD, allow_pickle=False)


@pytest.fixture
def wx_twin():
    import weno_ext
    weno_ext.set_variant("twin")
    yield weno_ext
    weno_ext.set_variant("exact")


def _faces(tg, prefix):
    return [tg[f"{prefix}/{k}"] for k in ("q_l", "f_l", "q_r", "f_r")]


@pytest.mark.parametrize("name", CASES)
def test_dropin_matches_twin(wx_twin, tg, name):
    wx = wx_twin
    u = tg[f"{name}/u"]
    print(f"[twin] drop-in {name}", flush=True)
    for d in ("x", "y"):
        got = (wx.weno_x_py if d == "x" else wx.weno_y_py)(u, G, True)
        for a, b in zip(got, _faces(tg, f"{name}/weno_{d}")):
            np.testing.assert_array_equal(a, b, err_msg=f"weno_{d}")
        ql, fl, qr, fr = _faces(tg, f"{name}/weno_{d}")
        hllc = wx.hllc_x_rs_py if d == "x" else wx.hllc_y_rs_py
        hlle = wx.hlle_x_rs_py if d == "x" else wx.hlle_y_rs_py
        blend = wx.hllc_hlle_x_rs_py if d == "x" else wx.hllc_hlle_y_rs_py
        np.testing.assert_array_equal(hllc(ql, fl, qr, fr, G), tg[f"{name}/hllc_{d}"], err_msg=f"hllc_{d}")
        np.testing.assert_array_equal(hlle(ql, fl, qr, fr, G), tg[f"{name}/hlle_{d}"], err_msg=f"hlle_{d}")
        np.testing.assert_array_equal(blend(ql, fl, qr, fr, G, *JP, False), tg[f"{name}/blend_{d}"],
                                      err_msg=f"blend_{d}")
    dx = float(tg[f"{name}/dx"])
    np.testing.assert_array_equal(wx.l_local_py(u, dx, G, False, JP, True), tg[f"{name}/L_blend"])
    np.testing.assert_array_equal(wx.l_local_py(u, dx, G, True, JP, True), tg[f"{name}/L_hlle"])


def test_dropin_riemann_faces(wx_twin, tg):
    """Supersonic and star-region faces (HLLC, HLLE) and the HLLE edge rows:
    near-vacuum faces and S_L = S_R = 0, where the twin keeps F_R."""
    wx = wx_twin
    print("[twin] drop-in Riemann faces", flush=True)
    for d in ("x", "y"):
        hllc = wx.hllc_x_rs_py if d == "x" else wx.hllc_y_rs_py
        hlle = wx.hlle_x_rs_py if d == "x" else wx.hlle_y_rs_py
        args = _faces(tg, f"riemann_rand/{d}")
        np.testing.assert_array_equal(hllc(*args, G), tg[f"riemann_rand/{d}/hllc"])
        np.testing.assert_array_equal(hlle(*args, G), tg[f"riemann_rand/{d}/hlle"])
        np.testing.assert_array_equal(hlle(*_faces(tg, f"riemann_edge/{d}"), G), tg[f"riemann_edge/{d}/hlle"])


def test_dropin_larger_grid_matches_oracle(wx_twin, oracle_mod):
    from wenox import problems
    u = problems.twin_quadrants(48)
    print("[twin] drop-in l_local 48^2 vs oracle", flush=True)
    with oracle_mod.twin():
        for force in (False, True):
            ref = oracle_mod.l_local_py(u, 1 / 48, G, force, JP, True)
            np.testing.assert_array_equal(wx_twin.l_local_py(u, 1 / 48, G, force, JP, True), ref)


def _solver_run(u0, n, n_step, force):
    from wenox import Solver
    with Solver(n, n, 1 / n, gamma=G, cfl=0.2, cons_switch=True, force_hlle=force, jp_cri=JP, bc="copy",
                engine="unfused", variant="twin") as s:
        s.set_cfl_rule(1e-10, 0.0)
        s.set_pfloor(1e-8)
        s.upload(u0)
        T, k = s.run(nstep=n_step - 1)
        assert k == n_step - 1
        return s.download(), T


@pytest.mark.parametrize("name", ["loop24", "loop20_hlle"])
def test_solver_matches_twin_loop(tg, name):
    n, n_step, force = (int(x) for x in tg[f"{name}/meta"])
    print(f"[twin] solver {name}: {n}^2, {n_step - 1} steps, force_hlle={bool(force)}", flush=True)
    got, T = _solver_run(tg[f"{name}/u0"], n, n_step, bool(force))
    want = tg[f"{name}/u"]
    assert T == tg[f"{name}/t"][-1]
    np.testing.assert_array_equal(got[..., :4], want[..., :4])
    # primitives: the floor did not act on this problem, so the download's
    # (rho_safe, u, v, p) equal the twin's channels
    np.testing.assert_array_equal(got[..., 4:], want[..., 4:])


def test_main_twin_matches_oracle_loop(oracle_mod):
    """wenox.drivers.main_twin (no snapshots) against the oracle's twin loop
    on the twin's own initial data, 64^2, 40 steps."""
    from wenox import drivers, problems
    n, n_step = 64, 41
    print(f"[twin] main_twin {n}^2 {n_step - 1} steps", flush=True)
    ref, ts = oracle_mod.run_twin(problems.twin_quadrants(n), n, n_step, False, (1.0, 4.0))
    got = drivers.main_twin(n, n_step, fileStorage=False)
    np.testing.assert_array_equal(got[..., :4], ref[..., :4])


def _low_pressure(n):
    """A density bump advected at speed 1e-4 (Mach about 1) in pressure
    1e-8 (1 + 0.9 sin 2 pi x cos 2 pi y): the floor raises the energy of
    hundreds of cells per step."""
    from wenox import problems
    x = (np.arange(n + 6) - 3 + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    q = np.zeros((n + 6, n + 6, 8))
    q[..., 4] = 1 + 0.2 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    q[..., 5] = q[..., 6] = 1e-4
    q[..., 7] = 1e-8 * (1 + 0.9 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y))
    problems._from_prims(q, G)
    return problems._copy_fill(q)


def test_pressure_floor_loop_matches_oracle(oracle_mod):
    """The floor acts on many cells each step; the next step's dt comes from
    the floored primitives (p = pmin exactly there), as in the twin."""
    n, n_step = 32, 16
    q0 = _low_pressure(n)
    print(f"[twin] pressure-floor loop {n}^2 {n_step - 1} steps", flush=True)
    ref, ts = oracle_mod.run_twin(q0.copy(), n, n_step, False, JP)
    assert (ref[..., 7] == 1e-8).sum() > 50
    got, T = _solver_run(q0, n, n_step, False)
    assert T == ts[-1]
    np.testing.assert_array_equal(got[..., :4], ref[..., :4])


def test_twin_solver_arguments_checked():
    from wenox import Solver
    from wenox._lib import WenoxError
    for kw in (dict(engine="fused"), dict(cons_switch=False), dict(precision="fp32split", engine="fused")):
        args = dict(cons_switch=True, engine="unfused", variant="twin")
        args.update(kw)
        with pytest.raises((ValueError, WenoxError)):
            Solver(16, 16, 1 / 16, **args)


def test_main_twin_snapshots(oracle_mod, tmp_path, monkeypatch):
    """main_twin with fileStorage writes the twin's files: data/<stamp>/0.npy
    (the initial state), {i}.npy for i in 1 .. N_STEP-1 and time.npy, the
    conserved channels and times equal to the oracle's twin loop."""
    import glob
    from wenox import drivers, problems
    monkeypatch.chdir(tmp_path)
    n, n_step = 24, 6
    print(f"[twin] main_twin snapshots {n}^2 {n_step - 1} steps", flush=True)
    q = drivers.main_twin(n, n_step, fileStorage=True)
    ref, ts = oracle_mod.run_twin(problems.twin_quadrants(n), n, n_step, False, (1.0, 4.0))
    folders = glob.glob(os.path.join("data", "*"))
    assert len(folders) == 1
    files = set(os.listdir(folders[0]))
    assert files == {f"{i}.npy" for i in range(n_step)} | {"time.npy"}
    np.testing.assert_array_equal(np.load(os.path.join(folders[0], "time.npy")), ts)
    np.testing.assert_array_equal(np.load(os.path.join(folders[0], "0.npy")), problems.twin_quadrants(n))
    last = np.load(os.path.join(folders[0], f"{n_step - 1}.npy"))
    np.testing.assert_array_equal(last, q)
    np.testing.assert_array_equal(last[..., :4], ref[..., :4])
