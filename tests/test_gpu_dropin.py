"""GPU parity of the drop-in functions (weno_ext surface) against the oracle.

The "exact" kernel variant follows the reference's operation order with no
FMA contraction and correctly rounded division/sqrt, so it must equal the
oracle bit for bit.  The "fast" variant (FMA, shared reciprocals) must agree to
a relative max-norm of 1e-12 per call.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "twin_golden.npz")
CASES = ["rand16x12", "rand20x20", "quad24x18", "smooth26x22"]
FAST_RTOL = 1e-12


@pytest.fixture(scope="module")
def wx():
    import weno_ext
    weno_ext.set_variant(False)
    yield weno_ext
    weno_ext.set_variant(False)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def _inputs(gold):
    rng = np.random.default_rng(7)
    out = {c: gold[f"{c}/u"] for c in CASES}
    from wenox import problems
    out["fq40"] = problems.four_quadrant(40)[0]
    out["ref32"] = problems.riemann_reference(32, 32)
    # non-contiguous view with 8 channels (the drivers pass (m, n, 8) arrays)
    big = rng.uniform(0.5, 1.5, (30, 28, 8))
    big[..., 3] += 3.0
    out["strided"] = big[:, ::-1, :][::1, :, :]
    return out


def _close(a, b, rtol):
    scale = max(np.abs(b).max(), 1e-300)
    assert a.shape == b.shape
    assert np.all(np.isfinite(a) == np.isfinite(b))
    ok = np.isfinite(b)
    err = np.abs(a[ok] - b[ok]).max() / scale if ok.any() else 0.0
    assert err <= rtol, err


@pytest.mark.parametrize("fast", [False, True])
def test_flux_con2primi(wx, oracle_mod, gold, fast):
    wx.set_variant(fast)
    for name, u in _inputs(gold).items():
        for fn in ("flux_x_py", "flux_y_py"):
            got, ref = getattr(wx, fn)(u, 1.4), getattr(oracle_mod, fn)(u, 1.4)
            if fast:
                _close(got, ref, FAST_RTOL)
            else:
                np.testing.assert_array_equal(got, ref, err_msg=f"{fn} {name}")
        for g, r in zip(wx.con2primi_py(u, 1.4), oracle_mod.con2primi_py(u, 1.4)):
            if fast:
                _close(g, r, FAST_RTOL)
            else:
                np.testing.assert_array_equal(g, r, err_msg=f"con2primi {name}")


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("cons", [False, True])
@pytest.mark.parametrize("d", ["x", "y"])
def test_weno(wx, oracle_mod, gold, fast, cons, d):
    wx.set_variant(fast)
    for name, u in _inputs(gold).items():
        got = getattr(wx, f"weno_{d}_py")(u, 1.4, cons)
        ref = getattr(oracle_mod, f"weno_{d}_py")(u, 1.4, cons)
        for g, r, k in zip(got, ref, ("q_l", "f_l", "q_r", "f_r")):
            if fast:
                _close(g, r, 1e-10 if not cons else FAST_RTOL)
            else:
                np.testing.assert_array_equal(g, r, err_msg=f"weno_{d} {k} {name} cons={cons}")


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("d", ["x", "y"])
def test_riemann(wx, oracle_mod, gold, fast, d):
    wx.set_variant(fast)
    for name, u in _inputs(gold).items():
        faces = getattr(oracle_mod, f"weno_{d}_py")(u, 1.4, False)
        if min(faces[0].shape[:2]) < 2:
            continue
        pairs = [
            (getattr(wx, f"hllc_{d}_rs_py")(*faces, 1.4), getattr(oracle_mod, f"hllc_{d}_rs_py")(*faces, 1.4)),
            (getattr(wx, f"hlle_{d}_rs_py")(*faces, 1.4), getattr(oracle_mod, f"hlle_{d}_rs_py")(*faces, 1.4)),
        ]
        for lo, hi in ((0.05, 0.6), (0.5, 2.0), (1e4, 1.1e4)):
            for fh in (False, True):
                pairs.append((getattr(wx, f"hllc_hlle_{d}_rs_py")(*faces, 1.4, lo, hi, fh),
                              getattr(oracle_mod, f"hllc_hlle_{d}_rs_py")(*faces, 1.4, lo, hi, fh)))
        for g, r in pairs:
            if fast:
                _close(g, r, 1e-10)
            else:
                np.testing.assert_array_equal(g, r, err_msg=f"riemann {d} {name}")


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("cons", [False, True])
@pytest.mark.parametrize("force_hlle", [False, True])
def test_l_local(wx, oracle_mod, gold, fast, cons, force_hlle):
    wx.set_variant(fast)
    for name, u in _inputs(gold).items():
        dx = 1.0 / (u.shape[0] - 6)
        got = wx.l_local_py(u, dx, 1.4, force_hlle, (0.5, 2.0), cons)
        ref = oracle_mod.l_local_py(u, dx, 1.4, force_hlle, (0.5, 2.0), cons)
        if fast:
            _close(got, ref, 1e-9)
        else:
            np.testing.assert_array_equal(got, ref, err_msg=f"l_local {name} cons={cons} hlle={force_hlle}")


def test_l_local_five_args(wx, oracle_mod):
    """double_mach_reflection.py calls l_local_py with five arguments; the
    default cons_switch=False selects the characteristic path."""
    from wenox import problems
    wx.set_variant(False)
    q, bound = problems.dmr_reference(40, 10)
    got = wx.l_local_py(q, 1 / 10, 1.4, False, (10000, 11000))
    ref = oracle_mod.l_local_py(q, 1 / 10, 1.4, False, (10000, 11000), False)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("shape,cons,force_hlle", [((1920, 480), False, False), ((1920, 480), True, True),
                                                    ((257, 1100), False, False), ((130, 2050), True, False)])
def test_l_local_large_grid_slabs(wx, oracle_mod, shape, cons, force_hlle):
    """Grids of a million output values and more run in row slabs (packing,
    the copies each way and the kernels overlapped; each slab's passes cover
    one row more on each inner side, the sensor's h that of the whole grid):
    exact variant bitwise against the oracle on an (m, n, 8) driver array,
    at the double Mach reflection's size and on uneven slab splits; fast within
    1e-9 of the largest value, as on small grids."""
    from wenox import problems
    nx, ny = shape
    if shape == (1920, 480):
        q = problems.double_mach(nx, ny)[0]
    else:
        rng = np.random.default_rng(nx + ny)
        q = problems.random_riemann_instance(rng, nx, ny)[0]
    assert q.shape == (nx + 6, ny + 6, 8)
    print(f"[dropin] l_local {shape} cons={cons} force_hlle={force_hlle}", flush=True)
    ref = oracle_mod.l_local_py(q, 1.0 / ny, 1.4, force_hlle, (0.5, 2.0), cons)
    wx.set_variant(False)
    got = wx.l_local_py(q, 1.0 / ny, 1.4, force_hlle, (0.5, 2.0), cons)
    np.testing.assert_array_equal(got, ref)
    wx.set_variant(True)
    try:
        fast = wx.l_local_py(q, 1.0 / ny, 1.4, force_hlle, (0.5, 2.0), cons)
    finally:
        wx.set_variant(False)
    _close(fast, ref, 1e-9)


def test_sqrt_div_correctly_rounded(wx):
    """con2primi's a = sqrt(g p / rho) and u = m/rho on random data exercise the
    device division and square root; they must round like the host's."""
    wx.set_variant(False)
    rng = np.random.default_rng(11)
    q = np.empty((257, 129, 4))
    q[..., 0] = np.exp(rng.uniform(-20, 20, q.shape[:2]))
    q[..., 1] = rng.standard_normal(q.shape[:2]) * np.exp(rng.uniform(-20, 20, q.shape[:2]))
    q[..., 2] = rng.standard_normal(q.shape[:2])
    q[..., 3] = np.exp(rng.uniform(-20, 30, q.shape[:2]))
    p, a, rho, u, v, h = wx.con2primi_py(q, 1.4)
    rs = np.maximum(q[..., 0], 1e-12)
    np.testing.assert_array_equal(u, q[..., 1] / rs)
    pp = np.maximum((1.4 - 1.0) * (q[..., 3] - 0.5 * rs * (u * u + v * v)), 1e-12)
    np.testing.assert_array_equal(p, pp)
    np.testing.assert_array_equal(a, np.sqrt(1.4 * pp / rs))
