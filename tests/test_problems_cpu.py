"""CPU checks of the host-side initial data, closed-form solutions and driver helpers."""
import numpy as np
import pytest

from wenox import problems


def test_sod_exact_star_state():
    # Toro (2009) table 4.3, test 1: p* = 0.30313, u* = 0.92745, rho*L = 0.42632, rho*R = 0.26557
    x = np.array([0.55, 0.75])
    rho, u, p = problems.sod_exact(x, 0.2, x0=0.5)
    np.testing.assert_allclose(p, [0.30313, 0.30313], rtol=1e-4)
    np.testing.assert_allclose(u, [0.92745, 0.92745], rtol=1e-4)
    np.testing.assert_allclose(rho, [0.42632, 0.26557], rtol=1e-4)


def test_four_quadrant_states():
    q, dx = problems.four_quadrant(16)
    assert dx == 1 / 16
    st = {(1, 1): (1.5, 0, 0, 1.5), (0, 1): (0.5323, 1.206, 0, 0.3),
          (0, 0): (0.138, 1.206, 1.206, 0.029), (1, 0): (0.5323, 0, 1.206, 0.3)}
    for (hx, hy), (r, u, v, p) in st.items():
        c = q[3 + 8 * hx + 3, 3 + 8 * hy + 3]
        np.testing.assert_allclose(c[4:], (r, u, v, p))
        np.testing.assert_allclose(c[:4], (r, r * u, r * v, p / 0.4 + 0.5 * r * (u * u + v * v)))


def test_vortex_cell_average_is_exact_at_t0():
    q, dx = problems.isentropic_vortex(32)
    avg = problems.vortex_cell_average(32, 0.0)
    np.testing.assert_array_equal(q[3:-3, 3:-3, :4], avg)
    # far field is the free stream (1, 1, 1, 1); the vortex centre is cooler
    assert abs(q[3, 3, 0] - 1.0) < 1e-6 and q[3 + 16, 3 + 16, 0] < 0.7
    # advected by (1, 1) over one period of the box the exact solution returns
    np.testing.assert_allclose(problems.vortex_cell_average(16, 10.0), problems.vortex_cell_average(16, 0.0),
                               atol=1e-12)


def test_kelvin_helmholtz_seeded():
    a, _ = problems.kelvin_helmholtz(64)
    b, _ = problems.kelvin_helmholtz(64)
    np.testing.assert_array_equal(a, b)
    rho = a[3:-3, 3:-3, 4]
    assert set(np.unique(rho)) == {1.0, 2.0}
    assert np.abs(a[3:-3, 3:-3, 6]).max() <= 0.01


def test_double_mach_geometry():
    q, dx = problems.double_mach(192, 48)
    assert dx == 4 / 192
    x = (np.arange(192) + 0.5) * dx
    # on the first interior row (y = dx/2) the shock sits at x = 1/6 + y/sqrt(3)
    post = q[3:-3, 3, 0] == 8.0
    xs = 1 / 6 + 0.5 * dx / np.sqrt(3)
    assert np.all(post == (x < xs))


def test_dmr_reference_bound():
    q, bound = problems.dmr_reference(400, 100)
    assert bound == 19     # the reference configuration (nx=400, dx=1/100)


def _reference_dmr_finite(oracle_mod, cfl, steps, nx=100, ny=25):
    q, bound = problems.dmr_reference(nx, ny)
    dx, T = 1.0 / ny, 0.0
    for _ in range(steps):
        dt = oracle_mod.cfl_dt(q, nx, ny, 1.4, cfl, dx, remaining=0.2 - T)
        q = oracle_mod.rk3_refac(q, nx, ny, dx, dt, 1.4, False, bound, (10000, 11000))
        T += dt
        if not np.isfinite(q[3:-3, 3:-3, :4]).all():
            return False
    return True


def test_reference_dmr_needs_its_own_cfl(oracle_mod):
    """The reference scheme (characteristic WENO5 + HLLC, its DMR ghost fill)
    stays finite at its driver's CFL 0.1 (double_mach_reflection.py:17) and
    turns non-finite within 70 steps at CFL 0.4, so double Mach reflection runs
    are made at CFL 0.1."""
    assert _reference_dmr_finite(oracle_mod, 0.1, 70)
    assert not _reference_dmr_finite(oracle_mod, 0.4, 70)


def test_snapshot_writer(tmp_path):
    """The drivers' background writer saves every frame it is given and
    reports a failed write to the stepping thread."""
    from wenox.drivers import _SnapshotWriter
    w = _SnapshotWriter(depth=2)
    for i in range(6):
        w.put(str(tmp_path / f"{i}.npy"), np.full((4, 3, 8), float(i)))
    w.close()
    for i in range(6):
        assert np.load(tmp_path / f"{i}.npy")[0, 0, 0] == float(i)
    w = _SnapshotWriter()
    w.put(str(tmp_path / "missing_dir" / "0.npy"), np.zeros(3))
    with pytest.raises(OSError):
        w.close()


def test_double_mach_needs_a_4_to_1_grid():
    """The solver places the exact shock trace on the top boundary for
    [0,4] x [0,1] with one cell size, so the standard double Mach reflection
    rejects other aspect ratios."""
    import pytest
    from wenox import problems
    with pytest.raises(ValueError):
        problems.double_mach(40, 34)
    q, dx = problems.double_mach(136, 34)
    assert q.shape == (142, 40, 8) and dx == 4.0 / 136
