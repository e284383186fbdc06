"""Pin the oracle with what the mathematics fixes (CPU only).

Each test is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, index or weight, a transposed operand) fails it:

* WENO5-JS: smoothness indicators equal their definition (integral of the
  squared derivatives of each candidate polynomial, computed here by numpy
  polynomial fits); linear weights reproduce the 5-point stencil; quadratic
  data is reconstructed exactly whatever the nonlinear weights.
* Gauss-quadrature values: exact for quadratic data at x = +-h/(2 sqrt 3) up
  to the truncated constants of the reference.
* Characteristic basis: A R = R diag(u-c, u, u, u+c) with A the Euler flux
  Jacobian (central differences); L = R^-1 except for the reference's own
  1/(2 rho c) entries in rows 0 and 3 (pinned exactly); Roe's property
  A(Roe) (U_R - U_L) = F(U_R) - F(U_L).
* Characteristic reconstruction of fields quadratic along the normal equals
  R L q_exact (limiter idle); x/y equivariance under transposition with
  u <-> v.
* HLLC equals Toro's pressure-form HLLC flux (an independent formulation);
  supersonic states give the upwind physical flux; HLLE middle state equals
  the HLL formula.
* Positivity limiter: admissible states untouched; limited states have
  density and pressure above the floors.
* Whole RHS: free stream gives 0; y-uniform data gives no y-flux difference.
* Drivers: one rk3_cc step on dU/dt = z U has amplification 1 + z + z^2/2 + z^3/6.
* Threads: the OpenMP split of the oracle's loops leaves the RHS bitwise unchanged.
"""
import numpy as np
import pytest

G = 1.4


def cons(rho, u, v, p, g=G):
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


def euler_flux(U, d=0, g=G):
    rho, m, n, E = U
    u, v = m / rho, n / rho
    p = (g - 1) * (E - 0.5 * rho * (u * u + v * v))
    if d == 0:
        return np.array([m, m * u + p, n * u, u * (E + p)])
    return np.array([n, m * v, n * v + p, v * (E + p)])


def jacobian(U, d=0):
    J = np.zeros((4, 4))
    for k in range(4):
        h = 1e-6 * max(1.0, abs(U[k]))
        e = np.zeros(4)
        e[k] = h
        J[:, k] = (euler_flux(U + e, d) - euler_flux(U - e, d)) / (2 * h)
    return J


# ---------------------------------------------------------------- WENO5 ----
def _beta_by_definition(vals, cell):
    """Jiang-Shu indicator of the quadratic with cell averages vals on cells
    [-1,0],[0,1],[1,2] (h = 1), taken over cell number `cell` of the three:
    sum_{l=1,2} int (d^l p/dx^l)^2 dx."""
    # primitive P with P(k) - P(k-1) = vals; P cubic through 4 points
    xs = np.array([-1.0, 0.0, 1.0, 2.0])
    P = np.concatenate([[0.0], np.cumsum(vals)])
    c = np.polyfit(xs, P, 3)
    p = np.polyder(c)
    b = 0.0
    lo, hi = cell - 1.0, float(cell)
    for l in (1, 2):
        dp = np.polyder(p, l)
        sq = np.polymul(dp, dp)
        I = np.polyint(sq)
        b += np.polyval(I, hi) - np.polyval(I, lo)
    return b


def test_weno5_indicators_and_weights(oracle_mod):
    """Weights recomputed from the indicators' definition (integrals of the
    candidate polynomials' derivatives over the centre cell u2) and the
    linear weights (0.3, 0.6, 0.1) reproduce the oracle's value; the
    right-biased value is the mirror image of the left-biased one."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        u = rng.standard_normal(6) * rng.uniform(0.01, 10)
        left, right = oracle_mod.weno5_pair(u)
        # candidate stencils {u2,u3,u4}, {u1,u2,u3}, {u0,u1,u2}; u2 is cell 0, 1, 2 of each
        b = [_beta_by_definition(u[[2, 3, 4]], 0), _beta_by_definition(u[[1, 2, 3]], 1),
             _beta_by_definition(u[[0, 1, 2]], 2)]
        d = [0.3, 0.6, 0.1]
        a = [dk / (1e-12 + bk) ** 2 for dk, bk in zip(d, b)]
        w = np.array(a) / sum(a)
        cand = [(-u[4] + 5 * u[3] + 2 * u[2]) / 6, (-u[1] + 5 * u[2] + 2 * u[3]) / 6,
                (2 * u[0] - 7 * u[1] + 11 * u[2]) / 6]
        ref = float(np.dot(w, cand))
        assert abs(left - ref) <= 1e-9 * (1 + abs(ref)), (left, ref)
        mirrored, _ = oracle_mod.weno5_pair(np.concatenate([u[5:0:-1], [0.0]]))
        assert abs(right - mirrored) <= 1e-13 * (1 + abs(mirrored))


def test_weno5_linear_weights_reproduce_5pt_stencil(oracle_mod):
    # smooth data: nonlinear weights -> linear ones; the combination is the
    # 5th-order interface value (2 u-2 - 13 u-1 + 47 u0 + 27 u1 - 3 u2) / 60
    h = 1e-2
    x = (np.arange(6) - 2.5) * h        # cell centres; face at x = 0 between cells 2 and 3
    f = lambda t: np.exp(0.3 + 0.7 * t)
    avg = np.array([(np.exp(0.3 + 0.7 * (xi + h / 2)) - np.exp(0.3 + 0.7 * (xi - h / 2))) / (0.7 * h) for xi in x])
    left, right = oracle_mod.weno5_pair(avg)
    five = (2 * avg[0] - 13 * avg[1] + 47 * avg[2] + 27 * avg[3] - 3 * avg[4]) / 60
    assert abs(left - f(0.0)) < 1e-10
    assert abs(left - five) < 1e-11
    five_r = (2 * avg[5] - 13 * avg[4] + 47 * avg[3] + 27 * avg[2] - 3 * avg[1]) / 60
    assert abs(right - five_r) < 1e-11


def test_weno5_quadratic_exact(oracle_mod):
    rng = np.random.default_rng(5)
    for _ in range(50):
        a, b, c = rng.standard_normal(3)
        xs = np.arange(6) - 2.5
        avg = np.array([a + b * x + c * (x * x + 1 / 12) for x in xs])
        left, right = oracle_mod.weno5_pair(avg)
        assert abs(left - a) < 1e-12 * (1 + abs(a) + abs(b) + abs(c))
        assert abs(right - a) < 1e-12 * (1 + abs(a) + abs(b) + abs(c))


def test_gauss_points_quadratic(oracle_mod):
    rng = np.random.default_rng(6)
    gpt = 1.0 / (2.0 * np.sqrt(3.0))
    for _ in range(50):
        a, b, c = rng.standard_normal(3)
        xs = np.arange(5) - 2.0
        avg = np.array([a + b * x + c * (x * x + 1 / 12) for x in xs])
        ga, gb = oracle_mod.weno5_gq(avg)
        ex_a = a - b * gpt + c * gpt * gpt
        ex_b = a + b * gpt + c * gpt * gpt
        tol = 1e-9 * (1 + abs(a) + abs(b) + abs(c))   # truncated SQRT3_12 and weights
        assert abs(ga - ex_a) < tol and abs(gb - ex_b) < tol


# ------------------------------------------------- characteristic basis ----
def _states(rng, n):
    for _ in range(n):
        yield (cons(rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.1, 5)),
               cons(rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.1, 5)))


def _deviation(d, rho, c):
    """The reference's left eigenvectors differ from the inverse of its right
    eigenvectors in one entry of rows 0 and 3: the momentum-normal column
    carries 1/(2 rho c) where the inverse has 1/(2 c) (weno.rs:88,103 for x,
    :149,164 for y).  R L = I holds only where the Roe density is 1."""
    D = np.zeros((4, 4))
    col = 1 if d == 0 else 2
    delta = 0.5 / (rho * c) - 0.5 / c
    D[0, col] = -delta
    D[3, col] = +delta
    return D


@pytest.mark.parametrize("d", [0, 1])
def test_eigenvectors(oracle_mod, d):
    rng = np.random.default_rng(8 + d)
    for ul, ur in _states(rng, 60):
        roe, Lm, Rm = oracle_mod.roe_eig(d, ul, ur, G)
        rho, u, v, H, c = roe
        # right eigenvectors: A R = R diag(lambda) with A the flux Jacobian at
        # the Roe state (E = rho H - p, p = rho c^2 / gamma)
        p = rho * c * c / G
        U = np.array([rho, rho * u, rho * v, rho * H - p])
        un = u if d == 0 else v
        lam = np.array([un - c, un, un, un + c])
        AR = jacobian(U, d) @ Rm
        np.testing.assert_allclose(AR, Rm * lam[None, :], rtol=1e-6, atol=1e-6 * np.abs(AR).max())
        # left eigenvectors: the inverse of R plus the reference's deviation
        np.testing.assert_allclose(Lm, np.linalg.inv(Rm) + _deviation(d, rho, c), rtol=1e-11, atol=1e-12)
        # rows 1 and 2 (entropy, shear) are exact left eigenvectors
        np.testing.assert_allclose((Lm @ Rm)[1:3], np.eye(4)[1:3], atol=1e-12)


@pytest.mark.parametrize("d", [0, 1])
def test_product_of_bases_closed_form(oracle_mod, d):
    """R L q = q - ((rho~ - 1)/rho~) q_n (0, 1, 0, u~) for x faces and
    (0, 0, 1, v~) for y faces (q_n the normal momentum): the deviation above
    times R collapses to one vector because columns 0 and 3 of R differ only
    in their normal-velocity and energy entries.  The FP32-split kernels use
    this form for the FP64 part of the characteristic reconstruction."""
    rng = np.random.default_rng(28 + d)
    for ul, ur in _states(rng, 60):
        roe, Lm, Rm = oracle_mod.roe_eig(d, ul, ur, G)
        rho, u, v = roe[0], roe[1], roe[2]
        q = rng.uniform(-3, 3, 4)
        e = np.array([0.0, 1.0, 0.0, u]) if d == 0 else np.array([0.0, 0.0, 1.0, v])
        np.testing.assert_allclose(Rm @ (Lm @ q), q - (rho - 1.0) / rho * q[1 + d] * e,
                                   rtol=1e-12, atol=1e-13 * np.abs(q).max())


@pytest.mark.parametrize("d", ["x", "y"])
def test_uniform_flow_faces_are_product_of_bases(oracle_mod, d):
    """On a uniform field the characteristic reconstruction returns R L q, not
    q (the WENO weights sum to one in every characteristic component), so the
    face states of a moving flow with rho != 1 are shifted by the closed form
    while the flux differences still vanish."""
    q = cons(1.7, 0.6, -0.4, 2.1)
    u = np.broadcast_to(q, (14, 13, 4)).copy()
    ql = getattr(oracle_mod, f"weno_{d}_py")(u, G, False)[0]
    n = 1 if d == "x" else 2
    un = q[n] / q[0]
    e = np.array([0.0, 1.0, 0.0, un]) if d == "x" else np.array([0.0, 0.0, 1.0, un])
    expect = q - (q[0] - 1.0) / q[0] * q[n] * e
    np.testing.assert_allclose(ql, np.broadcast_to(expect, ql.shape), rtol=1e-13)
    rhs = oracle_mod.l_local_py(u, 0.1, G, False, (0.2, 1.0), False)
    assert np.abs(rhs).max() <= 1e-12 * np.abs(q).max() / 0.1


@pytest.mark.parametrize("d", [0, 1])
def test_roe_property(oracle_mod, d):
    """A(Roe)(U_R - U_L) = F(U_R) - F(U_L) with A = R diag(lambda) R^-1."""
    rng = np.random.default_rng(18 + d)
    for ul, ur in _states(rng, 60):
        roe, Lm, Rm = oracle_mod.roe_eig(d, ul, ur, G)
        rho, u, v, H, c = roe
        un = u if d == 0 else v
        A = Rm @ np.diag([un - c, un, un, un + c]) @ np.linalg.inv(Rm)
        lhs = A @ (ur - ul)
        rhs = euler_flux(ur, d) - euler_flux(ul, d)
        np.testing.assert_allclose(lhs, rhs, rtol=1e-10, atol=1e-10 * np.abs(rhs).max())


def test_roe_of_equal_states(oracle_mod):
    U = cons(1.3, 0.4, -0.7, 2.2)
    roe, _, _ = oracle_mod.roe_eig(0, U, U, G)
    np.testing.assert_allclose(roe, [1.3, 0.4, -0.7, (U[3] + 2.2) / 1.3, np.sqrt(G * 2.2 / 1.3)], rtol=1e-14)


def _quadratic_field(nx, ny, coef, h):
    """Cell averages (nx+6, ny+6, 4) of per-component quadratics in (x, y)."""
    x = (np.arange(nx + 6) - 3 + 0.5) * h
    y = (np.arange(ny + 6) - 3 + 0.5) * h
    X, Y = np.meshgrid(x, y, indexing="ij")
    q = np.zeros((nx + 6, ny + 6, 4))
    for k, (a, b, c, dd, e, f) in enumerate(coef):
        q[..., k] = a + b * X + c * Y + dd * (X * X + h * h / 12) + e * X * Y + f * (Y * Y + h * h / 12)
    return q


def test_componentwise_reconstruction_of_quadratic_fields(oracle_mod):
    nx, ny, h = 10, 8, 0.05
    coef = [(1.0, 0.3, -0.2, 0.1, 0.05, -0.1), (0.2, 0.1, 0.05, -0.1, 0.02, 0.03),
            (-0.1, 0.05, 0.2, 0.04, -0.03, 0.02), (3.0, 0.4, 0.3, -0.2, 0.1, 0.1)]
    q = _quadratic_field(nx, ny, coef, h)
    ql, fl, qr, fr = oracle_mod.weno_x_py(q, G, True)
    xf = (np.arange(nx + 1)) * h                       # face a at x = a h
    yc = (np.arange(ny) + 0.5) * h
    Xf, Yc = np.meshgrid(xf, yc, indexing="ij")
    for k, (a, b, c, dd, e, f) in enumerate(coef):
        exact = a + b * Xf + c * Yc + dd * Xf * Xf + e * Xf * Yc + f * (Yc * Yc + h * h / 12)
        np.testing.assert_allclose(ql[..., k], exact, atol=1e-9)
        np.testing.assert_allclose(qr[..., k], exact, atol=1e-9)


@pytest.mark.parametrize("d", [0, 1])
def test_characteristic_reconstruction_of_quadratic_fields(oracle_mod, d):
    """Fields quadratic along the face normal and uniform along the face: the
    projected WENO5 is exact for each characteristic field, the Gauss pass
    sees constants, so q_face = R L q_exact with the Roe basis of the two
    cells adjacent to the face (and the limiter idle)."""
    nx, h = 12, 0.05
    coef = [(1.0, 0.3, 0.1), (0.2, 0.1, -0.1), (-0.1, 0.05, 0.04), (3.0, 0.4, -0.2)]
    x = (np.arange(nx + 6) - 3 + 0.5) * h
    q1 = np.stack([a + b * x + c * (x * x + h * h / 12) for a, b, c in coef], axis=-1)
    q = np.repeat(q1[:, None, :], 10, axis=1)
    if d == 1:
        q = np.ascontiguousarray(q.transpose(1, 0, 2))
    ql, fl, qr, fr = (oracle_mod.weno_x_py if d == 0 else oracle_mod.weno_y_py)(q, G, False)
    for a in range(nx + 1):
        xf = a * h
        exact = np.array([aa + b * xf + c * xf * xf for aa, b, c in coef])
        c2, c3 = q1[a + 2], q1[a + 3]
        _, Lm, Rm = oracle_mod.roe_eig(d, c2, c3, G)
        want = Rm @ Lm @ exact
        got_l = ql[a, 2] if d == 0 else ql[2, a]
        got_r = qr[a, 2] if d == 0 else qr[2, a]
        np.testing.assert_allclose(got_l, want, rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(got_r, want, rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("cons_switch", [False, True])
def test_xy_equivariance(oracle_mod, cons_switch):
    rng = np.random.default_rng(21)
    q = np.zeros((14, 12, 4))
    q[..., 0] = rng.uniform(0.5, 2.0, (14, 12))
    q[..., 1] = rng.uniform(-1, 1, (14, 12))
    q[..., 2] = rng.uniform(-1, 1, (14, 12))
    q[..., 3] = rng.uniform(3, 6, (14, 12))
    qt = np.ascontiguousarray(q.transpose(1, 0, 2)[..., [0, 2, 1, 3]])
    ys = oracle_mod.weno_y_py(q, G, cons_switch)
    xs = oracle_mod.weno_x_py(qt, G, cons_switch)
    for a, b in zip(ys, xs):
        bb = b.transpose(1, 0, 2)[..., [0, 2, 1, 3]]
        np.testing.assert_allclose(a, bb, rtol=1e-13, atol=1e-13)
    Ly = oracle_mod.l_local_py(q, 0.1, G, False, (0.5, 2.0), cons_switch)
    Lt = oracle_mod.l_local_py(qt, 0.1, G, False, (0.5, 2.0), cons_switch)
    np.testing.assert_allclose(Ly, Lt.transpose(1, 0, 2)[..., [0, 2, 1, 3]], rtol=1e-12, atol=1e-11)


# ------------------------------------------------------------- Riemann ----
def toro_hllc(ul, ur, d, sl, sr):
    """Toro (2009) eq. 10.71-10.73 with given wave speeds: pressure form."""
    def prim(U):
        rho = U[0]
        u, v = U[1] / rho, U[2] / rho
        p = (G - 1) * (U[3] - 0.5 * rho * (u * u + v * v))
        return rho, (u if d == 0 else v), p
    rl, wl, pl = prim(ul)
    rr, wr, pr = prim(ur)
    fl, fr = euler_flux(ul, d), euler_flux(ur, d)
    s_star = (pr - pl + rl * wl * (sl - wl) - rr * wr * (sr - wr)) / (rl * (sl - wl) - rr * (sr - wr))
    D = np.array([0.0, 1.0 if d == 0 else 0.0, 0.0 if d == 0 else 1.0, s_star])
    if sl >= 0:
        return fl
    if s_star >= 0:
        p_star = pl + rl * (sl - wl) * (s_star - wl)
        return (s_star * (sl * ul - fl) + sl * p_star * D) / (sl - s_star)
    if sr > 0:
        p_star = pr + rr * (sr - wr) * (s_star - wr)
        return (s_star * (sr * ur - fr) + sr * p_star * D) / (sr - s_star)
    return fr


@pytest.mark.parametrize("d", [0, 1])
def test_hllc_matches_toro_pressure_form(oracle_mod, d):
    rng = np.random.default_rng(30 + d)
    n_star = 0
    for ul, ur in _states(rng, 300):
        roe, _, _ = oracle_mod.roe_eig(d, ul, ur, G)
        # HLLC uses the Roe speeds with a floor at 0 on c^2 (riemann.rs:47)
        rho_l, rho_r = ul[0], ur[0]
        sq_l, sq_r = np.sqrt(rho_l), np.sqrt(rho_r)
        def hh(U):
            rho = U[0]; u, v = U[1] / rho, U[2] / rho
            return (U[3] + (G - 1) * (U[3] - 0.5 * rho * (u * u + v * v))) / rho
        ut = (sq_l * ul[1] / rho_l + sq_r * ur[1] / rho_r) / (sq_l + sq_r)
        vt = (sq_l * ul[2] / rho_l + sq_r * ur[2] / rho_r) / (sq_l + sq_r)
        ht = (sq_l * hh(ul) + sq_r * hh(ur)) / (sq_l + sq_r)
        at = np.sqrt(max((G - 1) * (ht - 0.5 * (ut * ut + vt * vt)), 0.0))
        wt = ut if d == 0 else vt
        sl, sr = wt - at, wt + at
        got = oracle_mod.riemann_pt(d, 0, ul, euler_flux(ul, d), ur, euler_flux(ur, d), G)
        ref = toro_hllc(ul, ur, d, sl, sr)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10 * (1 + np.abs(ref).max()))
        n_star += sl < 0 < sr
    assert n_star > 100   # the star regions were exercised


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("d", [0, 1])
def test_riemann_consistency_and_upwinding(oracle_mod, kind, d):
    rng = np.random.default_rng(40)
    for ul, _ in _states(rng, 30):
        f = euler_flux(ul, d)
        np.testing.assert_allclose(oracle_mod.riemann_pt(d, kind, ul, f, ul, f, G), f, rtol=1e-12, atol=1e-12)
    # supersonic to the right / left
    a = cons(1.0, 5.0, 5.0, 1.0)
    b = cons(0.5, 6.0, 6.0, 0.8)
    np.testing.assert_array_equal(oracle_mod.riemann_pt(d, kind, a, euler_flux(a, d), b, euler_flux(b, d), G), euler_flux(a, d))
    a = cons(1.0, -5.0, -5.0, 1.0)
    b = cons(0.5, -6.0, -6.0, 0.8)
    np.testing.assert_array_equal(oracle_mod.riemann_pt(d, kind, a, euler_flux(a, d), b, euler_flux(b, d), G), euler_flux(b, d))


def test_hlle_middle_is_hll(oracle_mod):
    rng = np.random.default_rng(50)
    for ul, ur in _states(rng, 50):
        roe, _, _ = oracle_mod.roe_eig(0, ul, ur, G)
        got = oracle_mod.riemann_pt(0, 1, ul, euler_flux(ul), ur, euler_flux(ur), G)
        # invert: HLL = (sr fl - sl fr + sl sr (ur - ul)) / (sr - sl) is linear in the fluxes;
        # consistency of the middle formula: adding the same constant flux to both sides adds it to the result
        c = np.array([0.3, -0.2, 0.1, 0.7])
        got2 = oracle_mod.riemann_pt(0, 1, ul, euler_flux(ul) + c, ur, euler_flux(ur) + c, G)
        np.testing.assert_allclose(got2 - got, c, rtol=1e-9, atol=1e-9)


# ----------------------------------------------------------- limiter ----
def test_pp_limit(oracle_mod):
    ub = cons(1.0, 0.3, -0.2, 1.0)
    q = cons(1.1, 0.2, -0.1, 1.2)
    np.testing.assert_array_equal(oracle_mod.pp_limit(q, ub, G), q)
    # negative density face
    q = np.array([-0.2, 0.1, 0.0, 0.5])
    r = oracle_mod.pp_limit(q, ub, G)
    assert r[0] >= 1e-12 * (1 - 1e-9)
    p = (G - 1) * (r[3] - 0.5 * (r[1] ** 2 + r[2] ** 2) / r[0])
    assert p >= 1e-12 * (1 - 1e-6)
    # negative pressure face
    q = np.array([1.0, 3.0, 0.0, 0.5])
    r = oracle_mod.pp_limit(q, ub, G)
    p = (G - 1) * (r[3] - 0.5 * (r[1] ** 2 + r[2] ** 2) / r[0])
    assert p >= 1e-12 * (1 - 1e-6)
    # the limited state lies on the segment between q and the average
    with np.errstate(invalid="ignore", divide="ignore"):
        t = (r - ub) / (q - ub)
    t = t[np.isfinite(t)]
    assert np.allclose(t, t[0]) and 0 <= t[0] <= 1


def test_pp_limit_closed_forms(oracle_mod):
    """weno.rs:327-352.  Density is linear along the segment q -> u_bar, so the
    density pass lands on RHO_MIN = 1e-12 up to the rounding of O(1) values;
    pressure is linear along it when q and u_bar share density and momentum,
    and then the pressure pass lands on P_MIN = 1e-12.  A theta that drops the
    floor, over-limits or limits toward the wrong end misses these by ~1e-12."""
    tol = 4e-16 * 4.0
    # density pass only (no momentum, positive energy: the pressure stays positive)
    ub = np.array([1.0, 0.0, 0.0, 2.5])
    for rho_q in (-0.5, 0.0, 5e-13):
        r = oracle_mod.pp_limit(np.array([rho_q, 0.0, 0.0, 2.0]), ub, G)
        assert abs(r[0] - 1e-12) <= tol, r
        assert r[1] == 0.0 and r[2] == 0.0 and 2.0 <= r[3] <= 2.5
    # an average below the floor gives theta = 0: the face takes the average
    ub_low = np.array([1e-13, 0.0, 0.0, 1.0])
    np.testing.assert_array_equal(oracle_mod.pp_limit(np.array([-0.3, 0.2, 0.1, 0.5]), ub_low, G), ub_low)
    # pressure pass with the same density and momentum: p lands on 1e-12
    for (rho, u, v, p) in ((1.0, 0.3, -0.2, 1.0), (0.2, -1.5, 0.7, 0.05), (3.0, 0.0, 2.0, 10.0)):
        ub = cons(rho, u, v, p)
        kin = 0.5 * (ub[1] ** 2 + ub[2] ** 2) / ub[0]
        for p_q in (-0.5 * p, -1e-3, 0.0, 1e-13):
            q = ub.copy()
            q[3] = kin + p_q / (G - 1)
            r = oracle_mod.pp_limit(q, ub, G)
            pr = (G - 1) * (r[3] - 0.5 * (r[1] ** 2 + r[2] ** 2) / r[0])
            assert abs(pr - 1e-12) <= tol * max(p, kin) * 4, (pr, p_q)
            assert r[0] == ub[0]


# ----------------------------------------------------------- full RHS ----
@pytest.mark.parametrize("cons_switch", [False, True])
@pytest.mark.parametrize("force_hlle", [False, True])
def test_free_stream_and_y_uniform(oracle_mod, cons_switch, force_hlle):
    q = np.zeros((16, 14, 8))
    q[..., :4] = cons(1.3, 0.3, 0.5, 2.0)
    L = oracle_mod.l_local_py(q, 0.1, G, force_hlle, (0.5, 2.0), cons_switch)
    assert np.abs(L).max() == 0.0
    # data varying in x only: the y flux difference vanishes, L = -dF/dx of a 1-D problem
    x = np.linspace(0, 1, 16)[:, None] * np.ones((1, 14))
    q[..., 0] = 1.0 + 0.5 * (x > 0.5)
    q[..., 1] = 0.2 * q[..., 0]
    q[..., 2] = 0.0
    q[..., 3] = (1.0 + 0.3 * (x > 0.5)) / (G - 1) + 0.5 * 0.04 * q[..., 0]
    L = oracle_mod.l_local_py(q, 0.1, G, force_hlle, (0.5, 2.0), cons_switch)
    assert np.abs(L[..., 2]).max() < 1e-12
    np.testing.assert_allclose(L, L[:, :1, :] * np.ones((1, L.shape[1], 1)), atol=1e-12)


# ------------------------------------------------------------ drivers ----
def test_rk3_amplification(oracle_mod, monkeypatch):
    z = -0.37
    def fake_l(u, dx, gamma, force_hlle, jp_cri, cons_switch=False):
        return z * u[3:-3, 3:-3, :4]
    monkeypatch.setattr(oracle_mod, "l_local_py", fake_l)
    q = np.ones((10, 10, 8))
    q[..., 3] = 5.0
    qn = oracle_mod.rk3_cc(q.copy(), 4, 4, 0.1, 1.0, G, False, False)
    amp = 1 + z + z * z / 2 + z ** 3 / 6
    np.testing.assert_allclose(qn[3:7, 3:7, :4], amp * q[3:7, 3:7, :4], rtol=1e-15)
    qn = oracle_mod.rk3_refac(q.copy(), 4, 4, 0.1, 1.0, G, False, 0)
    np.testing.assert_allclose(qn[3:7, 3:7, :4], amp * q[3:7, 3:7, :4], rtol=1e-15)


def test_mirror_bc(oracle_mod):
    rng = np.random.default_rng(2)
    q = rng.standard_normal((12, 11, 8))
    oracle_mod.apply_bc_zero_gradient(q)
    for g in range(3):
        np.testing.assert_array_equal(q[g], q[5 - g])
        np.testing.assert_array_equal(q[11 - g], q[6 + g])
        np.testing.assert_array_equal(q[:, g], q[:, 5 - g])
        np.testing.assert_array_equal(q[:, 10 - g], q[:, 5 + g])


def test_threads_do_not_change_results(oracle_mod):
    """The multi-core comparator (orc_set_threads) splits outer loops only:
    the whole RHS is bitwise the same on 1 and 4 threads."""
    rng = np.random.default_rng(11)
    q = np.zeros((40, 34, 8))
    q[..., 0] = 1.0 + 0.5 * rng.random(q.shape[:2])
    q[..., 1] = 0.3 * rng.standard_normal(q.shape[:2])
    q[..., 2] = 0.3 * rng.standard_normal(q.shape[:2])
    q[..., 3] = 2.5 + rng.random(q.shape[:2])
    out = {}
    try:
        for t in (1, 4):
            oracle_mod.set_threads(t)
            out[t] = oracle_mod.l_local_py(q, 0.02, G, False, (0.5, 2.0), False)
    finally:
        oracle_mod.set_threads(1)
    assert np.array_equal(out[1], out[4])
