"""Write golden vectors of the NumPy twin's own semantics (twin mode).

Run in the build container only (the reference tree is not present anywhere
else):  python tests/golden/make_twin_mode_golden.py [/path/to/reference]

``twin_golden.npz`` (make_golden.py) pins the Rust semantics through the twin
with its Gauss constant patched.  This script leaves ``WENO_HLLC_2d.py``
unpatched and records what its own functions return, for the solver's twin
mode: the exact sqrt(3)/12 (:5), its HLLC (:435-574, right-star coefficient
rho_R (S* - u_R) / (S_R - S*), wave speeds without a floor under the root,
masks assigned in order) and its HLLE (:576-669, raw S_R - S_L, unfloored
states, S_R <= 0 assigned last), the blend and L built on them (:706-757),
and short runs of its time loop (main :840-905: quadrant initial data,
CFL 0.2, RK3_correct with the constant-extrapolation ghost fill, pressure
floor after each step).

Inputs are restricted to faces where the twin's Riemann solvers assign every
output element: where (gamma-1)(H~ - |V~|^2/2) < 0 its HLLC takes the root
of a negative number, no mask holds and F keeps np.empty's contents, which is
no reference value.  ``riemann_edge`` holds HLLE-only rows for the cases that
separate its mask order and raw states from the Rust HLLE.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import _four_quadrant  # noqa: E402

SEED = 20261204
G = 1.4
JP = (1.0, 4.0)   # sensor window of WENO_HLLC_2d.main's default


def _prim(q):
    rho = np.maximum(q[..., 0], 1e-12)
    u, v = q[..., 1] / rho, q[..., 2] / rho
    p = np.maximum((G - 1) * (q[..., 3] - 0.5 * rho * (u * u + v * v)), 1e-12)
    return rho, u, v, p, (q[..., 3] + p) / rho


def _assigned(ql, qr, direction):
    """Faces where the twin's HLLC assigns its output: positive squared Roe
    sound speed and a star-speed denominator outside the guard band."""
    rl, ul, vl, pl, hl = _prim(ql)
    rr, ur, vr, pr, hr = _prim(qr)
    a, b = np.sqrt(rl), np.sqrt(rr)
    ut, vt, ht = (a * ul + b * ur) / (a + b), (a * vl + b * vr) / (a + b), (a * hl + b * hr) / (a + b)
    c2 = (G - 1) * (ht - 0.5 * (ut ** 2 + vt ** 2))
    c = np.sqrt(np.maximum(c2, 0))
    w, wl, wr = (ut, ul, ur) if direction == 0 else (vt, vl, vr)
    den = rl * (w - c - wl) - rr * (w + c - wr)
    return (c2 > 0) & (np.abs(den) >= 1e-12)


def _field(rng, shape, rho, vel, p):
    """Random cell states with density, velocity components and pressure
    drawn uniformly from the given ranges."""
    r = rng.uniform(*rho, shape)
    u = rng.uniform(*vel, shape)
    v = rng.uniform(*vel, shape)
    pp = rng.uniform(*p, shape)
    q = np.zeros(shape + (8,))
    q[..., 0], q[..., 1], q[..., 2] = r, r * u, r * v
    q[..., 3] = pp / (G - 1) + 0.5 * r * (u ** 2 + v ** 2)
    q[..., 4:] = np.stack([r, u, v, pp], axis=-1)
    return q


def _supersonic(rng, shape):
    """Mean flow (2.25, -2.25) at sound speed about 0.84 with small random
    perturbations: supersonic toward +x and -y (HLLC/HLLE's S_L >= 0 and
    S_R <= 0 branches)."""
    q = _field(rng, shape, (0.8, 1.2), (-0.25, 0.25), (0.4, 0.6))
    r, u, v, p = q[..., 4], q[..., 5] + 2.25, q[..., 6] - 2.25, q[..., 7]
    q[..., 1], q[..., 2] = r * u, r * v
    q[..., 3] = p / (G - 1) + 0.5 * r * (u ** 2 + v ** 2)
    q[..., 5], q[..., 6] = u, v
    return q


def _twin_ic(n):
    """WENO_HLLC_2d.main's initial data (:846-880), written out here."""
    u = np.zeros((n + 6, n + 6, 8))
    h = int((n + 6) / 2)
    u[:h, :h, 4:] = (0.125, 1, 1, 0.125)
    u[h:, :h, 4:] = (0.5, 0, 1, 0.25)
    u[:h, h:, 4:] = (0.5, 1, 0, 0.25)
    u[h:, h:, 4:] = (1, 0, 0, 1)
    u[:, :, 0] = u[:, :, 4]
    u[:, :, 1] = u[:, :, 4] * u[:, :, 5]
    u[:, :, 2] = u[:, :, 4] * u[:, :, 6]
    u[:, :, 3] = u[:, :, 7] / (G - 1) + 0.5 * u[:, :, 4] * (u[:, :, 5] ** 2 + u[:, :, 6] ** 2)
    return u


def _edge_rows(rng, k):
    """Face pairs (1, k, 4) for HLLE: supersonic both ways, a near-vacuum
    side (rho 1e-14, below the 1e-12 floor) and negative-enthalpy pairs with
    zero Roe velocity (S_L = S_R = 0: both masks hold)."""
    ql = np.zeros((k, 4))
    qr = np.zeros((k, 4))
    for i in range(k):
        kind = i % 4
        r1, r2 = rng.uniform(0.3, 2.0, 2)
        p1, p2 = rng.uniform(0.1, 1.0, 2)
        if kind == 0:       # supersonic, sign alternating
            s = 1.0 if (i // 4) % 2 == 0 else -1.0
            v1, v2 = s * rng.uniform(3, 5), s * rng.uniform(3, 5)
        elif kind == 1:     # near-vacuum left face
            r1, v1, v2 = 1e-14, rng.uniform(-1, 1) * 1e-14, rng.uniform(-1, 1)
        elif kind == 2:     # negative energy, opposite velocities, equal density
            r2 = r1
            v1 = rng.uniform(0.5, 1.5)
            v2 = -v1
        else:               # ordinary jump
            v1, v2 = rng.uniform(-1, 1, 2)
        t1 = rng.uniform(-0.5, 0.5)
        t2 = -t1 if kind == 2 else rng.uniform(-0.5, 0.5)   # kind 2: zero Roe velocity in x and y
        for q, r, v, t, p in ((ql, r1, v1, t1, p1), (qr, r2, v2, t2, p2)):
            q[i] = (r, r * v, r * t, p / (G - 1) + 0.5 * r * (v * v + t * t))
            if kind == 2:
                q[i, 3] = -1.0
    return ql[None], qr[None]


def main(ref):
    sys.path.insert(0, ref)
    import matplotlib
    matplotlib.use("Agg")
    import WENO_HLLC_2d as tw  # noqa: E402  (reference twin, this container only)
    assert tw.SQRT3_12 == np.sqrt(3) / 12
    rng = np.random.default_rng(SEED)
    out = {"sqrt3_12": np.array(tw.SQRT3_12)}
    cases = {
        "rand16x12": _field(rng, (16 + 6, 12 + 6), (0.7, 1.3), (-0.6, 0.6), (0.6, 1.4)),
        "quad24x18": _four_quadrant(24, 18),
        "super20x16": _supersonic(rng, (20 + 6, 16 + 6)),
    }
    for name, u in cases.items():
        u = np.ascontiguousarray(u)
        out[f"{name}/u"] = u
        faces = {"x": tw.weno_x(u, G), "y": tw.weno_y(u, G)}
        for d, (ql, fl, qr, fr) in faces.items():
            assert _assigned(ql, qr, 0 if d == "x" else 1).all(), (name, d)
            out.update({f"{name}/weno_{d}/{k}": v for k, v in zip(("q_l", "f_l", "q_r", "f_r"), (ql, fl, qr, fr))})
        (qlx, flx, qrx, frx), (qly, fly, qry, fry) = faces["x"], faces["y"]
        out[f"{name}/hllc_x"] = tw.HLLC_x(qlx, qrx, flx, frx, G)
        out[f"{name}/hllc_y"] = tw.HLLC_y(qly, qry, fly, fry, G)
        out[f"{name}/hlle_x"] = tw.HLLE_x(qlx, qrx, flx, frx, G)
        out[f"{name}/hlle_y"] = tw.HLLE_y(qly, qry, fly, fry, G)
        out[f"{name}/blend_x"] = tw.HLLC_HLLE_x(qlx, qrx, flx, frx, G, *JP)
        out[f"{name}/blend_y"] = tw.HLLC_HLLE_y(qly, qry, fly, fry, G, *JP)
        dx = 1.0 / (u.shape[0] - 6)
        out[f"{name}/dx"] = np.array(dx)
        out[f"{name}/L_blend"] = tw.L(u, dx, G, False, JP)
        out[f"{name}/L_hlle"] = tw.L(u, dx, G, True, JP)
        pl = tw.con2primi(qlx, G)[0]
        pr = tw.con2primi(qrx, G)[0]
        phi = tw.shock_sensor_grad_p(pr, pl, *JP)
        print(f"{name}: x-face blend weights 0 / mixed / 1: "
              f"{(phi == 0).sum()} / {((phi > 0) & (phi < 1)).sum()} / {(phi == 1).sum()}")

    # supersonic and star-region faces of HLLC, drawn at random and filtered
    ql = _field(rng, (1, 400), (0.1, 3.0), (-3.0, 3.0), (0.05, 2.0))[..., :4]
    qr = _field(rng, (1, 400), (0.1, 3.0), (-3.0, 3.0), (0.05, 2.0))[..., :4]
    for d, fl_fn in ((0, tw.flux_x), (1, tw.flux_y)):
        keep = _assigned(ql, qr, d)[0]
        a, b = ql[:, keep], qr[:, keep]
        fa, fb = fl_fn(a, G), fl_fn(b, G)
        nm = "x" if d == 0 else "y"
        out[f"riemann_rand/{nm}/q_l"], out[f"riemann_rand/{nm}/q_r"] = a, b
        out[f"riemann_rand/{nm}/f_l"], out[f"riemann_rand/{nm}/f_r"] = fa, fb
        hl = tw.HLLC_x if d == 0 else tw.HLLC_y
        he = tw.HLLE_x if d == 0 else tw.HLLE_y
        out[f"riemann_rand/{nm}/hllc"] = hl(a, b, fa, fb, G)
        out[f"riemann_rand/{nm}/hlle"] = he(a, b, fa, fb, G)

    # HLLE-only edge rows
    el, er = _edge_rows(rng, 48)
    for d, fl_fn in ((0, tw.flux_x), (1, tw.flux_y)):
        nm = "x" if d == 0 else "y"
        fa, fb = fl_fn(el, G), fl_fn(er, G)
        out[f"riemann_edge/{nm}/q_l"], out[f"riemann_edge/{nm}/q_r"] = el, er
        out[f"riemann_edge/{nm}/f_l"], out[f"riemann_edge/{nm}/f_r"] = fa, fb
        out[f"riemann_edge/{nm}/hlle"] = (tw.HLLE_x if d == 0 else tw.HLLE_y)(el, er, fa, fb, G)

    # the pressure floor (:813-837) on cells around and below pmin = 1e-8,
    # some with density under its 1e-12 floor
    pf = _field(rng, (12, 10), (0.2, 2.0), (-1.0, 1.0), (0.1, 1.0))
    low = rng.random((12, 10)) < 0.4
    kin = 0.5 * pf[..., 1] ** 2 / pf[..., 0] + 0.5 * pf[..., 2] ** 2 / pf[..., 0]
    pf[..., 3] = np.where(low, kin + rng.uniform(-1e-6, 3e-8, (12, 10)) / (G - 1), pf[..., 3])
    pf[0, :3, 0] = (1e-13, 0.0, -1e-3)
    out["pfloor/in"] = pf.copy()
    out["pfloor/out"] = tw.con_var2primi_pfloor(pf, G)

    # the time loop of main (:895-905), I/O left out
    for name, n, n_step, force in (("loop24", 24, 8, False), ("loop20_hlle", 20, 6, True)):
        u = tw.apply_bc_corrected(_twin_ic(n), n)
        out[f"{name}/u0"] = u.copy()
        dx = 1 / n
        T, ts = 0, [0.0]
        for _ in range(1, n_step):
            speed = np.sqrt(u[3:n + 3, 3:n + 3, 5] ** 2 + u[3:n + 3, 3:n + 3, 6] ** 2)
            c_max = np.max(np.sqrt(G * u[3:n + 3, 3:n + 3, 7] / np.maximum(u[3:n + 3, 3:n + 3, 0], 1e-10)) + speed)
            dt = 0.2 * dx / c_max
            u = tw.RK3_correct(u, n, dt, G, force, JP)
            u = tw.con_var2primi_pfloor(u, G)
            T += dt
            ts.append(T)
        out[f"{name}/u"] = u
        out[f"{name}/t"] = np.array(ts)
        out[f"{name}/meta"] = np.array([n, n_step, int(force)])
    dst = os.path.join(HERE, "twin_mode_golden.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {dst}: {len(out)} arrays, {os.path.getsize(dst)} bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
