"""Write golden input/output vectors from the reference's NumPy twin.

Run in the build container only (the reference tree is not present anywhere
else):  python tests/golden/make_golden.py [/path/to/reference]

The reference ships no fixtures or tests for its hot path.  Its pure-NumPy
twin ``WENO_HLLC_2d.py`` is an implementation of the same component-wise
WENO5 + Gauss-quadrature reconstruction, fluxes, HLLE and pressure-gradient
sensor written independently of the Rust crate.  With its ``SQRT3_12`` global
set to the Rust constant 0.1443375673 (``weno_ext/src/utils.rs:4``; the twin
uses the exact sqrt(3)/12 at ``WENO_HLLC_2d.py:5``) its outputs equal the Rust
semantics operation for operation on the inputs below (no density below 1e-12,
no degenerate wave-speed denominators).  These vectors pin the parts of the
oracle that closed forms and invariants leave open: the WENO5-JS smoothness
indicators and weights, the Gauss-point weights, the flux averaging, HLLE's
middle state, and the sensor.

Functions used (twin file:line): flux_x/flux_y :7-60, weno_x :62-228,
weno_y :231-398, con2primi :400-432, HLLE_x/HLLE_y :576-669,
shock_sensor_grad_p :682-704, HLLC_HLLE_x/y :706-741 (with a sensor window
that forces phi = 0, i.e. the HLLE branch), L :743-757 (force_HLLE=True).

Not used: the twin's HLLC (its right-star coefficient differs from the Rust,
``WENO_HLLC_2d.py:476,547`` versus ``riemann.rs:67,160``).
"""
import os
import sys

import numpy as np

SQRT3_12_RUST = 0.1443375673
SEED = 20261203


def _admissible(rng, shape, gamma=1.4):
    rho = rng.uniform(0.2, 2.0, shape)
    u = rng.uniform(-1.2, 1.2, shape)
    v = rng.uniform(-1.2, 1.2, shape)
    p = rng.uniform(0.1, 2.0, shape)
    q = np.zeros(shape + (8,))
    q[..., 0] = rho
    q[..., 1] = rho * u
    q[..., 2] = rho * v
    q[..., 3] = p / (gamma - 1.0) + 0.5 * rho * (u ** 2 + v ** 2)
    q[..., 4], q[..., 5], q[..., 6], q[..., 7] = rho, u, v, p
    return q


def _four_quadrant(nx, ny, gamma=1.4):
    q = np.zeros((nx + 6, ny + 6, 8))
    xs = np.arange(nx + 6)[:, None] < 3 + nx // 2
    ys = np.arange(ny + 6)[None, :] < 3 + ny // 2
    states = {(False, False): (1.5, 0, 0, 1.5), (True, False): (0.5323, 1.206, 0, 0.3),
              (True, True): (0.138, 1.206, 1.206, 0.029), (False, True): (0.5323, 0, 1.206, 0.3)}
    for (lx, ly), (r, u, v, p) in states.items():
        mask = (xs == lx) & (ys == ly)
        q[..., 0][mask] = r
        q[..., 1][mask] = r * u
        q[..., 2][mask] = r * v
        q[..., 3][mask] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return q


def main(ref):
    sys.path.insert(0, ref)
    import matplotlib
    matplotlib.use("Agg")
    import WENO_HLLC_2d as tw  # noqa: E402  (reference twin, this container only)
    tw.SQRT3_12 = SQRT3_12_RUST
    rng = np.random.default_rng(SEED)
    g = 1.4
    out = {}
    cases = {
        "rand16x12": _admissible(rng, (16 + 6, 12 + 6)),
        "rand20x20": _admissible(rng, (20 + 6, 20 + 6)),
        "quad24x18": _four_quadrant(24, 18),
    }
    # a smooth field with a sharp front (exercises nonlinear weights)
    x = np.linspace(0, 1, 26)[:, None]
    y = np.linspace(0, 1, 22)[None, :]
    sm = _admissible(np.random.default_rng(SEED + 1), (26, 22))
    rho = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.3 * (x > 0.55)
    sm[..., 0] = rho
    sm[..., 1] = rho * (0.3 + 0.2 * np.cos(2 * np.pi * y))
    sm[..., 2] = rho * (-0.4 + 0.1 * np.sin(2 * np.pi * x))
    sm[..., 3] = (1.0 + 0.2 * (y > 0.4)) / (g - 1) + 0.5 * (sm[..., 1] ** 2 + sm[..., 2] ** 2) / rho
    cases["smooth26x22"] = sm
    for name, u in cases.items():
        u = np.ascontiguousarray(u)
        out[f"{name}/u"] = u
        qlx, flx, qrx, frx = tw.weno_x(u, g)
        qly, fly, qry, fry = tw.weno_y(u, g)
        out.update({f"{name}/weno_x/{k}": v for k, v in zip(("q_l", "f_l", "q_r", "f_r"), (qlx, flx, qrx, frx))})
        out.update({f"{name}/weno_y/{k}": v for k, v in zip(("q_l", "f_l", "q_r", "f_r"), (qly, fly, qry, fry))})
        out[f"{name}/flux_x"] = tw.flux_x(u[..., :4], g)
        out[f"{name}/flux_y"] = tw.flux_y(u[..., :4], g)
        p, a, rs, uu, vv = tw.con2primi(u, g)
        for k, v in zip(("p", "a", "rho", "u", "v"), (p, a, rs, uu, vv)):
            out[f"{name}/con2primi/{k}"] = v
        out[f"{name}/hlle_x"] = tw.HLLE_x(qlx, qrx, flx, frx, g)
        out[f"{name}/hlle_y"] = tw.HLLE_y(qly, qry, fly, fry, g)
        pl, _, _, _, _ = tw.con2primi(qlx, g)
        pr, _, _, _, _ = tw.con2primi(qrx, g)
        out[f"{name}/sensor_x/p_l"] = pl
        out[f"{name}/sensor_x/p_r"] = pr
        out[f"{name}/sensor_x/phi"] = tw.shock_sensor_grad_p(pr, pl, 0.05, 0.6)
        pl, _, _, _, _ = tw.con2primi(qly, g)
        pr, _, _, _, _ = tw.con2primi(qry, g)
        out[f"{name}/sensor_y/p_l"] = pl
        out[f"{name}/sensor_y/p_r"] = pr
        out[f"{name}/sensor_y/phi"] = tw.shock_sensor_grad_p(pr, pl, 0.05, 0.6)
        out[f"{name}/blend0_x"] = tw.HLLC_HLLE_x(qlx, qrx, flx, frx, g, -2.0, -1.0)
        out[f"{name}/blend0_y"] = tw.HLLC_HLLE_y(qly, qry, fly, fry, g, -2.0, -1.0)
        dx = 1.0 / (u.shape[0] - 6)
        out[f"{name}/L_hlle/dx"] = np.array(dx)
        out[f"{name}/L_hlle"] = tw.L(u, dx, g, True, (0.5, 2.0))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "twin_golden.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {dst}: {len(out)} arrays, {os.path.getsize(dst)} bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
