"""jp_cri tuning on the device: the objective of ``metrics.py`` and the search
of ``opt_jp_cro.py``.

``compute_2d_objective`` has the signature and results of
``metrics.compute_2d_objective`` (``metrics.py:1-77``) and runs on the GPU
(``wenox_objective_2d``).  Its shock-mask thinning is the Zhang-Suen
algorithm, which ``skimage.morphology.skeletonize`` implements for 2-D images.

``evaluate_jp_pair`` and ``optimize_jp`` follow ``opt_jp_cro.py:9-80``: one
``main_rs`` run per candidate (``jp_lo >= jp_hi`` or a failed run scores
1e6), its final state scored by the objective -- here on the device, from the
solver's state, with no copy to the host -- and a Bayesian search over
jp_lo in [0.02, 1], jp_hi in [0.2, 10] (30 runs, 8 initial points, expected
improvement, ``random_state`` 0).  With scikit-optimize installed the search
is its ``gp_minimize`` with the reference's arguments; without it, the
search below (a Gaussian process with a Matern 5/2 kernel, expected
improvement over random candidates) takes its place, and its proposals
differ from gp_minimize's.  The reference's script imports a module that does
not exist (``Rie_prob_2D_rs``, :6) and calls an unqualified ``main_rs``
(:20); here the run is ``wenox``'s device-resident ``main_rs``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import Solver, _lib, problems

METRIC_NAMES = ("I_curve", "E_perp", "O", "TV", "cI", "cE", "cO", "cTV")
PENALTY = 1e6
SPACE = ((0.02, 1.0), (0.2, 10.0))   # opt_jp_cro.py:55-58


def _weights(weights):
    if weights is None:
        return None
    w = (ctypes.c_double * 4)(float(weights["I"]), float(weights["E"]), float(weights["O"]), float(weights["TV"]))
    return w


def _result(J, m):
    return float(J.value), dict(zip(METRIC_NAMES, (float(x) for x in m)))


def compute_2d_objective(rho, u, v, p=None, dx=1.0, dy=1.0, weights=None):
    """metrics.compute_2d_objective on the GPU.  Returns (J, metrics) with the
    reference's keys; ``p`` is accepted and not read, as in the reference."""
    rv, uv, vv = _lib.view2(rho), _lib.view2(u), _lib.view2(v)
    J = ctypes.c_double()
    m = (ctypes.c_double * 8)()
    w = _weights(weights)
    _lib.check(_lib.load().wenox_objective_2d(ctypes.byref(rv), ctypes.byref(uv), ctypes.byref(vv), float(dx),
                                              float(dy), w, ctypes.byref(J), m))
    return _result(J, m)


def objective_fields(rho, dx=1.0, dy=1.0):
    """Intermediates of the objective of ``rho``: (S, C, band) as bool arrays
    and the 7 x 7 median of rho."""
    rv = _lib.view2(rho)
    nx, ny = rho.shape
    S = np.empty((nx, ny), np.uint8)
    C = np.empty((nx, ny), np.uint8)
    band = np.empty((nx, ny), np.uint8)
    med = np.empty((nx, ny))
    _lib.check(_lib.load().wenox_objective_fields(ctypes.byref(rv), float(dx), float(dy), S.ctypes.data,
                                                  C.ctypes.data, band.ctypes.data, med.ctypes.data))
    return S.astype(bool), C.astype(bool), band.astype(bool), med


def solver_objective(solver: Solver, dx=1.0, dy=1.0, weights=None):
    """The objective of a solver's current state (rho, and the con2primi u, v
    of the interior), computed where the state lives."""
    J = ctypes.c_double()
    m = (ctypes.c_double * 8)()
    _lib.check(_lib.load().wenox_solver_objective(solver._h, float(dx), float(dy), _weights(weights),
                                                  ctypes.byref(J), m))
    return _result(J, m)


def evaluate_jp_pair(jp_lo, jp_hi, grid_dict, control_dict, phys_dict, engine="fused", variant="exact",
                     q0=None):
    """opt_jp_cro.py:9-47: run main_rs with jp_cri = (jp_lo, jp_hi) (the
    reference drivers' initial data unless q0 is given) and score its final
    state.  jp_lo >= jp_hi, a failed run or a non-finite score give 1e6."""
    if jp_lo >= jp_hi:
        return PENALTY
    gamma = phys_dict.get("gamma", 1.4)
    nx, ny = int(grid_dict["nx"]), int(grid_dict["ny"])
    dx = float(grid_dict.get("dx", 1.0))
    t_final = control_dict.get("t_final", None)
    nstep = control_dict.get("nstep", None)
    if q0 is None:
        q0 = problems.riemann_reference(nx, ny, gamma)
    try:
        with Solver(nx, ny, dx, gamma=gamma, cfl=float(control_dict.get("CFL", 0.5)),
                    cons_switch=bool(control_dict.get("weno type")),
                    force_hlle=bool(control_dict.get("force_hlle", False)), jp_cri=(float(jp_lo), float(jp_hi)),
                    bc="mirror", engine=engine, variant=variant) as s:
            s.upload(q0)
            s.run(nstep=nstep, t_final=t_final)
            score, _ = solver_objective(s)
    except Exception:
        return PENALTY
    return score if math.isfinite(score) else PENALTY


def _matern52(a, b, length):
    d = np.sqrt(np.maximum(((a[:, None, :] - b[None, :, :]) ** 2).sum(-1), 0.0)) / length
    return (1.0 + math.sqrt(5.0) * d + 5.0 / 3.0 * d * d) * np.exp(-math.sqrt(5.0) * d)


def gp_ei_minimize(func, bounds, n_calls=30, n_initial_points=8, random_state=0, xi=0.01, n_candidates=4096):
    """Minimise func over a box by Bayesian optimisation: n_initial_points
    uniform draws, then n_calls - n_initial_points proposals maximising the
    expected improvement (exploration xi) of a Gaussian process (Matern 5/2,
    length scale by marginal likelihood on a grid, 1e-6 nugget) fitted to the
    standardised scores.  Returns (x_best, f_best, xs, fs)."""
    from scipy.linalg import cho_factor, cho_solve, solve_triangular
    from scipy.special import ndtr

    rng = np.random.default_rng(random_state)
    lo = np.array([b[0] for b in bounds], float)
    hi = np.array([b[1] for b in bounds], float)
    xs, fs = [], []

    def ev(z):
        x = lo + z * (hi - lo)
        xs.append(x)
        fs.append(float(func(list(x))))

    for _ in range(min(n_initial_points, n_calls)):
        ev(rng.random(len(bounds)))
    while len(xs) < n_calls:
        Z = (np.array(xs) - lo) / (hi - lo)
        Y = np.array(fs)
        sd = Y.std() or 1.0
        Yn = (Y - Y.mean()) / sd
        best = None
        for length in (0.05, 0.1, 0.2, 0.3, 0.5, 0.8, 1.2):
            K = _matern52(Z, Z, length) + 1e-6 * np.eye(len(Z))
            try:
                cf = cho_factor(K, lower=True)
            except np.linalg.LinAlgError:
                continue
            alpha = cho_solve(cf, Yn)
            nll = 0.5 * Yn @ alpha + np.log(np.diag(cf[0])).sum()
            if best is None or nll < best[0]:
                best = (nll, length, cf, alpha)
        _, length, cf, alpha = best
        cand = rng.random((n_candidates, len(bounds)))
        Kc = _matern52(cand, Z, length)
        mean = Kc @ alpha
        w = solve_triangular(cf[0], Kc.T, lower=True)
        s = np.sqrt(np.maximum(1.0 - (w * w).sum(0), 1e-12))
        imp = Yn.min() - mean - xi
        zz = imp / s
        ei = imp * ndtr(zz) + s * np.exp(-0.5 * zz * zz) / math.sqrt(2 * math.pi)
        ev(cand[int(np.argmax(ei))])
    k = int(np.argmin(fs))
    return xs[k], fs[k], xs, fs


def optimize_jp(grid_dict, control_dict, phys_dict, n_calls=30, n_initial_points=8, random_state=0,
                engine="fused", variant="exact", q0=None):
    """opt_jp_cro.py:50-80.  Returns (best_jp_lo, best_jp_hi, best_score)."""
    def objective(x):
        return evaluate_jp_pair(x[0], x[1], grid_dict, control_dict, phys_dict, engine, variant, q0)

    try:
        from skopt import gp_minimize
        from skopt.space import Real
    except ImportError:
        x, f, _, _ = gp_ei_minimize(objective, SPACE, n_calls, n_initial_points, random_state)
        return float(x[0]), float(x[1]), float(f)
    space = [Real(*SPACE[0], name="jp_lo"), Real(*SPACE[1], name="jp_hi")]
    res = gp_minimize(objective, space, n_calls=n_calls, n_initial_points=n_initial_points,
                      random_state=random_state, acq_func="EI")
    return float(res.x[0]), float(res.x[1]), float(res.fun)
