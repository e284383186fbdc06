"""CPU oracle for the 2D WENO5/HLLC hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker (never as the thing that
is measured as the GPU product or shipped).

Two layers:

* ``libweno_oracle.so`` (``weno_oracle.c``): a plain-C, FP64, single-threaded
  restatement of the reference Rust crate ``weno_ext`` (``weno_ext/src/*.rs``),
  reached here through ctypes with the same call surface as the reference
  Python module (``weno_ext/src/lib.rs:18-153``).
* numpy restatements of the reference's Python drivers around it: ghost fill,
  primitive sync, SSP-RK3 and the CFL time step
  (``Rie_prob_2d_rs.py:42-207,311-319``, ``double_mach_reflection.py:97-134,
  229-260``, ``unitest.py:169-227``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libweno_oracle.so")


class _View3(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p),
                ("d0", ctypes.c_int64), ("d1", ctypes.c_int64), ("d2", ctypes.c_int64),
                ("s0", ctypes.c_int64), ("s1", ctypes.c_int64), ("s2", ctypes.c_int64)]


_lib = None


def build() -> str:
    """Compile the C oracle in place (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(_View3)
        d = ctypes.c_double
        dp = ctypes.c_void_p
        i = ctypes.c_int
        L.orc_flux.argtypes = [P, d, i, dp]
        L.orc_con2primi.argtypes = [P, d, dp, dp, dp, dp, dp, dp]
        L.orc_weno.argtypes = [P, d, i, i, dp, dp, dp, dp]
        L.orc_riemann.argtypes = [P, P, P, P, d, i, i, dp]
        L.orc_sensor.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, d, d, dp]
        L.orc_blend.argtypes = [P, P, P, P, d, i, d, d, dp]
        L.orc_l_local.argtypes = [P, d, d, i, d, d, i, dp]
        L.orc_test_roe_eig.argtypes = [i, dp, dp, d, dp, dp, dp]
        L.orc_test_pp_limit.argtypes = [dp, dp, d]
        L.orc_test_weno5.argtypes = [dp, dp, dp]
        L.orc_test_gq.argtypes = [dp, dp, dp]
        L.orc_test_riemann_pt.argtypes = [i, i, dp, dp, dp, dp, d, dp]
        L.orc_set_threads.argtypes = [i]
        L.orc_pp_counts.argtypes = [dp]
        L.orc_pp_reset.argtypes = []
        L.orc_set_twin.argtypes = [i]
        for f in (L.orc_pp_counts, L.orc_pp_reset, L.orc_set_threads, L.orc_set_twin, L.orc_flux,
                  L.orc_con2primi, L.orc_weno, L.orc_riemann,
                  L.orc_sensor, L.orc_blend, L.orc_l_local, L.orc_test_roe_eig,
                  L.orc_test_pp_limit, L.orc_test_weno5, L.orc_test_gq, L.orc_test_riemann_pt):
            f.restype = None
        _lib = L
    return _lib


_lib32 = None


def lib32():
    """libweno_oracle_f32.so: the same restatement built with every
    floating-point value in FP32 (test instrumentation; float32 arrays)."""
    global _lib32
    if _lib32 is None:
        path = os.path.join(_HERE, "libweno_oracle_f32.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(_View3)
        f, dp, i = ctypes.c_float, ctypes.c_void_p, ctypes.c_int
        L.orc_weno.argtypes = [P, f, i, i, dp, dp, dp, dp]
        L.orc_blend.argtypes = [P, P, P, P, f, i, f, f, dp]
        L.orc_riemann.argtypes = [P, P, P, P, f, i, i, dp]
        for fn in (L.orc_weno, L.orc_blend, L.orc_riemann):
            fn.restype = None
        _lib32 = L
    return _lib32


def _view32(a):
    s = [st // 4 for st in a.strides]
    return _View3(a.ctypes.data, a.shape[0], a.shape[1], a.shape[2], s[0], s[1], s[2])


def blend_f32(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, direction):
    """hllc_hlle_{x,y}_rs_py (riemann.rs:412-486) in FP32 arithmetic on the
    face states rounded to FP32; returns float64."""
    fs = [np.ascontiguousarray(x, dtype=np.float32) for x in (q_l, f_l, q_r, f_r)]
    o = np.zeros(fs[0].shape, dtype=np.float32)
    lib32().orc_blend(*[ctypes.byref(_view32(x)) for x in fs], gamma, direction, jp_lo, jp_hi, o.ctypes.data)
    return o.astype(np.float64)


class count_limiter:
    """Context manager counting the faces where the positivity limiter's
    density pass and pressure pass acted (weno.rs:327-352) in the oracle
    calls inside the block: ``with count_limiter() as c: ...; c.density,
    c.pressure``."""

    def __enter__(self):
        lib().orc_pp_reset()
        self.density = self.pressure = 0
        return self

    def __exit__(self, *exc):
        out = np.zeros(2, dtype=np.int64)
        lib().orc_pp_counts(out.ctypes.data)
        self.density, self.pressure = int(out[0]), int(out[1])


def set_threads(n: int) -> None:
    """Threads of the oracle's outer loops (1, the default, is the
    reference's single thread; results do not depend on n)."""
    lib().orc_set_threads(int(n))


class twin:
    """Context manager: the C functions follow the reference's pure-NumPy
    twin WENO_HLLC_2d.py (exact sqrt(3)/12, the twin's HLLC and HLLE) inside
    the block, the Rust crate outside it."""

    def __enter__(self):
        lib().orc_set_twin(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_twin(0)


def _view(a: np.ndarray) -> _View3:
    assert a.dtype == np.float64 and a.ndim == 3
    s = [st // 8 for st in a.strides]
    return _View3(a.ctypes.data, a.shape[0], a.shape[1], a.shape[2], s[0], s[1], s[2])


def _out(*shape):
    return np.zeros(shape, dtype=np.float64)


# --------------------------------------------------------------------------
# weno_ext surface (lib.rs:18-153), oracle semantics
# --------------------------------------------------------------------------
def flux_x_py(q, gamma):
    o = _out(q.shape[0], q.shape[1], 4)
    lib().orc_flux(ctypes.byref(_view(q)), gamma, 0, o.ctypes.data)
    return o


def flux_y_py(q, gamma):
    o = _out(q.shape[0], q.shape[1], 4)
    lib().orc_flux(ctypes.byref(_view(q)), gamma, 1, o.ctypes.data)
    return o


def con2primi_py(q, gamma):
    outs = [_out(q.shape[0], q.shape[1]) for _ in range(6)]
    lib().orc_con2primi(ctypes.byref(_view(q)), gamma, *[o.ctypes.data for o in outs])
    return tuple(outs)


def _weno(u, gamma, direction, cons):
    m, n = u.shape[0], u.shape[1]
    shp = (m - 5, n - 6, 4) if direction == 0 else (m - 6, n - 5, 4)
    outs = [_out(*shp) for _ in range(4)]
    lib().orc_weno(ctypes.byref(_view(u)), gamma, direction, int(cons),
                   *[o.ctypes.data for o in outs])
    return tuple(outs)


def weno_x_py(u, gamma, cons=False):
    return _weno(u, gamma, 0, cons)


def weno_y_py(u, gamma, cons=False):
    return _weno(u, gamma, 1, cons)


def _riemann(q_l, f_l, q_r, f_r, gamma, direction, kind):
    o = _out(q_l.shape[0], q_l.shape[1], 4)
    lib().orc_riemann(*[ctypes.byref(_view(a)) for a in (q_l, f_l, q_r, f_r)],
                      gamma, direction, kind, o.ctypes.data)
    return o


def hllc_x_rs_py(q_l, f_l, q_r, f_r, gamma):
    return _riemann(q_l, f_l, q_r, f_r, gamma, 0, 0)


def hllc_y_rs_py(q_l, f_l, q_r, f_r, gamma):
    return _riemann(q_l, f_l, q_r, f_r, gamma, 1, 0)


def hlle_x_rs_py(q_l, f_l, q_r, f_r, gamma):
    return _riemann(q_l, f_l, q_r, f_r, gamma, 0, 1)


def hlle_y_rs_py(q_l, f_l, q_r, f_r, gamma):
    return _riemann(q_l, f_l, q_r, f_r, gamma, 1, 1)


def _blend(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, force_hlle, direction):
    if force_hlle:
        return _riemann(q_l, f_l, q_r, f_r, gamma, direction, 1)
    o = _out(q_l.shape[0], q_l.shape[1], 4)
    lib().orc_blend(*[ctypes.byref(_view(a)) for a in (q_l, f_l, q_r, f_r)],
                    gamma, direction, jp_lo, jp_hi, o.ctypes.data)
    return o


def hllc_hlle_x_rs_py(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, force_hlle):
    return _blend(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, force_hlle, 0)


def hllc_hlle_y_rs_py(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, force_hlle):
    return _blend(q_l, f_l, q_r, f_r, gamma, jp_lo, jp_hi, force_hlle, 1)


def shock_sensor_grad_p(p, jp_lo, jp_hi):
    """riemann.rs:352-410 on a C-contiguous 2D face-grid pressure."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    o = np.zeros_like(p)
    lib().orc_sensor(p.ctypes.data, p.shape[0], p.shape[1], jp_lo, jp_hi, o.ctypes.data)
    return o


def l_local_py(u, dx, gamma, force_hlle, jp_cri, cons_switch=False):
    o = _out(u.shape[0] - 6, u.shape[1] - 6, 4)
    lib().orc_l_local(ctypes.byref(_view(u)), dx, gamma, int(force_hlle),
                      float(jp_cri[0]), float(jp_cri[1]), int(cons_switch), o.ctypes.data)
    return o


# --------------------------------------------------------------------------
# point-function hooks (for the oracle's own invariant tests)
# --------------------------------------------------------------------------
def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def roe_eig(direction, ul, ur, gamma):
    """Roe state (rho, u, v, H, c) and the left/right eigenvector matrices
    (weno.rs:13-189) of the characteristic basis between states ul | ur."""
    ul, ur = _c(ul), _c(ur)
    roe, Lm, Rm = np.zeros(5), np.zeros((4, 4)), np.zeros((4, 4))
    lib().orc_test_roe_eig(direction, ul.ctypes.data, ur.ctypes.data, gamma,
                           roe.ctypes.data, Lm.ctypes.data, Rm.ctypes.data)
    return roe, Lm, Rm


def pp_limit(q, ub, gamma):
    """weno.rs:327-352 applied to a copy of q."""
    q = _c(q).copy()
    ub = _c(ub)
    lib().orc_test_pp_limit(q.ctypes.data, ub.ctypes.data, gamma)
    return q


def weno5_pair(u6):
    """(weno5_left(u0..u4), weno5_right(u1..u5)) (weno.rs:214-257)."""
    u6 = _c(u6)
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_test_weno5(u6.ctypes.data, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def weno5_gq(q5):
    """(ga, gb) Gauss-point values (weno.rs:262-303)."""
    q5 = _c(q5)
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_test_gq(q5.ctypes.data, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def riemann_pt(direction, kind, ql, fl, qr, fr, gamma):
    """One face of HLLC (kind 0, riemann.rs:6-189) or HLLE (kind 1, :191-349)."""
    arrs = [_c(x) for x in (ql, fl, qr, fr)]
    out = np.zeros(4)
    lib().orc_test_riemann_pt(direction, kind, *[a.ctypes.data for a in arrs], gamma, out.ctypes.data)
    return out


# --------------------------------------------------------------------------
# Drivers (numpy restatements of the reference's Python)
# --------------------------------------------------------------------------
NG = 3


def apply_bc_zero_gradient(u):
    """Even mirror, y first then x, all channels (Rie_prob_2d_rs.py:123-152)."""
    nx, ny = u.shape[0], u.shape[1]
    for g in range(NG):
        u[:, g, :] = u[:, 2 * NG - 1 - g, :]
    for g in range(NG):
        u[:, ny - 1 - g, :] = u[:, ny - 2 * NG + g, :]
    for g in range(NG):
        u[g, :, :] = u[2 * NG - 1 - g, :, :]
    for g in range(NG):
        u[nx - 1 - g, :, :] = u[nx - 2 * NG + g, :, :]
    return u


_S_Y = np.array([1, 1, -1, 1, 1, 1, -1, 1], dtype=np.float64)
_S_X = np.array([1, -1, 1, 1, 1, -1, 1, 1], dtype=np.float64)
_S_C = np.array([1, -1, -1, 1, 1, -1, -1, 1], dtype=np.float64)


def apply_bc_reflecting(u):
    """Rie_prob_2d_rs.py:42-120, including its (m, n) = (shape[1], shape[0])."""
    m, n = u.shape[1], u.shape[0]
    S_Y, S_X, S_C = (_S_Y[:u.shape[2]], _S_X[:u.shape[2]], _S_C[:u.shape[2]])
    for g in range(NG):
        u[g, NG:n - NG, :] = u[2 * NG - 1 - g, NG:n - NG, :] * S_Y
    for g in range(NG):
        u[m - 1 - g, NG:n - NG, :] = u[m - 2 * NG + g, NG:n - NG, :] * S_Y
    for g in range(NG):
        u[NG:m - NG, g, :] = u[NG:m - NG, 2 * NG - 1 - g, :] * S_X
    for g in range(NG):
        u[NG:m - NG, n - 1 - g, :] = u[NG:m - NG, n - 2 * NG + g, :] * S_X
    for gi in range(NG):
        for gj in range(NG):
            mi, mj = 2 * NG - 1 - gi, 2 * NG - 1 - gj
            int_j, gh_j = n - 2 * NG + gj, n - NG + gj
            int_i, gh_i = m - 2 * NG + gi, m - NG + gi
            u[gi, gj, :] = u[mi, mj, :] * S_C
            u[gi, gh_j, :] = u[mi, int_j, :] * S_C
            u[gh_i, gj, :] = u[int_i, mj, :] * S_C
            u[gh_i, gh_j, :] = u[int_i, int_j, :] * S_C
    return u


DMR_POST_ROW = [8, 33, -57.1576, 563.499, 8, 4.125, -7.1447, 116.5]


def apply_bc_refac(q, bound, nx, ny):
    """Double-Mach ghost fill of the reference (unitest.py:207-227)."""
    for i in range(3):
        q[i, :, :] = DMR_POST_ROW
        q[:, ny + 3 + i, :] = q[:, ny + 2, :]
        q[nx + 3 + i, :, :] = q[nx + 2, :, :]
        q[:bound + 1, i, :] = DMR_POST_ROW
        q[bound + 1:bound + 10, i, :] = q[bound + 1:bound + 10, 3, :]
        q[bound + 10:, i, :] = q[bound + 10:, 5 - i, :]
        q[bound + 10:, i, 2] *= -1
        q[bound + 10:, i, 6] *= -1
    return q


def sync_primitives(q, gamma):
    """Rie_prob_2d_rs.py:157-163."""
    p, a, rho, u, v, h = con2primi_py(q, gamma)
    q[:, :, 4] = rho
    q[:, :, 5] = u
    q[:, :, 6] = v
    q[:, :, 7] = p
    return q


def rk3_cc(q, nx, ny, dx, dt, gamma, force_hlle, weno_switch, jp_cri=(0.2, 1.0)):
    """Shu-Osher SSP-RK3 with mirror ghost fill (Rie_prob_2d_rs.py:169-207)."""
    sl = np.s_[3:nx + 3, 3:ny + 3, :4]
    apply_bc_zero_gradient(q)
    l0 = l_local_py(q, dx, gamma, force_hlle, jp_cri, weno_switch)
    q1 = q.copy()
    q1[sl] = q[sl] + dt * l0
    apply_bc_zero_gradient(q1)
    sync_primitives(q1, gamma)
    apply_bc_zero_gradient(q1)
    l1 = l_local_py(q1, dx, gamma, force_hlle, jp_cri, weno_switch)
    q2 = q.copy()
    q2[sl] = 0.75 * q[sl] + 0.25 * q1[sl] + 0.25 * dt * l1
    apply_bc_zero_gradient(q2)
    sync_primitives(q2, gamma)
    apply_bc_zero_gradient(q2)
    l2 = l_local_py(q2, dx, gamma, force_hlle, jp_cri, weno_switch)
    qn = q.copy()
    qn[sl] = (1 / 3) * q[sl] + (2 / 3) * q2[sl] + (2 / 3) * dt * l2
    apply_bc_zero_gradient(qn)
    sync_primitives(qn, gamma)
    apply_bc_zero_gradient(qn)
    return qn


def rk3_refac(q, nx, ny, dx, dt, gamma, force_hlle, bound, jp_cri=(0.2, 1.0),
              cons_switch=False):
    """DMR SSP-RK3 (double_mach_reflection.py:97-134); 5-arg l_local calls
    select the characteristic path."""
    sl = np.s_[3:nx + 3, 3:ny + 3, :4]
    l0 = l_local_py(q, dx, gamma, force_hlle, jp_cri, cons_switch)
    q1 = q.copy()
    q1[sl] = q[sl] + dt * l0
    q1 = apply_bc_refac(q1, bound, nx, ny)
    l1 = l_local_py(q1, dx, gamma, force_hlle, jp_cri, cons_switch)
    q2 = q.copy()
    q2[sl] = 0.75 * q[sl] + 0.25 * q1[sl] + 0.25 * dt * l1
    q2 = apply_bc_refac(q2, bound, nx, ny)
    l2 = l_local_py(q2, dx, gamma, force_hlle, jp_cri, cons_switch)
    qn = q.copy()
    qn[sl] = 1 / 3 * q[sl] + 2 / 3 * q2[sl] + 2 / 3 * dt * l2
    qn = apply_bc_refac(qn, bound, nx, ny)
    p, a, rho, u, v, h = con2primi_py(qn, gamma)
    qn[:, :, 4] = rho
    qn[:, :, 5] = u
    qn[:, :, 6] = v
    qn[:, :, 7] = p
    return apply_bc_refac(qn, bound, nx, ny)


def apply_bc_copy(u, nx, ny):
    """Constant extrapolation (Rie_prob_1D_rs.py:59-76)."""
    for i in (0, 1, 2):
        u[i, :, :] = u[3, :, :]
    for i in (nx + 3, nx + 4, nx + 5):
        u[i, :, :] = u[nx + 2, :, :]
    for j in (0, 1, 2):
        u[:, j, :] = u[:, 3, :]
    for j in (ny + 3, ny + 4, ny + 5):
        u[:, j, :] = u[:, ny + 2, :]
    return u


def rk3_1d(q, nx, ny, dx, dt, gamma, force_hlle=True, jp_cri=(1, 5)):
    """Rie_prob_1D_rs.rk3 (:20-57): component-wise l_local, ghost fill after
    each stage, primitive sync at the end."""
    sl = np.s_[3:nx + 3, 3:ny + 3, :4]
    l0 = l_local_py(q, dx, gamma, force_hlle, jp_cri, True)
    q1 = q.copy()
    q1[sl] = q[sl] + dt * l0
    q1 = apply_bc_copy(q1, nx, ny)
    l1 = l_local_py(q1, dx, gamma, force_hlle, jp_cri, True)
    q2 = q.copy()
    q2[sl] = 0.75 * q[sl] + 0.25 * q1[sl] + 0.25 * dt * l1
    q2 = apply_bc_copy(q2, nx, ny)
    l2 = l_local_py(q2, dx, gamma, force_hlle, jp_cri, True)
    qn = q.copy()
    qn[sl] = 1 / 3 * q[sl] + 2 / 3 * q2[sl] + 2 / 3 * dt * l2
    qn = apply_bc_copy(qn, nx, ny)
    p, a, rho, u, v, h = con2primi_py(qn, gamma)
    qn[:, :, 4] = rho
    qn[:, :, 5] = u
    qn[:, :, 6] = v
    qn[:, :, 7] = p
    return apply_bc_copy(qn, nx, ny)


def run_1d(q, nx, ny, dx, gamma, cfl, nstep):
    """The time loop of Rie_prob_1D_rs.main (:159-175) without I/O."""
    T = 0.0
    for _ in range(1, nstep):
        speed = np.sqrt(q[3:nx + 3, 3:ny + 3, 5] ** 2 + q[3:nx + 3, 3:ny + 3, 6] ** 2)
        c_max = np.max(np.sqrt(gamma * q[3:nx + 3, 3:ny + 3, 7]
                               / np.maximum(q[3:nx + 3, 3:ny + 3, 0], 1e-10)) + speed)
        dt = cfl * dx / c_max
        q = rk3_1d(q, nx, ny, dx, dt, gamma)
        T += dt
    return q, T


def cfl_dt(q, nx, ny, gamma, cfl, dx, remaining=None):
    """Rie_prob_2d_rs.py:311-319 (and double_mach_reflection.py:233-240)."""
    speed = np.sqrt(q[3:3 + nx, 3:3 + ny, 5] ** 2 + q[3:3 + nx, 3:3 + ny, 6] ** 2)
    c = np.sqrt(gamma * q[3:3 + nx, 3:3 + ny, 7]
                / np.maximum(q[3:3 + nx, 3:3 + ny, 0], 1e-14))
    c_max = float(np.max(c + speed)) + 1e-16
    dt = float(cfl) * dx / c_max
    if remaining is not None:
        dt = min(dt, remaining)
    return dt


def run_dmr_reference(q, nx, ny, dx, bound, gamma, cfl, t_final, nstep, force_hlle=False,
                      jp_cri=(1e4, 1.1e4)):
    """Time-controlled loop of double_mach_reflection.main_rs (:229-260)
    without I/O: CFL dt clamped to the remaining time, rk3_refac, T += dt,
    while T < t_final - 1e-12 and step < nstep.  Returns (q, list of T)."""
    T, step, ts = 0.0, 0, [0.0]
    while T < t_final - 1e-12 and step < int(nstep):
        dt = cfl_dt(q, nx, ny, gamma, cfl, dx, remaining=t_final - T)
        q = rk3_refac(q, nx, ny, dx, dt, gamma, force_hlle, bound, jp_cri)
        T += dt
        step += 1
        ts.append(T)
    return q, ts


def run_riemann(q, nx, ny, dx, gamma, cfl, force_hlle, weno_switch, jp_cri,
                nstep=None, t_final=None):
    """Time loop of Rie_prob_2d_rs.main_rs (:328-365) without I/O."""
    T, step, ts = 0.0, 0, [0.0]
    max_steps = int(nstep) if nstep is not None else int(1e8)
    if t_final is not None:
        while T < t_final - 1e-12 and step < max_steps:
            dt = cfl_dt(q, nx, ny, gamma, cfl, dx, remaining=t_final - T)
            q = rk3_cc(q, nx, ny, dx, dt, gamma, force_hlle, weno_switch, jp_cri)
            T += dt
            step += 1
            ts.append(T)
    else:
        for _ in range(max_steps):
            dt = cfl_dt(q, nx, ny, gamma, cfl, dx)
            q = rk3_cc(q, nx, ny, dx, dt, gamma, force_hlle, weno_switch, jp_cri)
            T += dt
            step += 1
            ts.append(T)
    return q, ts


# --------------------------------------------------------------------------
# The NumPy twin's time loop (WENO_HLLC_2d.py:786-905)
# --------------------------------------------------------------------------
def twin_quadrants(n, gamma=1.4):
    """Initial data of WENO_HLLC_2d.main (:846-880): four constant quadrants
    split at padded index int((n+6)/2), ghost cells included, then the
    constant-extrapolation ghost fill."""
    u = np.zeros((n + 6, n + 6, 8))
    h = int((n + 6) / 2)
    for (si, sj), (r, vx, vy, p) in (((slice(None, h), slice(None, h)), (0.125, 1, 1, 0.125)),
                                     ((slice(h, None), slice(None, h)), (0.5, 0, 1, 0.25)),
                                     ((slice(None, h), slice(h, None)), (0.5, 1, 0, 0.25)),
                                     ((slice(h, None), slice(h, None)), (1, 0, 0, 1))):
        u[si, sj, 4], u[si, sj, 5], u[si, sj, 6], u[si, sj, 7] = r, vx, vy, p
    u[:, :, 0] = u[:, :, 4]
    u[:, :, 1] = u[:, :, 4] * u[:, :, 5]
    u[:, :, 2] = u[:, :, 4] * u[:, :, 6]
    u[:, :, 3] = u[:, :, 7] / (gamma - 1) + 0.5 * u[:, :, 4] * (u[:, :, 5] ** 2 + u[:, :, 6] ** 2)
    return apply_bc_copy(u, n, n)


def twin_pfloor(u, gamma, pmin=1e-8):
    """con_var2primi_pfloor (:813-837): raise E where p < pmin, then write
    the primitive channels (rho floored at 1e-12, u, v, p) on every cell."""
    rho = np.maximum(u[:, :, 0], 1e-12)
    uvel = u[:, :, 1] / rho
    vvel = u[:, :, 2] / rho
    E = u[:, :, 3]
    kinetic = 0.5 * rho * (uvel ** 2 + vvel ** 2)
    e_int = E - kinetic
    p = (gamma - 1) * e_int
    mask = p < pmin
    if np.any(mask):
        e_int[mask] = pmin / (gamma - 1)
        E[mask] = kinetic[mask] + e_int[mask]
        p[mask] = pmin
    u[:, :, 4] = rho
    u[:, :, 5] = uvel
    u[:, :, 6] = vvel
    u[:, :, 7] = p
    return u


def twin_rk3(u, n, dt, gamma, force_hlle, jp_cri):
    """RK3_correct (:786-811): component-wise L of the twin semantics, ghost
    copy after every stage; the caller holds the twin context."""
    dx = 1 / n
    sl = np.s_[3:n + 3, 3:n + 3, :4]
    l0 = l_local_py(u, dx, gamma, force_hlle, jp_cri, True)
    u1 = u.copy()
    u1[sl] = u[sl] + dt * l0
    u1 = apply_bc_copy(u1, n, n)
    l1 = l_local_py(u1, dx, gamma, force_hlle, jp_cri, True)
    u2 = u.copy()
    u2[sl] = 0.75 * u[sl] + 0.25 * u1[sl] + 0.25 * dt * l1
    u2 = apply_bc_copy(u2, n, n)
    l2 = l_local_py(u2, dx, gamma, force_hlle, jp_cri, True)
    un = u.copy()
    un[sl] = 1 / 3 * u[sl] + 2 / 3 * u2[sl] + 2 / 3 * dt * l2
    return apply_bc_copy(un, n, n)


def run_twin(u, n, n_step, force_hlle=False, jp_cri=(1.0, 4.0), gamma=1.4, cfl=0.2):
    """The loop of WENO_HLLC_2d.main (:895-905) without I/O: for i in
    range(1, n_step), dt = CFL dx / max(sqrt(g p / max(rho, 1e-10)) + |V|)
    from the primitive channels, RK3_correct, pressure floor.  Returns (u,
    list of T)."""
    dx = 1 / n
    T, ts = 0, [0.0]
    with twin():
        for _ in range(1, n_step):
            speed = np.sqrt(u[3:n + 3, 3:n + 3, 5] ** 2 + u[3:n + 3, 3:n + 3, 6] ** 2)
            c_max = np.max(np.sqrt(gamma * u[3:n + 3, 3:n + 3, 7] / np.maximum(u[3:n + 3, 3:n + 3, 0], 1e-10))
                           + speed)
            dt = cfl * dx / c_max
            u = twin_rk3(u, n, dt, gamma, force_hlle, jp_cri)
            u = twin_pfloor(u, gamma)
            T += dt
            ts.append(float(T))
    return u, ts
