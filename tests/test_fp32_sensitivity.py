"""How far FP32 arithmetic in the Riemann stage moves the reference scheme
(CPU, test instrumentation: the oracle built with every value in FP32).

The FP32-split path keeps the fluxes, the Riemann solvers and the sensor in
FP64 and runs only the reconstruction in FP32, on increments.  This test
records why: on the 64^2 four-quadrant problem (characteristic WENO5 + HLLC,
CFL 0.4) the oracle's RHS with only its HLLC/HLLE/blend stage in FP32 moves
the density by more than 1e-5 relative L1 within 40 steps (measured 4.3e-5),
forty times the 1e-6 the FP32-split precision is held to; the same RHS in
FP64 composed from the same stages equals l_local bit for bit, and the FP32
stage alone differs from the FP64 one by FP32 rounding (about 1e-7 relative
per call) -- the scheme amplifies it.
"""
import numpy as np

G = 1.4
JP = (1e4, 1.1e4)


def _rhs(oracle_mod, u, dx, riemann32):
    out = []
    for d in (0, 1):
        faces = oracle_mod._weno(u, G, d, False)
        if riemann32:
            out.append(oracle_mod.blend_f32(*faces, G, JP[0], JP[1], d))
        else:
            out.append(oracle_mod._blend(*faces, G, JP[0], JP[1], False, d))
    fx, fy = out
    return -((fx[1:] - fx[:-1]) / dx + (fy[:, 1:] - fy[:, :-1]) / dx)


def _run(oracle_mod, q, n, dx, steps, riemann32):
    """Rie_prob_2d_rs.rk3_cc's stage structure (the oracle's restatement)
    around the composed RHS."""
    sl = np.s_[3:n + 3, 3:n + 3, :4]
    bc = oracle_mod.apply_bc_zero_gradient
    for _ in range(steps):
        dt = oracle_mod.cfl_dt(q, n, n, G, 0.4, dx)
        bc(q)
        q1 = q.copy()
        q1[sl] = q[sl] + dt * _rhs(oracle_mod, q, dx, riemann32)
        bc(q1); oracle_mod.sync_primitives(q1, G); bc(q1)
        q2 = q.copy()
        q2[sl] = 0.75 * q[sl] + 0.25 * q1[sl] + 0.25 * dt * _rhs(oracle_mod, q1, dx, riemann32)
        bc(q2); oracle_mod.sync_primitives(q2, G); bc(q2)
        qn = q.copy()
        qn[sl] = (1 / 3) * q[sl] + (2 / 3) * q2[sl] + (2 / 3) * dt * _rhs(oracle_mod, q2, dx, riemann32)
        bc(qn); oracle_mod.sync_primitives(qn, G); bc(qn)
        q = qn
    return q


def test_fp32_riemann_stage(oracle_mod):
    from wenox import problems
    n = 64
    q0, dx = problems.four_quadrant(n)
    # the composition is the oracle's own l_local
    np.testing.assert_array_equal(_rhs(oracle_mod, q0, dx, False),
                                  oracle_mod.l_local_py(q0, dx, G, False, JP, False))
    # one call: FP32 rounding of the stage
    one = _rhs(oracle_mod, q0, dx, True)
    ref1 = _rhs(oracle_mod, q0, dx, False)
    assert 1e-9 < np.abs(one - ref1).max() / np.abs(ref1).max() < 1e-5
    # 40 steps: amplified far beyond the FP32-split bound
    ref = _run(oracle_mod, q0.copy(), n, dx, 40, False)
    got = _run(oracle_mod, q0.copy(), n, dx, 40, True)
    rel = np.abs(got[3:-3, 3:-3, 0] - ref[3:-3, 3:-3, 0]).sum() / np.abs(ref[3:-3, 3:-3, 0]).sum()
    print("FP32 Riemann stage, 64^2 four-quadrant, 40 steps: rel L1 density", rel)
    assert rel > 1e-5
