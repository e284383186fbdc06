"""Compensated global sums on the device, and conservation with periodic
ghost cells.

* ``Solver.conserved_sums`` (TwoSum-compensated FP64 sums of rho, rho u,
  rho v, E over the interior) equals the correctly rounded sum of the
  downloaded state (``math.fsum``) to one unit in the last place, for FP64
  and FP32-split states.
* With periodic ghost cells and HLLC everywhere (phi = 1) the scheme is in
  flux form, so the interior sums stay constant to round-off: relative to the
  sum of magnitudes, <= 1e-14 per conserved variable for FP64.  FP32-split
  stores its flux differences as FP32 numbers, whose sum telescopes only to
  FP32 rounding of each difference: <= 1e-7.
* Kelvin-Helmholtz at its full 8192^2 size (seeded noise, periodic,
  characteristic WENO5 + HLLC): finite, conservative, and FP32-split within
  1e-6 relative L1 density of FP64 after the same steps.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONS_TOL = {"fp64": 1e-14, "fp32split": 1e-7}


def _fsums(q):
    return np.array([math.fsum(q[3:-3, 3:-3, k].ravel().tolist()) for k in range(4)])


def _ulp_close(got, ref, mags, ulps=1.0):
    for k in range(4):
        tol = ulps * math.ulp(abs(ref[k])) + 1e-30 * mags[k]
        assert abs(got[k] - ref[k]) <= tol, (k, got[k], ref[k], got[k] - ref[k])


@pytest.mark.parametrize("precision,variant", [("fp64", "fast"), ("fp64", "exact"), ("fp32split", "fast")])
def test_conserved_sums_and_periodic_conservation(precision, variant):
    from wenox import Solver, problems
    q0, dx = problems.kelvin_helmholtz(96)
    print(f"[sums] KH 96^2 {precision} {variant}", flush=True)
    with Solver(96, 96, dx, cfl=0.4, jp_cri=(1e4, 1.1e4), bc="periodic", engine="fused",
                variant=variant, precision=precision) as s:
        s.upload(q0)
        m0 = s.conserved_sums(magnitudes=True)
        _ulp_close(m0[:4], _fsums(s.download()), m0[4:])
        s.run(nstep=40)
        m1 = s.conserved_sums(magnitudes=True)
        q1 = s.download()
    assert np.isfinite(q1).all()
    _ulp_close(m1[:4], _fsums(q1), m1[4:])
    drift = np.abs(m1[:4] - m0[:4]) / m0[4:]
    print(f"[sums] relative drift after 40 steps {drift}", flush=True)
    assert np.all(drift <= CONS_TOL[precision]), drift


def test_conserved_sums_argument_checks():
    from wenox import Solver, _lib, problems
    import ctypes
    q0, dx = problems.kelvin_helmholtz(16)
    with Solver(16, 16, dx, bc="periodic", engine="fused") as s:
        s.upload(q0)
        out = np.zeros(8)
        rc = _lib.load().wenox_solver_conserved_sums(s._h, out.ctypes.data, 5)
        assert rc == -1
        assert b"n_out" in _lib.load().wenox_last_error()
        assert np.all(out == 0)
        # the unfused engine shares the state layout
    with Solver(16, 16, dx, bc="periodic", engine="unfused") as s:
        s.upload(q0)
        _ulp_close(s.conserved_sums(), _fsums(s.download()), s.conserved_sums(True)[4:])
    del ctypes


def test_kelvin_helmholtz_8192():
    """The bandwidth-bound large grid at its full size: 8192^2 periodic
    Kelvin-Helmholtz (seeded white-noise v of amplitude 0.01), characteristic
    WENO5 + HLLC, 20 SSP-RK3 steps in FP64 and in FP32-split."""
    from wenox import Solver, problems
    n, steps = 8192, 20
    q0, dx = problems.kelvin_helmholtz(n)
    out = {}
    for precision in ("fp64", "fp32split"):
        print(f"[kh8192] {precision} {steps} steps", flush=True)
        with Solver(n, n, dx, cfl=0.4, jp_cri=(1e4, 1.1e4), bc="periodic", engine="fused",
                    variant="fast", precision=precision) as s:
            s.upload(q0)
            m0 = s.conserved_sums(magnitudes=True)
            T, k = s.run(nstep=steps)
            m1 = s.conserved_sums(magnitudes=True)
            rho = s.download()[3:-3, 3:-3, 0]
        assert k == steps and np.isfinite(rho).all()
        drift = np.abs(m1[:4] - m0[:4]) / m0[4:]
        print(f"[kh8192] {precision} T={T:.6g} relative drift {drift}", flush=True)
        assert np.all(drift <= CONS_TOL[precision]), drift
        # the density sum against an extended-precision pairwise sum of the download
        ref = float(rho.astype(np.longdouble).sum())
        assert abs(m1[0] - ref) <= 2 * math.ulp(ref), (m1[0], ref)
        out[precision] = (rho, T)
    (r64, t64), (r32, t32) = out["fp64"], out["fp32split"]
    rel = float(np.abs(r32 - r64).sum() / np.abs(r64).sum())
    print(f"[kh8192] FP32-split rel L1 density vs FP64 {rel:.3g}; T {t64} / {t32}", flush=True)
    assert rel <= 1e-6, rel
