"""GPU parity of the device-resident solver against the oracle's time loop.

Criterion: relative L1 = sum|U_gpu - U_ref| / sum|U_ref| per conserved
variable over the interior, <= 1e-11 after the stated number of SSP-RK3
steps (the "exact" variant is expected to be bitwise equal; the bound is the
contract).  The oracle runs the reference drivers' numpy time loop around the
C restatement of the Rust RHS.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-11


def rel_l1(a, b, nx, ny):
    out = []
    for k in range(4):
        d = np.abs(a[3:nx + 3, 3:ny + 3, k] - b[3:nx + 3, 3:ny + 3, k]).sum()
        s = np.abs(b[3:nx + 3, 3:ny + 3, k]).sum()
        out.append(d / s)
    return np.array(out)


def run_pair(oracle_mod, q0, nx, ny, dx, steps, cons, force_hlle, jp, engine, variant, cfl=0.4,
             t_final=None):
    from wenox import Solver
    ref, ts = oracle_mod.run_riemann(q0.copy(), nx, ny, dx, 1.4, cfl, force_hlle, cons, jp,
                                     nstep=steps, t_final=t_final)
    with Solver(nx, ny, dx, gamma=1.4, cfl=cfl, cons_switch=cons, force_hlle=force_hlle,
                jp_cri=jp, bc="mirror", engine=engine, variant=variant) as s:
        s.upload(q0)
        T, k = s.run(nstep=steps, t_final=t_final)
        got = s.download()
    return ref, ts, got, T, k


@pytest.mark.parametrize("engine", ["unfused", "fused"])
@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_four_quadrant_50_steps(oracle_mod, engine, variant):
    from wenox import problems
    n = 128
    q0, dx = problems.four_quadrant(n)
    print(f"[solver] four-quadrant {n}^2 char+HLLC 50 steps engine={engine} variant={variant}", flush=True)
    ref, ts, got, T, k = run_pair(oracle_mod, q0, n, n, dx, 50, False, False, (1e4, 1.1e4), engine, variant)
    assert k == 50
    if variant == "exact":
        assert T == ts[-1]
    assert abs(T - ts[-1]) <= TOL * ts[-1]
    err = rel_l1(got, ref, n, n)
    print("rel L1", err)
    assert np.all(err <= TOL), err
    if variant == "exact":
        np.testing.assert_array_equal(got[..., :4], ref[..., :4])
    # primitives of the download match the reference's synced channels (the
    # fast variant's pointwise differences near the shocks reach ~1e-8)
    rtol = 1e-12 if variant == "exact" else 1e-6
    np.testing.assert_allclose(got[3:-3, 3:-3, 4:], ref[3:-3, 3:-3, 4:], rtol=rtol, atol=1e-12)


@pytest.mark.parametrize("engine", ["unfused", "fused"])
def test_componentwise_blend(oracle_mod, engine):
    from wenox import problems
    q0 = problems.riemann_reference(64, 64)
    print(f"[solver] reference IC 64^2 cons+blend 20 steps engine={engine}", flush=True)
    ref, ts, got, T, k = run_pair(oracle_mod, q0, 64, 64, 1 / 64, 20, True, False, (0.5, 2.0),
                                  engine, "exact", cfl=0.3)
    err = rel_l1(got, ref, 64, 64)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("engine", ["unfused", "fused"])
def test_t_final_loop(oracle_mod, engine):
    """The time-controlled loop: same number of steps and final time as
    Rie_prob_2d_rs.main_rs (:335-344), the last dt clipped to t_final - T."""
    from wenox import problems
    q0, dx = problems.four_quadrant(48)
    print(f"[solver] t_final loop 48^2 engine={engine}", flush=True)
    ref, ts, got, T, k = run_pair(oracle_mod, q0, 48, 48, dx, None, False, True, (0.2, 1.0),
                                  engine, "exact", t_final=0.05)
    assert k == len(ts) - 1
    assert T == ts[-1]
    err = rel_l1(got, ref, 48, 48)
    assert np.all(err <= TOL), err


def test_dmr_reference_bc(oracle_mod):
    from wenox import Solver, problems
    nx, ny = 80, 20
    q0, bound = problems.dmr_reference(nx, ny)
    dx = 1.0 / ny
    print(f"[solver] DMR reference BC {nx}x{ny} bound={bound}", flush=True)
    q = q0.copy()
    T = 0.0
    for _ in range(8):
        dt = oracle_mod.cfl_dt(q, nx, ny, 1.4, 0.1, dx, remaining=0.2 - T)
        q = oracle_mod.rk3_refac(q, nx, ny, dx, dt, 1.4, False, bound, (10000, 11000))
        T += dt
    with Solver(nx, ny, dx, gamma=1.4, cfl=0.1, cons_switch=False, force_hlle=False,
                jp_cri=(10000, 11000), bc="dmr_ref", dmr_bound=bound) as s:
        s.upload(q0)
        T2, k = s.run(nstep=8, t_final=0.2)
        got = s.download()
    assert k == 8 and T2 == T
    err = rel_l1(got, q, nx, ny)
    assert np.all(err <= TOL), err
    np.testing.assert_array_equal(got, q)


def test_solver_rhs_matches_l_local(oracle_mod):
    from wenox import Solver, problems
    q0, dx = problems.four_quadrant(40)
    ref = oracle_mod.l_local_py(q0, dx, 1.4, False, (0.5, 2.0), False)
    for engine in ("unfused", "fused"):
        with Solver(40, 40, dx, cons_switch=False, force_hlle=False, jp_cri=(0.5, 2.0),
                    engine=engine) as s:
            s.upload(q0)
            got = s.rhs()
        np.testing.assert_array_equal(got, ref, err_msg=engine)


def test_fp32_split_accuracy(oracle_mod):
    """FP32-split (FP32 WENO arithmetic on increments, hi+lo state, FP64
    fluxes, Riemann solver, combine and CFL; faces with an inadmissible FP32
    Gauss-point state, or whose stencils span a large relative jump,
    recomputed in FP64) after 50 steps on the coarse 128^2 four-quadrant
    grid: the same state bit for bit for every sweep tiling (the FP64
    fallback is chosen per face), and relative L1 density error against
    the FP64 oracle <= 1e-5 (measured 1.6e-6; 3.1e-6 without the jump test).  This grid is the hard case:
    the reference scheme itself turns it non-finite at step 293, and
    perturbations of FP32 size grow by about 50x within 50 steps (an FP32
    Riemann stage alone, on the oracle, gives 1.6e-5: test_fp32_sensitivity);
    the 4096^2 grid and the 1920x480 double Mach reflection of the benchmark
    stay near 1e-7 and 1e-9."""
    from wenox import Solver, problems
    n = 128
    q0, dx = problems.four_quadrant(n)
    print(f"[solver] fp32-split four-quadrant {n}^2 50 steps", flush=True)
    ref, ts = oracle_mod.run_riemann(q0.copy(), n, n, dx, 1.4, 0.4, False, False, (1e4, 1.1e4), nstep=50)
    outs = []
    for tiling in ((8, 8, 1, 1), (48, 48, 1, 1), (10, 9, 4, 4), (13, 7, 2, 1)):
        with Solver(n, n, dx, cfl=0.4, jp_cri=(1e4, 1.1e4), engine="fused", variant="fast",
                    precision="fp32split") as s:
            s.set_tiling(*tiling)
            s.upload(q0)
            T, k = s.run(nstep=50)
            assert k == 50
            outs.append(s.download())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    err = rel_l1(outs[0], ref, n, n)
    print("fp32-split rel L1 per variable", err)
    assert err[0] <= 1e-5, err


@pytest.mark.parametrize("cons", [False, True])
def test_fp32_split_every_tiling_with_blend(cons):
    """With the HLLC/HLLE blend active (the sensor reads p̄ of the faces
    beside each face, including the wave's edge lanes) the FP32-split step
    still gives the same state bit for bit for every tiling: the FP64
    fallback is chosen per face from that face's own states."""
    from wenox import Solver, problems
    n = 96
    q0, dx = problems.four_quadrant(n)
    outs = []
    for tiling in ((8, 8, 1, 1), (40, 40, 1, 1), (10, 9, 4, 4), (13, 7, 2, 1)):
        with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), cons_switch=cons, engine="fused", variant="fast",
                    precision="fp32split") as s:
            s.set_tiling(*tiling)
            s.upload(q0)
            s.run(nstep=30)
            outs.append(s.download())
    assert np.isfinite(outs[0]).all()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_fp32_split_rhs(oracle_mod):
    from wenox import Solver, problems
    q0, dx = problems.four_quadrant(64)
    ref = oracle_mod.l_local_py(q0, dx, 1.4, False, (0.5, 2.0), False)
    with Solver(64, 64, dx, jp_cri=(0.5, 2.0), engine="fused", variant="fast", precision="fp32split") as s:
        s.upload(q0)
        got = s.rhs()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-4 * scale


def test_fp32_split_jump_threshold(oracle_mod):
    """The FP32-split split-jump threshold: with a tiny threshold every face
    whose stencils vary takes FP64 states, and the RHS is the oracle's up to
    the FP32 rounding of the stored flux differences; with the test off (inf)
    the FP32 reconstruction of the four-quadrant jumps shows; non-positive
    and NaN thresholds are rejected."""
    from wenox import Solver, problems
    q0, dx = problems.four_quadrant(64)
    ref = oracle_mod.l_local_py(q0, dx, 1.4, False, (0.5, 2.0), False)
    scale = np.abs(ref).max()
    err = {}
    with Solver(64, 64, dx, jp_cri=(0.5, 2.0), engine="fused", variant="fast", precision="fp32split") as s:
        s.upload(q0)
        for tau in (1e-30, float("inf"), 0.25):
            s.set_split_jump(tau)
            assert s.split_jump == np.float32(tau)
            err[tau] = np.abs(s.rhs() - ref).max() / scale
        for bad in (0.0, -1.0, float("nan")):
            with pytest.raises(ValueError, match="tau must be > 0"):
                s.set_split_jump(bad)
        assert s.split_jump == 0.25
    print(f"[solver] fp32-split rhs max rel error by split-jump threshold {err}", flush=True)
    assert err[1e-30] <= 1e-7, err
    assert err[float("inf")] <= 1e-4 and err[float("inf")] > err[1e-30], err


@pytest.mark.parametrize("cons,force_hlle", [(False, False), (True, False), (False, True)])
def test_sweep_tilings_bitwise(oracle_mod, cons, force_hlle):
    """Every sweep tiling -- 1, 2 or 4 wavefronts per workgroup sharing the
    sensor halo, any rows per wavefront, grids that are not a multiple of the
    group -- gives the oracle's RHS bit for bit (exact variant)."""
    from wenox import Solver, problems
    nx, ny = 45, 37
    q0, dx = problems.four_quadrant(nx, ny=ny)
    ref = oracle_mod.l_local_py(q0, dx, 1.4, force_hlle, (0.5, 2.0), cons)
    assert np.isfinite(ref).all()   # (NaN would compare equal to NaN below)
    print(f"[solver] tilings 45x37 cons={cons} hlle={force_hlle}", flush=True)
    with Solver(nx, ny, dx, cons_switch=cons, force_hlle=force_hlle, jp_cri=(0.5, 2.0), engine="fused") as s:
        s.upload(q0)
        for wx, cx in ((1, 8), (1, 45), (2, 3), (2, 7), (4, 3), (4, 4), (4, 11)):
            for wy, cy in ((1, 5), (4, 3), (2, 10)):
                s.set_tiling(cx, cy, wx, wy)
                assert s.tiling == (cx, cy, wx, wy)
                np.testing.assert_array_equal(s.rhs(), ref, err_msg=f"tiling {(cx, cy, wx, wy)}")


@pytest.mark.parametrize("blocked", [False, True])
@pytest.mark.parametrize("cons,force_hlle", [(False, False), (True, False), (False, True)])
def test_fast_step_same_for_every_tiling(oracle_mod, cons, force_hlle, blocked):
    """The fast FP64 step lets the x sweep form its part of each Shu-Osher
    stage, a Un + b Us - (c dt) Dx, inside the march or -- for the cells next
    to the faces a workgroup's wavefronts share -- after it, from reloaded
    states.  Its state after 6 steps is the same bit for bit for every tiling
    (x tilings with 1, 2 and 4 wavefronts, with both merged and separate
    launches), and within 1e-12 relative L1 of the oracle's time loop; a
    wrong coefficient or a cell finished with another cell's state would
    give O(dt) differences."""
    from wenox import Solver, problems
    nx, ny = 45, 37
    q0, dx = problems.four_quadrant(nx, ny=ny)
    ref, ts = oracle_mod.run_riemann(q0.copy(), nx, ny, dx, 1.4, 0.4, force_hlle, cons, (0.5, 2.0), nstep=6)
    assert np.isfinite(ref).all()
    print(f"[solver] fast step, tilings 45x37 cons={cons} hlle={force_hlle} blocked={blocked}", flush=True)
    first = None
    for wx, cx in ((1, 8), (2, 3), (2, 7), (4, 3), (4, 4), (4, 11)):
        for wy, cy in ((1, 5), (4, 3)):
            with Solver(nx, ny, dx, cfl=0.4, cons_switch=cons, force_hlle=force_hlle, jp_cri=(0.5, 2.0),
                        engine="fused", variant="fast") as s:
                s.set_blocked(blocked)
                s.set_tiling(cx, cy, wx, wy)
                s.upload(q0)
                T, k = s.run(nstep=6)
                got = s.download()[..., :4]
            assert k == 6 and T == pytest.approx(ts[6], rel=1e-12)
            if first is None:
                first = got
                err = rel_l1(got, ref, nx, ny)
                assert max(err) <= 1e-12, err
            else:
                np.testing.assert_array_equal(got, first, err_msg=f"tiling {(cx, cy, wx, wy)}")


def test_tiling_arguments_checked():
    from wenox import Solver, problems
    from wenox._lib import WenoxError
    q0, dx = problems.four_quadrant(16)
    with Solver(16, 16, dx, engine="fused") as s:
        for bad in ((2, 8, 4, 1), (8, 8, 3, 1), (0, 8, 1, 1)):
            with pytest.raises((ValueError, WenoxError)):
                s.set_tiling(*bad)


def test_combine_tile_widths_bitwise(oracle_mod):
    """The RK combine with 32- and 64-wide tiles (grids that are not a
    multiple of either) gives the oracle's 5-step trajectory bit for bit."""
    from wenox import Solver, problems
    nx, ny = 70, 45
    q0, dx = problems.four_quadrant(nx, ny=ny)
    ref, _ = oracle_mod.run_riemann(q0.copy(), nx, ny, dx, 1.4, 0.4, False, False, (0.5, 2.0), nstep=5)
    assert np.isfinite(ref).all()
    print("[solver] combine tile widths 70x45", flush=True)
    for cols in (32, 64):
        with Solver(nx, ny, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="fused") as s:
            s.set_combine_cols(cols)
            assert s.combine_cols == cols
            s.upload(q0)
            s.run(nstep=5)
            np.testing.assert_array_equal(s.download()[..., :4], ref[..., :4], err_msg=f"combine cols {cols}")


def test_fused_step_control_across_run_calls(oracle_mod):
    """The fused engine forms the next step's CFL maxima in the last RK
    combine and applies a step's time advance at the next step's start (or
    when run() returns).  Whatever the split of the steps over run() calls
    -- one call, one step per call (the snapshot drivers), a t_final run,
    multi-step graph replays -- T, the step count and the state equal the
    oracle's bit for bit (exact variant); so does a CFL rule changed after
    upload."""
    from wenox import Solver, problems
    n, steps = 40, 19
    q0, dx = problems.four_quadrant(n)
    ref, ts = oracle_mod.run_riemann(q0.copy(), n, n, dx, 1.4, 0.4, False, False, (0.5, 2.0), nstep=steps)
    print("[solver] fused step control: one call, per-step calls, t_final, cfl rule", flush=True)
    for split in ([steps], [1] * steps, [8, 3, 8], [17, 2]):
        with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="fused") as s:
            s.upload(q0)
            total = 0
            for k in split:
                T, got = s.run(nstep=k)
                total += got
                assert T == ts[total], (split, total)
            assert total == steps
            np.testing.assert_array_equal(s.download()[..., :4], ref[..., :4], err_msg=str(split))
    t_end = ts[12] + 0.5 * (ts[13] - ts[12])
    ref_t, ts_t = oracle_mod.run_riemann(q0.copy(), n, n, dx, 1.4, 0.4, False, False, (0.5, 2.0), t_final=t_end)
    with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="fused") as s:
        s.upload(q0)
        T, k = s.run(t_final=t_end)
        assert k == len(ts_t) - 1 and T == ts_t[-1]
        np.testing.assert_array_equal(s.download()[..., :4], ref_t[..., :4])
        # stepping on after the end changes nothing
        T2, k2 = s.run(nstep=3, t_final=t_end)
        assert k2 == 0 and T2 == T
    # a CFL rule set after upload (the 1-D drivers' floor and no epsilon)
    with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="fused") as s:
        s.upload(q0[..., :4].copy())
        s.set_cfl_rule(0.3, 0.0)
        T, k = s.run(nstep=4)
    with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="unfused") as s:
        s.upload(q0[..., :4].copy())
        s.set_cfl_rule(0.3, 0.0)
        T_ref, _ = s.run(nstep=4)
    assert T == T_ref


@pytest.mark.parametrize("bc", ["mirror", "periodic", "copy", "dmr_ref", "dmr"])
def test_fused_matches_unfused_for_every_boundary(bc):
    """The fused engine (sweeps on the state and its transposed copy, ghost
    fill of both orientations, CFL maxima from the last RK combine) and the
    unfused one (the reference's pass structure) agree bit for bit (exact
    variant) after 7 steps for every boundary kind: the full downloaded
    arrays, ghost cells included, the time, and rhs() of the final state."""
    from wenox import Solver, problems
    kw = dict(cfl=0.3, jp_cri=(0.5, 2.0), bc=bc, variant="exact")
    if bc == "dmr_ref":
        nx, ny = 80, 20
        q0, bound = problems.dmr_reference(nx, ny)
        dx = 1.0 / ny
        kw.update(dmr_bound=bound, cfl=0.1, jp_cri=(10000, 11000))
    elif bc == "dmr":
        nx, ny = 96, 24
        q0, dx = problems.double_mach(nx, ny)
        kw.update(cfl=0.1)
    else:
        nx, ny = 45, 37
        q0, dx = problems.four_quadrant(nx, ny=ny)
    print(f"[solver] fused vs unfused, bc={bc} {nx}x{ny}", flush=True)
    out = {}
    for engine in ("unfused", "fused"):
        with Solver(nx, ny, dx, engine=engine, **kw) as s:
            s.upload(q0)
            T, k = s.run(nstep=7)
            assert k == 7
            out[engine] = (s.download(), T, s.rhs())
    assert np.isfinite(out["unfused"][0]).all() and np.isfinite(out["unfused"][2]).all()
    assert out["fused"][1] == out["unfused"][1]
    np.testing.assert_array_equal(out["fused"][0], out["unfused"][0])
    np.testing.assert_array_equal(out["fused"][2], out["unfused"][2])


GHOST_GRIDS = [(33, 40, 32), (40, 34, 32), (66, 65, 64), (64, 35, 32), (35, 64, 64)]
# double Mach reflection grids are 4:1; their edge tiles: 33 -> 1 cell (32-wide
# tiles), 34 -> 2 cells, 65 -> 1 cell (64-wide), 132 x 33 -> 4 x 1, 140 x 35 -> 12 x 3
DMR_GHOST_GRIDS = [(132, 33, 32), (136, 34, 32), (260, 65, 64), (140, 35, 32), (136, 34, 64)]


@pytest.mark.parametrize("bc", ["mirror", "periodic", "copy", "dmr_ref", "dmr"])
@pytest.mark.parametrize("grid", range(5))
def test_fused_ghost_cells_for_every_edge_tile_width(bc, grid):
    """The fused FP64 step lets the RK combine write the ghost cells where
    each ghost's source lies in the combine tile of its nearest interior
    cell, and keeps the ghost-fill kernel where an edge tile is 1 or 2 cells
    wide (33, 34, 65 wide grids) or the rule is periodic.  Either way the full
    arrays, ghost cells included, equal the unfused engine's bit for bit
    after 5 steps (exact variant), for every boundary rule, with 32- and
    64-wide combine tiles."""
    from wenox import Solver, problems
    kw = dict(cfl=0.3, jp_cri=(0.5, 2.0))
    if bc in ("dmr", "dmr_ref"):
        nx, ny, cols = DMR_GHOST_GRIDS[grid]
        if bc == "dmr":
            q0, dx = problems.double_mach(nx, ny)
        else:
            q0, bound = problems.dmr_reference(nx, ny)
            dx = 1.0 / ny
            kw.update(dmr_bound=bound)
        kw.update(cfl=0.1, jp_cri=(1e4, 1.1e4))
    else:
        nx, ny, cols = GHOST_GRIDS[grid]
        q0, dx = problems.four_quadrant(nx, ny=ny)
    print(f"[solver] ghost cells {bc} {nx}x{ny} combine cols {cols}", flush=True)
    out = {}
    for engine in ("unfused", "fused"):
        with Solver(nx, ny, dx, bc=bc, engine=engine, variant="exact", **kw) as s:
            if engine == "fused":
                s.set_combine_cols(cols)
            s.upload(q0)
            T, k = s.run(nstep=5)
            out[engine] = (s.download(), T)
    assert np.isfinite(out["unfused"][0]).all()
    assert out["fused"][1] == out["unfused"][1]
    np.testing.assert_array_equal(out["fused"][0], out["unfused"][0], err_msg=bc)


# (nx, ny, tiling): the blocked y sweep's tiles are 58 x rows on the lanes and
# groups of the march along y (chunk_y, or waves_y * chunk_y - 3 with the
# blend); an edge tile of 1 or 2 cells leaves the ghost fill to its kernel
BLK_GHOST_CASES = [
    (61, 40, (16, 12, 1, 1), True),    # x edge 3 cells, y edge 4
    (59, 40, (16, 12, 1, 1), False),   # x edge 1 cell
    (60, 40, (16, 12, 1, 1), False),   # x edge 2 cells
    (70, 40, (16, 13, 1, 1), False),   # y edge 1 cell
    (70, 41, (16, 8, 2, 2), False),    # y groups of 13: edge 2 cells
    (70, 42, (16, 8, 2, 2), True),     # y groups of 13: edge 3 cells
    (130, 37, (12, 6, 4, 4), True),    # x edges 14, y groups of 21: edge 16
]
BLK_GHOST_DMR_CASES = [
    (136, 34, (16, 10, 1, 1), True),
    (136, 34, (16, 16, 1, 1), False),  # y edge 2 cells
    (120, 30, (12, 9, 4, 4), True),    # y groups of 33: one group of 30
    (60, 15, (8, 8, 1, 1), False),     # x edge 2 cells
]


@pytest.mark.parametrize("bc", ["mirror", "periodic", "copy", "dmr_ref", "dmr"])
@pytest.mark.parametrize("case", range(4))
def test_blocked_y_sweep_ghost_cells(bc, case):
    """The blocked FP64 step's y sweep writes the ghost cells of the state it
    forms where each ghost's source lies in the sweep tile of its nearest
    interior cell (local rules, edge tiles of 3 cells or more along x and y),
    and the ghost-fill kernel runs otherwise (profile()["ghost_fill"] > 0):
    either way the full arrays, ghost cells included, and the times equal the
    row-major step's bit for bit after 6 steps, for every boundary rule."""
    from wenox import Solver, problems
    dmr = bc in ("dmr", "dmr_ref")
    cases = BLK_GHOST_DMR_CASES if dmr else BLK_GHOST_CASES
    for nx, ny, tiling, in_sweep in cases[case::4]:
        kw = dict(cfl=0.3, jp_cri=(0.5, 2.0))
        if bc == "dmr":
            q0, dx = problems.double_mach(nx, ny)
            kw.update(cfl=0.1, jp_cri=(1e4, 1.1e4))
        elif bc == "dmr_ref":
            q0, bound = problems.dmr_reference(nx, ny)
            dx = 1.0 / ny
            kw.update(cfl=0.1, jp_cri=(1e4, 1.1e4), dmr_bound=bound)
        else:
            q0, dx = problems.four_quadrant(nx, ny=ny)
        print(f"[solver] blocked ghost cells {bc} {nx}x{ny} tiling {tiling}", flush=True)
        out = {}
        for blocked in (False, True):
            with Solver(nx, ny, dx, bc=bc, engine="fused", variant="fast", **kw) as s:
                s.set_blocked(blocked)
                s.set_tiling(*tiling)
                s.upload(q0)
                if blocked:
                    # profile() leaves the state and the step control as they were
                    fill_ms = s.profile(reps=1)["ghost_fill"]
                    assert (fill_ms == 0.0) == (in_sweep and bc != "periodic"), (fill_ms, in_sweep)
                T, k = s.run(nstep=6)
                out[blocked] = (s.download(), T)
        assert np.isfinite(out[False][0]).all()
        assert out[True][1] == out[False][1]
        np.testing.assert_array_equal(out[True][0], out[False][0], err_msg=f"{bc} {nx}x{ny} {tiling}")


@pytest.mark.parametrize("bc", ["mirror", "copy", "dmr", "dmr_ref"])
def test_fp32_split_ghost_cells_on_store(bc):
    """The FP32-split step's combine writes the ghost cells too where
    ghost_on_store allows it.  On a 40 x 34 grid (136 x 34 for the double
    Mach reflection rules) 32-wide combine tiles leave a 2-cell edge tile (the
    fill kernel runs) and 64-wide tiles do not (the combine writes the
    ghosts); the tile width changes nothing else, so the two runs agree bit
    for bit, ghost cells included."""
    from wenox import Solver, problems
    nx, ny = 40, 34
    if bc == "dmr":
        nx = 4 * ny
        q0, dx = problems.double_mach(nx, ny)
        kw = dict(cfl=0.1, jp_cri=(1e4, 1.1e4))
    elif bc == "dmr_ref":
        nx = 4 * ny
        q0, bound = problems.dmr_reference(nx, ny)
        dx = 1.0 / ny
        kw = dict(cfl=0.1, jp_cri=(1e4, 1.1e4), dmr_bound=bound)
    else:
        q0, dx = problems.four_quadrant(nx, ny=ny)
        kw = dict(cfl=0.3, jp_cri=(0.5, 2.0))
    print(f"[solver] fp32-split ghost cells {nx}x{ny} bc={bc}", flush=True)
    out = {}
    for cols in (32, 64):
        with Solver(nx, ny, dx, bc=bc, engine="fused", variant="fast", precision="fp32split", **kw) as s:
            s.set_combine_cols(cols)
            s.upload(q0)
            T, k = s.run(nstep=6)
            out[cols] = (s.download(), T)
    assert np.isfinite(out[32][0]).all()
    assert out[32][1] == out[64][1]
    np.testing.assert_array_equal(out[64][0], out[32][0])


def test_fp32_split_step_control_across_run_calls():
    """The FP32-split step takes dt from the previous step's last RK combine
    (or the upload) and applies its time advance at the next step's start or
    when run() returns, like FP64: one call, one step per call, graph-sized
    and odd splits give bitwise the same state, time and step count, and a
    t_final run stops exactly at t_final with the last dt clipped."""
    from wenox import Solver, problems
    n, steps = 48, 13
    q0, dx = problems.four_quadrant(n)
    print("[solver] fp32-split step control across run calls", flush=True)
    res = []
    for split in ([steps], [1] * steps, [8, 5], [3, 9, 1]):
        with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="fused", variant="fast",
                    precision="fp32split") as s:
            s.upload(q0)
            total, T = 0, 0.0
            for k in split:
                T, got = s.run(nstep=k)
                total += got
            assert total == steps
            res.append((T, s.download()))
    for T, q in res[1:]:
        assert T == res[0][0]
        np.testing.assert_array_equal(q, res[0][1])
    t_end = 0.5 * res[0][0]
    with Solver(n, n, dx, cfl=0.4, jp_cri=(0.5, 2.0), engine="fused", variant="fast", precision="fp32split") as s:
        s.upload(q0)
        T, k = s.run(t_final=t_end)
        assert T == pytest.approx(t_end, rel=0, abs=1e-15) and 0 < k < steps
        T2, k2 = s.run(nstep=3, t_final=t_end)
        assert k2 == 0 and T2 == T


@pytest.mark.parametrize("bc,nx,ny", [("mirror", 45, 37), ("mirror", 130, 96), ("periodic", 96, 70),
                                      ("dmr", 136, 34)])
def test_pipeline_parts_bitwise(bc, nx, ny):
    """The slab-pipelined FP64 stage (sweeps of slab k on the solver's
    stream, the RK combines of the finished slabs on a second stream) gives
    the same state, ghost cells included, and the same time as the plain
    stage for every slab count, over steps whose third stage writes the state
    in place."""
    from wenox import Solver, problems
    if bc == "dmr":
        q0, dx = problems.double_mach(nx, ny)
    elif bc == "periodic":
        q0, dx = problems.kelvin_helmholtz(nx)
        q0 = q0[:, :ny + 6].copy()
    else:
        q0, dx = problems.four_quadrant(nx, ny=ny)
    print(f"[pipeline] {bc} {nx}x{ny}", flush=True)
    out = {}
    for parts in (1, 2, 3, 5, 8):
        with Solver(nx, ny, dx, cfl=0.4 if bc != "dmr" else 0.1, jp_cri=(0.5, 2.0), bc=bc, engine="fused",
                    variant="fast") as s:
            s.set_tiling(7, 9, 1, 1)
            s.set_blocked(False)   # the pipeline stages the row-major step's combine
            s.set_pipeline(parts)
            assert s.pipeline == parts
            s.upload(q0)
            T, k = s.run(nstep=12)
            out[parts] = (s.download(), T, k)
    q1, T1, k1 = out[1]
    assert np.isfinite(q1).all()
    for parts, (q, T, k) in out.items():
        assert k == k1 and T == T1, (parts, T, T1)
        np.testing.assert_array_equal(q, q1, err_msg=f"parts {parts}")


def _layout_case(bc):
    from wenox import problems
    kw = dict(cfl=0.4, jp_cri=(0.5, 2.0), bc=bc, engine="fused", variant="fast")
    if bc == "dmr_ref":
        nx, ny = 80, 20
        q0, bound = problems.dmr_reference(nx, ny)
        dx = 1.0 / ny
        kw.update(dmr_bound=bound, cfl=0.1, jp_cri=(10000, 11000))
    elif bc == "dmr":
        nx, ny = 136, 34
        q0, dx = problems.double_mach(nx, ny)
        kw.update(cfl=0.1)
    elif bc == "periodic":
        nx, ny = 96, 70
        q0, dx = problems.kelvin_helmholtz(nx)
        q0 = q0[:, :ny + 6].copy()
    else:
        nx, ny = 45, 37
        q0, dx = problems.four_quadrant(nx, ny=ny)
    return nx, ny, dx, q0, kw


@pytest.mark.parametrize("bc", ["mirror", "periodic", "copy", "dmr_ref", "dmr"])
@pytest.mark.parametrize("cons,force_hlle", [(False, False), (True, False), (False, True)])
def test_blocked_layout_same_states(bc, cons, force_hlle):
    """The blocked FP64 fast step (both sweeps on one blocked copy of the
    state, the y sweep finishing each stage, the CFL maxima from its stage-3
    workgroups) gives the same states as the row-major step bit for bit:
    the downloaded arrays (ghost cells included), the time and the step
    count after 9 steps, for every boundary rule, both reconstructions,
    HLLE alone, and tilings with 1, 2 and 4 wavefronts per workgroup."""
    from wenox import Solver
    nx, ny, dx, q0, kw = _layout_case(bc)
    kw.update(cons_switch=cons, force_hlle=force_hlle)
    print(f"[blocked] bc={bc} {nx}x{ny} cons={cons} hlle={force_hlle}", flush=True)
    ref = None
    for blocked, tiling in ((False, (9, 7, 1, 1)), (True, (9, 7, 1, 1)), (True, (5, 6, 2, 2)),
                            (True, (4, 3, 4, 4)), (True, (11, 4, 1, 4))):
        with Solver(nx, ny, dx, **kw) as s:
            s.set_tiling(*tiling)
            s.set_blocked(blocked)
            assert s.blocked == blocked
            s.upload(q0)
            T, k = s.run(nstep=9)
            got = (s.download(), T, k)
        if ref is None:
            ref = got
            # component-wise reconstruction has no positivity limiter, and the
            # double Mach reflection turns non-finite with it: there the
            # comparison takes NaN for NaN
            if not (cons and bc.startswith("dmr")):
                assert np.isfinite(ref[0]).all()
            continue
        assert got[2] == ref[2], (tiling, got[2], ref[2])
        assert got[1] == ref[1] or (got[1] != got[1] and ref[1] != ref[1]), (tiling, got[1], ref[1])
        np.testing.assert_array_equal(got[0], ref[0], err_msg=f"tiling {tiling}")


def test_blocked_layout_switches_and_reads():
    """Switching the layout between runs, and reading the state of a blocked
    run (download, conserved sums, RHS) in between, leaves the trajectory
    that of the row-major step; t_final runs end at the same time."""
    from wenox import Solver
    nx, ny, dx, q0, kw = _layout_case("mirror")
    print("[blocked] switching layouts between runs", flush=True)
    with Solver(nx, ny, dx, **kw) as s:
        assert s.blocked   # the fast FP64 step's default
        s.set_blocked(False)
        s.upload(q0)
        s.run(nstep=5)
        s.run(t_final=0.05)
        ref = (s.download(), s.conserved_sums(), s.rhs(), s.run(nstep=4))
        ref_q = s.download()
    with Solver(nx, ny, dx, **kw) as s:
        s.set_blocked(True)
        s.upload(q0)
        s.run(nstep=2)
        s.set_blocked(False)
        s.run(nstep=3)
        s.set_blocked(True)
        s.run(t_final=0.05)
        got = (s.download(), s.conserved_sums(), s.rhs(), s.run(nstep=4))
        got_q = s.download()
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[2], ref[2])
    assert got[3] == ref[3]
    np.testing.assert_array_equal(got_q, ref_q)


def test_blocked_layout_argument_checks():
    from wenox import Solver, problems
    q0, dx = problems.four_quadrant(16)
    for kw in (dict(engine="unfused"), dict(engine="fused", variant="exact"),
               dict(engine="fused", precision="fp32split")):
        with Solver(16, 16, dx, **kw) as s:
            with pytest.raises(ValueError, match="blocked layout"):
                s.set_blocked(True)
            assert not s.blocked


def test_autotune_picks_a_layout():
    """The autotuner tunes the tiling of both state layouts of the FP64 fast
    step, keeps the one with the faster tuned step, and the run that follows
    gives the row-major step's states (both layouts and every tiling round
    alike)."""
    from wenox import Solver, problems
    n = 128
    q0, dx = problems.four_quadrant(n)
    print(f"[autotune] layouts, four-quadrant {n}^2", flush=True)
    kw = dict(cfl=0.4, jp_cri=(0.5, 2.0), bc="mirror", engine="fused", variant="fast")
    with Solver(n, n, dx, **kw) as s:
        s.upload(q0)
        r = s.autotune(candidates=((1, 8), (1, 16), (2, 6), (4, 4)), reps=1)
        assert set(r["step_ms"]) == {"row", "blocked"}
        assert s.blocked == r["blocked"]
        assert all(v > 0 for v in r["step_ms"].values())
        s.run(nstep=6)
        got = s.download()
    with Solver(n, n, dx, **kw) as s:
        s.upload(q0)
        s.run(nstep=6)
        ref = s.download()
    np.testing.assert_array_equal(got, ref)
    # FP32-split and the exact variant have one layout to tune
    with Solver(n, n, dx, **dict(kw, precision="fp32split")) as s:
        s.upload(q0)
        r = s.autotune(candidates=((1, 8), (2, 6)), reps=1)
        assert set(r["step_ms"]) == {"row"} and not r["blocked"]
