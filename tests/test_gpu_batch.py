"""GPU checks of the batched engine (many uneven instances, per-instance dt).

Each instance of a batch must evolve exactly as the same problem run alone
(the single-instance fused engine, itself bitwise equal to the oracle in the
exact variant), with its own time step.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(24, 24), (32, 40), (48, 32), (40, 40), (64, 24), (30, 62)]


def _instances(seed=20261203):
    from wenox import problems
    rng = np.random.default_rng(seed)
    return [problems.random_riemann_instance(rng, nx, ny) for nx, ny in SIZES]


@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_batch_matches_single_instances(variant):
    from wenox import Batch, Solver
    inst = _instances()
    print(f"[batch] {len(inst)} instances, 12 steps, variant={variant}", flush=True)
    with Batch([s for s in SIZES], [q[1] for q in inst], cfl=0.4, jp_cri=(1e4, 1.1e4),
               variant=variant) as b:
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        total = b.run(nstep=12)
        T, k = b.status()
        got = [b.download(i) for i in range(len(inst))]
    assert total == 12 * len(inst) and np.all(k == 12)
    assert len(set(np.round(T, 14))) > 1      # per-instance dt
    for i, ((nx, ny), (q, dx)) in enumerate(zip(SIZES, inst)):
        with Solver(nx, ny, dx, cfl=0.4, jp_cri=(1e4, 1.1e4), engine="fused", variant=variant) as s:
            s.upload(q)
            Ts, ks = s.run(nstep=12)
            ref = s.download()
        assert Ts == T[i]
        np.testing.assert_array_equal(got[i][..., :4], ref[..., :4], err_msg=f"instance {i}")


def test_batch_against_oracle(oracle_mod):
    from wenox import Batch
    inst = _instances(7)
    with Batch(SIZES, [q[1] for q in inst], cfl=0.4, jp_cri=(0.5, 2.0), variant="exact") as b:
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        b.run(nstep=5)
        got = [b.download(i) for i in range(len(inst))]
    for i in (0, 3, 5):
        nx, ny = SIZES[i]
        q, dx = inst[i]
        ref, _ = oracle_mod.run_riemann(q.copy(), nx, ny, dx, 1.4, 0.4, False, False, (0.5, 2.0), nstep=5)
        np.testing.assert_array_equal(got[i][..., :4], ref[..., :4], err_msg=f"instance {i}")


def test_batch_fp32_split():
    from wenox import Batch
    inst = _instances(11)
    res = {}
    for prec in ("fp64", "fp32split"):
        with Batch(SIZES, [q[1] for q in inst], cfl=0.4, jp_cri=(1e4, 1.1e4), variant="fast", precision=prec) as b:
            for i, (q, dx) in enumerate(inst):
                b.upload(i, q)
            b.run(nstep=20)
            res[prec] = [b.download(i) for i in range(len(inst))]
    for i, (nx, ny) in enumerate(SIZES):
        a = res["fp32split"][i][3:-3, 3:-3, 0]
        r = res["fp64"][i][3:-3, 3:-3, 0]
        err = np.abs(a - r).sum() / np.abs(r).sum()
        assert err <= 1e-6, (i, err)


@pytest.mark.parametrize("tau", [0.25, float("inf"), 0.05])
def test_batch_fp32_split_matches_single_instances(tau):
    """The batched FP32-split step with a split-jump threshold gives each
    instance the state of a single FP32-split solver with that threshold,
    bit for bit."""
    from wenox import Batch, Solver
    inst = _instances(7)
    with Batch(SIZES, [q[1] for q in inst], cfl=0.4, jp_cri=(1e4, 1.1e4), variant="fast",
               precision="fp32split") as b:
        b.set_split_jump(tau)
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        b.run(nstep=6)
        got = [b.download(i) for i in range(len(inst))]
    for i in (0, 2, 5):
        (nx, ny), (q, dx) = SIZES[i], inst[i]
        with Solver(nx, ny, dx, cfl=0.4, jp_cri=(1e4, 1.1e4), engine="fused", variant="fast",
                    precision="fp32split") as s:
            s.set_split_jump(tau)
            s.upload(q)
            s.run(nstep=6)
            np.testing.assert_array_equal(got[i][..., :4], s.download()[..., :4], err_msg=f"instance {i}")


def test_batch_t_final():
    from wenox import Batch
    inst = _instances(3)
    with Batch(SIZES, [q[1] for q in inst], cfl=0.4, jp_cri=(1e4, 1.1e4), variant="fast") as b:
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        b.run(t_final=0.02)
        T, k = b.status()
    assert np.all(np.abs(T - 0.02) <= 1e-12)
    assert len(set(k.tolist())) > 1           # instances needed different step counts


@pytest.mark.parametrize("variant", ["exact", "fast"])
def test_batch_chunks_and_autotune(variant):
    """Every sweep chunk (rows per workgroup), set by hand or picked by the
    autotuner, gives the same instance states bit for bit; the stage profile
    leaves the state unchanged."""
    from wenox import Batch
    inst = _instances(5)
    out = {}
    for chunks in ((16, 16), (5, 9), (64, 64), "auto"):
        print(f"[batch] chunks {chunks} variant={variant}", flush=True)
        with Batch(SIZES, [q[1] for q in inst], cfl=0.4, jp_cri=(0.5, 2.0), variant=variant) as b:
            for i, (q, dx) in enumerate(inst):
                b.upload(i, q)
            if chunks == "auto":
                tuned = b.autotune(reps=2)
                assert tuned["best"] == b.chunks[0]
            else:
                b.set_chunks(*chunks)
                assert b.chunks == chunks
                assert b.profile(2) > 0.0
            b.run(nstep=6)
            out[chunks] = ([b.download(i) for i in range(len(inst))], b.status())
    ref, (T0, k0) = out[(16, 16)]
    for chunks, (got, (T, k)) in out.items():
        assert np.array_equal(T, T0) and np.array_equal(k, k0), chunks
        for i in range(len(inst)):
            np.testing.assert_array_equal(got[i], ref[i], err_msg=f"chunks {chunks} instance {i}")


@pytest.mark.parametrize("bc,cons", [("mirror", False), ("periodic", False), ("copy", True)])
def test_batch_blocked_layout_same_states(bc, cons):
    """The batched fast FP64 step on blocked states (the default) and on
    row-major states give every instance the same state bit for bit, ghost
    cells included, and the same times, also when the layout is switched
    between runs and an instance is uploaded again in between."""
    from wenox import Batch
    inst = _instances(11)
    kw = dict(cfl=0.4, jp_cri=(0.5, 2.0), bc=bc, cons_switch=cons, variant="fast")
    print(f"[batch] blocked vs row-major, bc={bc} cons={cons}", flush=True)
    out = {}
    for mode in ("row", "blocked", "switch"):
        with Batch(SIZES, [q[1] for q in inst], **kw) as b:
            assert b.blocked
            if mode == "row":
                b.set_blocked(False)
                assert not b.blocked
            for i, (q, dx) in enumerate(inst):
                b.upload(i, q)
            if mode == "switch":
                b.run(nstep=3)
                b.set_blocked(False)
                b.run(nstep=2)
                b.set_blocked(True)
                b.upload(2, inst[2][0])
                b.run(nstep=4)
            else:
                b.run(nstep=5)
                b.upload(2, inst[2][0])
                b.run(nstep=4)
            T, k = b.status()
            out[mode] = (T, k, [b.download(i) for i in range(len(inst))])
    T0, k0, q0 = out["row"]
    assert all(np.isfinite(q).all() for q in q0)
    for mode in ("blocked", "switch"):
        T1, k1, q1 = out[mode]
        np.testing.assert_array_equal(k1, k0)
        np.testing.assert_array_equal(T1, T0)
        for i in range(len(inst)):
            np.testing.assert_array_equal(q1[i], q0[i], err_msg=f"{mode} instance {i}")


def _mixed_instances(count, seed):
    """count instances of sizes from {16, 24, ..., 96} x {16, ..., 96}."""
    from wenox import problems
    rng = np.random.default_rng(seed)
    sizes = [(int(a), int(b)) for a, b in rng.choice(np.arange(16, 97, 8), size=(count, 2))]
    return sizes, [problems.random_riemann_instance(rng, nx, ny) for nx, ny in sizes]


def _run_modes(sizes, inst, modes, script, **kw):
    from wenox import Batch
    out = {}
    for mode, graph in modes:
        print(f"[batch] {len(sizes)} instances, step structure {mode}, graphs {graph}, {kw}", flush=True)
        with Batch(sizes, [q[1] for q in inst], use_graph=graph, variant="fast", **kw) as b:
            b.set_persistent(mode == "persistent")
            assert b.persistent == (mode == "persistent")
            for i, (q, dx) in enumerate(inst):
                b.upload(i, q)
            script(b)
            T, k = b.status()
            out[(mode, graph)] = (T, k, [b.download(i) for i in range(len(sizes))])
    return out


def _assert_same(out, finite=True):
    keys = list(out)
    T0, k0, q0 = out[keys[0]]
    if finite:
        assert all(np.isfinite(q[3:-3, 3:-3, :4]).all() for q in q0)
    for key in keys[1:]:
        T1, k1, q1 = out[key]
        np.testing.assert_array_equal(k1, k0, err_msg=str(key))
        np.testing.assert_array_equal(T1, T0, err_msg=str(key))
        for i in range(len(q0)):
            np.testing.assert_array_equal(q1[i], q0[i], err_msg=f"{key} instance {i}")


@pytest.mark.parametrize("cons,force_hlle,bc", [(False, False, "mirror"), (False, True, "periodic"),
                                                (True, False, "copy"), (True, True, "mirror")])
def test_batch_persistent_same_states(cons, force_hlle, bc):
    """The persistent step (work items from a device queue, each waiting for
    its own instance's earlier phases) gives every instance the same state,
    time and step count bit for bit as the launch-per-phase step: single
    steps, eight-step launches with a remainder, a t_final run in which the
    instances stop at different steps inside one launch, and without graphs."""
    inst = _instances(13)

    def script(b):
        b.run(nstep=3)
        b.run(nstep=11)
        T, _ = b.status()
        b.run(t_final=float(T.max()) * 1.7)

    out = _run_modes(SIZES, inst, (("phases", True), ("persistent", True), ("persistent", False)), script,
                     cfl=0.4, jp_cri=(0.5, 2.0), bc=bc, cons_switch=cons, force_hlle=force_hlle)
    T, k, _ = out[("phases", True)]
    assert len(set(k.tolist())) > 1        # the t_final run stopped instances at different steps
    _assert_same(out)


@pytest.mark.parametrize("chunk", [5, 48])
def test_batch_persistent_many_items(chunk):
    """Forty uneven instances with short chunks (thousands of work items per
    step) and long ones: the persistent step equals the launch-per-phase step
    bit for bit over 17 steps (two eight-step launches and one single)."""
    sizes, inst = _mixed_instances(40, 29)
    out = _run_modes(sizes, inst, (("phases", True), ("persistent", True)), lambda b: b.run(nstep=17),
                     cfl=0.4, jp_cri=(1e4, 1.1e4), chunk=chunk)
    _assert_same(out)


def test_batch_persistent_wait_limit_error_path():
    """A dependency wait that passes its spin limit (forced here with a limit
    of zero polls, so waits give up at once) ends in an error from run(); the
    launch still drains, the flag is cleared, and with the normal limit the
    next run is correct again."""
    from wenox import Batch, WenoxError
    sizes, inst = _mixed_instances(40, 31)
    ref = _run_modes(sizes, inst, (("phases", True),), lambda b: b.run(nstep=8), cfl=0.4, jp_cri=(1e4, 1.1e4))
    with Batch(sizes, [q[1] for q in inst], cfl=0.4, jp_cri=(1e4, 1.1e4), variant="fast") as b:
        b.set_persistent(True)
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        b.set_spin_limit(0)
        print("[batch] persistent step with a zero spin limit (error path)", flush=True)
        with pytest.raises(WenoxError, match="spin limit"):
            b.run(nstep=8)
        b.set_spin_limit(1 << 22)
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        b.run(nstep=8)
        T, k = b.status()
        got = [b.download(i) for i in range(len(sizes))]
    T0, k0, q0 = ref[("phases", True)]
    np.testing.assert_array_equal(k, k0)
    np.testing.assert_array_equal(T, T0)
    for i in range(len(sizes)):
        np.testing.assert_array_equal(got[i], q0[i], err_msg=f"instance {i}")


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_profile_steps_restores_state(persistent):
    """profile_steps times whole steps and leaves states and times as they
    were; autotune picks a (structure, chunk) pair that gives the same states."""
    from wenox import Batch
    inst = _instances(17)

    def script(b):
        b.set_persistent(persistent)
        b.run(nstep=4)
        assert b.profile_steps(9) > 0.0
        b.run(nstep=5)

    out = _run_modes(SIZES, inst, (("phases", True),), script, cfl=0.4, jp_cri=(0.5, 2.0))
    ref = _run_modes(SIZES, inst, (("phases", True),), lambda b: b.run(nstep=9), cfl=0.4, jp_cri=(0.5, 2.0))
    _assert_same({"profiled": out[("phases", True)], "plain": ref[("phases", True)]})
    with Batch(SIZES, [q[1] for q in inst], cfl=0.4, jp_cri=(0.5, 2.0), variant="fast") as b:
        for i, (q, dx) in enumerate(inst):
            b.upload(i, q)
        tuned = b.autotune(candidates=(8, 16, 48), reps=1, steps=2)
        assert set(tuned["step_ms"]) == {"phases", "persistent"} and tuned["persistent"] == b.persistent
        b.run(nstep=9)
        T, k = b.status()
        got = [b.download(i) for i in range(len(SIZES))]
    _assert_same({"plain": ref[("phases", True)], "tuned": (T, k, got)})


def test_batch512_persistent_full_size():
    """The configured batched workload at its full size -- 512 instances of
    sizes drawn from {64, 96, 128, 192, 256, 384}^2, the bench's seed and
    state ranges -- advanced 16 steps by the persistent step (two eight-step
    launches) and by the launch-per-phase step: every instance's state, time
    and step count agree bit for bit, and each instance advanced by its own
    dt."""
    sizes_all = np.array([64, 96, 128, 192, 256, 384])
    from wenox import problems
    rng = np.random.default_rng(20261203)
    sizes = [(int(n), int(n)) for n in rng.choice(sizes_all, size=512)]
    inst = [problems.random_riemann_instance(rng, nx, ny) for nx, ny in sizes]
    out = _run_modes(sizes, inst, (("phases", True), ("persistent", True)), lambda b: b.run(nstep=16),
                     cfl=0.4, jp_cri=(1e4, 1.1e4), chunk=48)
    T, k, q = out[("phases", True)]
    assert np.all(k == 16) and len(set(np.round(T, 12).tolist())) > 100
    T1, k1, q1 = out[("persistent", True)]
    np.testing.assert_array_equal(k1, k)
    np.testing.assert_array_equal(T1, T)
    for i in range(len(sizes)):
        np.testing.assert_array_equal(q1[i], q[i], err_msg=f"instance {i} {sizes[i]}")


@pytest.mark.parametrize("bc", ["mirror", "periodic"])
def test_batch_persistent_tiny_and_lone_instances(bc):
    """Instances down to the smallest grid (3 x 3) beside a large one, and a
    batch of a single instance (one work queue busy, the other seven empty):
    the persistent step equals the launch-per-phase step bit for bit (the
    scheme itself may turn the tiniest grids non-finite: NaN compares equal
    to NaN here)."""
    from wenox import problems
    rng = np.random.default_rng(41)
    for sizes in ([(3, 3), (5, 7), (130, 6), (4, 64), (70, 70)], [(37, 29)]):
        inst = [problems.random_riemann_instance(rng, nx, ny) for nx, ny in sizes]
        out = _run_modes(sizes, inst, (("phases", True), ("persistent", True)), lambda b: b.run(nstep=11),
                         cfl=0.4, jp_cri=(0.5, 2.0), bc=bc)
        _assert_same(out, finite=False)


@pytest.mark.parametrize("cons,force_hlle,bc", [(False, False, "mirror"), (True, True, "periodic")])
def test_batch_packs_same_states(cons, force_hlle, bc):
    """Instances of one shape packed side by side along the wavefront's lanes
    (at most 8 or 3 a pack, or none) give every instance the same state, time
    and step count bit for bit as the launch-per-phase step: single steps,
    eight-step launches, and a t_final run in which members of one pack stop
    at different steps (their lanes then write nothing).  24 x 24 instances
    on another dx form packs of their own; a width of 58 (one wavefront's
    outputs) stays unpacked where packing would add lanes."""
    from wenox import Batch, problems
    rng = np.random.default_rng(43)
    # the work queues (one per XCD) split every shape's instances about
    # evenly, so each shape comes in multiples of the eight queues (five
    # 64 x 64 instances a queue pack by five under pack_max 8, by three
    # under 3)
    shapes = ([(24, 24, 1.0)] * 48 + [(40, 16, 1.0)] * 24 + [(64, 64, 1.0)] * 40 + [(58, 58, 1.0)] * 8 +
              [(24, 24, 1.5)] * 16 + [(7, 5, 1.0)] * 8)
    sizes = [(nx, ny) for nx, ny, _ in shapes]
    inst = []
    for nx, ny, f in shapes:
        q, dx = problems.random_riemann_instance(rng, nx, ny)
        inst.append((q, dx * f))

    def script(b):
        b.run(nstep=3)
        b.run(nstep=11)
        T, _ = b.status()
        # the scheme may turn the 7 x 5 grids non-finite (NaN times)
        b.run(t_final=float(np.median(T[np.isfinite(T)])) * 1.25)

    out, packs = {}, {}
    for mode, pack_max in (("phases", 8), ("persistent", 1), ("persistent", 3), ("persistent", 8)):
        print(f"[batch] {len(sizes)} instances, {mode}, pack_max {pack_max}, cons={cons}, bc={bc}", flush=True)
        with Batch(sizes, [dx for _, dx in inst], cfl=0.4, jp_cri=(0.5, 2.0), bc=bc, cons_switch=cons,
                   force_hlle=force_hlle, variant="fast") as b:
            b.set_persistent(mode == "persistent")
            b.set_pack_max(pack_max)
            packs[(mode, pack_max)] = b.packs
            for i, (q, _) in enumerate(inst):
                b.upload(i, q)
            script(b)
            T, k = b.status()
            out[(mode, pack_max)] = (T, k, [b.download(i) for i in range(len(sizes))])
    assert packs[("phases", 8)] == (8, 0)
    assert packs[("persistent", 1)] == (1, len(sizes))
    assert packs[("persistent", 8)][1] < packs[("persistent", 3)][1] < len(sizes)
    T, k, _ = out[("phases", 8)]
    assert len(set(k.tolist())) > 1        # the t_final run stopped instances at different steps
    _assert_same(out, finite=False)


def test_batch_persistent_on_a_subset_of_xcds():
    """The probe launch at creation reaches all eight XCDs of the MI355X; the
    persistent step serving only some of them (as on a partitioned device,
    where a launch reaches fewer) gives the same states bit for bit as the
    launch-per-phase step; masks outside the probed XCDs are refused."""
    from wenox import Batch
    sizes, inst = _mixed_instances(24, 37)
    ref = _run_modes(sizes, inst, (("phases", True),), lambda b: b.run(nstep=11), cfl=0.4, jp_cri=(0.5, 2.0))
    for mask in (0b00100001, 0b10000000, 0):
        with Batch(sizes, [q[1] for q in inst], cfl=0.4, jp_cri=(0.5, 2.0), variant="fast") as b:
            in_use, probed = b.persist_xcds
            assert probed == 0xFF and in_use == 0xFF
            with pytest.raises(ValueError):
                b.set_persist_xcds(1 << 9)
            b.set_persist_xcds(mask)
            assert b.persist_xcds[0] == (mask or 0xFF) and b.persistent
            print(f"[batch] persistent step on XCD mask {mask:#x}", flush=True)
            for i, (q, _) in enumerate(inst):
                b.upload(i, q)
            b.run(nstep=11)
            T, k = b.status()
            got = [b.download(i) for i in range(len(sizes))]
        _assert_same({"phases": ref[("phases", True)], "masked": (T, k, got)})
