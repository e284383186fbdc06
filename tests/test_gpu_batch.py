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
