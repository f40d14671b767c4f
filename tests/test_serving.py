"""Graph-replayed stepping and the continuous-batching serving loop on device 0.

Graph replays must reproduce eager stepping exactly (same launches, same arguments), and a
solve's result must be the same alone and in the mix: every solve runs the same
deterministic kernels on its own buffers, so the bar is bit-for-bit equality.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = dict(tau=0.6833, re=50.0, mesh_mode=0, u_ref=0.0, ndtfrc=10000)


def params(pkg, **kw):
    d = dict(BASE)
    d.update(kw)
    return pkg.default_params(**d)


@pytest.mark.parametrize("kw", [dict(nx=64, ny=32, nz=48, precision="fp32"),
                                dict(nx=16, ny=32, nz=32, precision="fp64"),
                                dict(nx=48, ny=32, nz=32, precision="bf16", cs_smag=0.1)])
def test_graph_replays_equal_eager_steps(pkg, kw):
    a = pkg.Solver(params(pkg, **kw))
    b = pkg.Solver(params(pkg, **kw))
    for s in (a, b):
        s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=11)
    for rep in range(4):
        a.step(10)
        b.step_graph(10)
        # eager work and synchronisation between the replays
        fa, fb = a.f(), b.f()
        print(f"{kw['precision']} replay {rep}: max |graph - eager| {np.max(np.abs(fa - fb)):.3e}", flush=True)
        assert np.isfinite(fa).all() and np.array_equal(fa, fb), rep
        assert a.step_count() == b.step_count() == 10 * (rep + 1)
    ra, rb = a.macros(), b.macros()
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("kw", [dict(nx=64, ny=32, nz=48, precision="fp32"),
                                dict(nx=16, ny=32, nz=32, precision="fp64")])
def test_graph_after_an_odd_step_and_a_restart_equals_eager(pkg, kw):
    """A chunk recorded at one buffer parity is never replayed from the other: step(1)
    then step_graph(10) equals step(11), and so does a replay after a fresh start."""
    a = pkg.Solver(params(pkg, **kw))
    b = pkg.Solver(params(pkg, **kw))
    for s in (a, b):
        s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=11)
    a.step(10)
    b.step_graph(10)  # recorded at parity 0
    a.step(1)
    b.step(1)
    a.step(10)
    b.step_graph(10)  # parity 1: recorded afresh, not the parity-0 chunk
    assert np.array_equal(a.f(), b.f())
    assert a.flux_sum() == b.flux_sum()
    for s in (a, b):
        s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=12)
    a.step(1)
    b.step(1)
    a.step(10)
    b.step_graph(10)
    fa, fb = a.f(), b.f()
    print(f"{kw['precision']}: max |graph - eager| {np.max(np.abs(fa - fb)):.3e}", flush=True)
    assert np.isfinite(fa).all() and np.array_equal(fa, fb)
    assert a.flux_sum() == b.flux_sum()


def test_graph_chunk_with_forcing_update_steps_eagerly(pkg):
    kw = dict(nx=64, ny=32, nz=32, precision="fp32", ndtfrc=7)
    a = pkg.Solver(params(pkg, **kw))
    b = pkg.Solver(params(pkg, **kw))
    for s in (a, b):
        s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=3)
    for _ in range(3):
        a.step(10)
        b.step_graph(10)
        assert a.force == b.force
        assert np.array_equal(a.f(), b.f())


def run_alone(pkg, p, nsteps, seed):
    s = pkg.Solver(p)
    s.init_equilibrium(u0=(0.0, 0.05, 0.0), amp=1e-3, seed=seed)
    s.step(nsteps)
    out = s.macros()
    s.close()
    return out


def test_results_are_the_same_alone_and_in_the_mix(pkg):
    reqs = [(params(pkg, nx=32, ny=64, nz=48, precision="fp32", cs_smag=0.1), 36, 1),
            (params(pkg, nx=48, ny=96, nz=64, precision="fp32", hill_alpha=1.7, ly=5.142 + 3.858 * 1.7), 40, 2),
            (params(pkg, nx=16, ny=32, nz=32, precision="fp64"), 30, 3),
            (params(pkg, nx=24, ny=48, nz=32, precision="bf16", re=200.0), 24, 4)]
    alone = [run_alone(pkg, p, n, sd) for p, n, sd in reqs]
    with pkg.Batch(max_active=3, chunk=10) as b:
        ids = [b.submit(p, n, u0=(0.0, 0.05, 0.0), amp=1e-3, seed=sd, keep_fields=True) for p, n, sd in reqs]
        done, failed = b.run()
        assert (done, failed) == (4, 0)
        for rid, ref in zip(ids, alone):
            st = b.status(rid)
            assert st["state"] == pkg.REQ_DONE
            got = b.result(rid)
            for x, y in zip(got, ref):
                assert np.array_equal(x, y, equal_nan=True), rid
            print(f"request {rid}: identical alone and in the mix, {st['steps']} steps", flush=True)


def test_grouped_launches_equal_solo_runs(pkg):
    """Eight FP32 solves of different grids advance together (one launch per step for all
    of them, chunks replayed from a recorded graph); odd step counts end chunks early and
    flip buffer parities, and forcing updates every 7 steps fall inside chunks.  Each
    result equals the same solve run alone, bit for bit."""
    grids = [(32, 64, 48), (48, 96, 64), (64, 128, 32), (32, 64, 96), (96, 64, 40), (16, 32, 32), (64, 48, 56),
             (128, 96, 24)]
    steps = [33, 41, 27, 50, 19, 36, 45, 23]
    # the near-wall xi treatment the serving workload uses (the reference's weights diverge
    # on some of these grids within 40 steps)
    reqs = [(params(pkg, nx=g[0], ny=g[1], nz=g[2], precision="fp32", cs_smag=0.1 * (k % 2), ndtfrc=7,
                    xi_near_wall=1), n, 50 + k)
            for k, (g, n) in enumerate(zip(grids, steps))]
    alone = [run_alone(pkg, p, n, sd) for p, n, sd in reqs]
    with pkg.Batch(max_active=8, chunk=10) as b:
        ids = [b.submit(p, n, u0=(0.0, 0.05, 0.0), amp=1e-3, seed=sd, keep_fields=True) for p, n, sd in reqs]
        done, failed = b.run()
        assert (done, failed) == (8, 0)
        gs = b.group_stats()
        # chunks end on the forcing updates, so their limits change from chunk to chunk
        print(f"grouped chunks {gs['chunks']}, graph recordings {gs['records']}", flush=True)
        assert gs["chunks"] >= 5
        for rid, ref in zip(ids, alone):
            got = b.result(rid)
            for x, y in zip(got, ref):
                assert np.isfinite(x).all() and np.array_equal(x, y), rid
    # without forcing updates the chunk graph is recorded once per change of members
    with pkg.Batch(max_active=8, chunk=10) as b:
        for k, g in enumerate(grids[:4]):
            b.submit(params(pkg, nx=g[0], ny=g[1], nz=g[2], precision="fp32"), 60, u0=(0.0, 0.05, 0.0), amp=1e-3,
                     seed=k)
        assert b.run() == (4, 0)
        gs = b.group_stats()
        print(f"no forcing: grouped chunks {gs['chunks']}, graph recordings {gs['records']}", flush=True)
        assert gs["chunks"] == 6 and gs["records"] == 1


def test_admission_respects_budget_and_slots(pkg):
    p = params(pkg, nx=32, ny=64, nz=64, precision="fp32")
    one = pkg.required_bytes(p)
    budget = 2 * one + one // 2
    with pkg.Batch(hbm_budget=budget, max_active=3, chunk=10) as b:
        big = b.submit(params(pkg, nx=128, ny=256, nz=128, precision="fp32"), 10)
        assert b.status(big)["state"] == pkg.REQ_FAILED  # larger than the whole budget
        ids = [b.submit(p, 40, u0=(0.0, 0.05, 0.0), amp=1e-3, seed=k) for k in range(5)]
        seen_active = 0
        while True:
            q, a, d, f = b.advance(True)
            assert a <= 2, "more solves admitted than the budget holds"
            assert b.bytes_in_use() <= budget
            seen_active = max(seen_active, a)
            if q == 0 and a == 0:
                break
        assert seen_active == 2
        assert (d, f) == (5, 1)
        st = [b.status(r) for r in ids]
        # continuous batching: later requests are admitted as earlier ones retire
        admits = sorted(s["t_admit"] for s in st)
        dones = sorted(s["t_done"] for s in st)
        assert admits[2] >= dones[0]
        for s in st:
            assert s["t_submit"] <= s["t_admit"] <= s["t_done"]
            assert s["steps"] == 40
        sm = b.summary(ids[0])
        assert sm["nodes"] == 32 * 64 * 64 and np.isfinite(sm["max_u"])


def test_requests_submitted_while_others_run_are_admitted(pkg):
    p = params(pkg, nx=32, ny=64, nz=48, precision="fp32")
    with pkg.Batch(max_active=2, chunk=10) as b:
        first = b.submit(p, 200, u0=(0.0, 0.05, 0.0), amp=1e-3, seed=1)
        b.advance(False)
        assert b.status(first)["state"] == pkg.REQ_ACTIVE
        late = b.submit(p, 20, u0=(0.0, 0.05, 0.0), amp=1e-3, seed=2)
        b.advance(False)
        # a free slot: the late request starts while the first one still runs
        assert b.status(late)["state"] == pkg.REQ_ACTIVE
        b.run()
        s1, s2 = b.status(first), b.status(late)
        assert s1["state"] == s2["state"] == pkg.REQ_DONE
        assert s2["t_admit"] < s1["t_done"]
