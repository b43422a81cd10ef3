"""Serving loop (mgo_serve_*) on the device: a request's result is the same alone and in any mix of
other requests (bitwise), matches the oracle to the parity bars, and a scenario update applied to a
resident request equals the oracle run with the same damping added to every node's contact matrix
(ContactMatrix::add_damping, cpp/memilio/epidemiology/contact_matrix.h:190-194)."""
import numpy as np
import pytest

from memilio_amd import capi, synthetic
from memilio_amd.engine import Batch, Context, MgoError, Server
from tests import setups

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6))) if a.size else 0.0


def _requests():
    """Mixed classes: (G, integrator) differ, sizes differ, one single-node model, one fixed-step run."""
    a = synthetic.graph(num_nodes=60, G=3, seed=11, mean_degree=8.0)
    b = synthetic.graph(num_nodes=25, G=6, seed=12, mean_degree=5.0)
    c = synthetic.single_model(seed=13, G=2)
    d = synthetic.graph(num_nodes=90, G=1, seed=14, mean_degree=10.0)
    e = synthetic.graph(num_nodes=40, G=3, seed=15, mean_degree=6.0)
    return [
        (a, synthetic.ck_cfg(1e-8), 30.0, 1e-6),
        (b, synthetic.fixed_cfg(0.1), 21.0, 1e-12),
        (c, setups.rkf45_cfg(1e-8, 1e-8, 1e-6, 1.0), 45.0, 1e-6),
        (d, synthetic.ck_cfg(1e-10), 17.0, 1e-6),
        (e, synthetic.ck_cfg(1e-6), 33.0, 1e-6),
    ]


def _oracle_daily(oracle, spec, cfg, tmax):
    ref, _ = oracle.graph_run(spec, cfg, tmax)
    return ref


def test_serve_alone_vs_mix(oracle):
    reqs = _requests()
    with Context() as ctx:
        srv = Server(ctx)
        t = srv.submit(*reqs[0][:3])
        srv.wait([t])
        alone = srv.take(t)
        # the same request in a mix, submitted third; the others arrive around it
        order = [1, 3, 0, 2, 4]
        tickets = {i: srv.submit(*reqs[i][:3]) for i in order}
        infos = srv.wait(tickets.values())
        mixed = {i: srv.take(tickets[i]) for i in order}
        srv.stop()
    for i, tk in tickets.items():
        assert infos[tk].status == capi.OK
        assert infos[tk].admit_ms >= infos[tk].submit_ms
        assert infos[tk].done_ms >= infos[tk].admit_ms
    assert np.array_equal(alone, mixed[0]), "result differs alone vs in a mix"
    for i, (spec, cfg, tmax, tol) in enumerate(reqs):
        ref = _oracle_daily(oracle, spec, cfg, tmax)
        assert rel_err(mixed[i], ref) < tol, i


def test_serve_equals_batch_bitwise():
    """With MGO_SERVE_BATCH_KERNELS serving launches the per-(integrator, G) kernels a batch uses: a served
    request's daily values equal the same request run in a device batch, bit for bit."""
    reqs = _requests()
    with Context() as ctx:
        srv = Server(ctx, flags=capi.SERVE_BATCH_KERNELS)
        tickets = [srv.submit(*r[:3]) for r in reqs]
        infos = srv.wait(tickets)
        served = [srv.take(tk) for tk in tickets]
        srv.stop()
        assert all(infos[tk].status == capi.OK for tk in tickets)
        b = Batch(ctx, [r[0] for r in reqs], [r[1] for r in reqs], [r[2] for r in reqs])
        b.run()
        batched = b.read_daily()
        b.close()
    for i, (s, r) in enumerate(zip(served, batched)):
        assert np.array_equal(s.reshape(r.shape), r), i


def test_serve_batch_kernels_mix_and_small_pool(oracle):
    """MGO_SERVE_BATCH_KERNELS: alone equals in a mix bitwise and every request meets the parity bars; a
    result pool too small for some results stages those in pageable memory with the same values."""
    reqs = _requests()
    with Context() as ctx:
        srv = Server(ctx, flags=capi.SERVE_BATCH_KERNELS, result_pool_bytes=256 << 10)
        t = srv.submit(*reqs[4][:3])
        srv.wait([t])
        alone = srv.take(t)
        tickets = {i: srv.submit(*reqs[i][:3]) for i in range(len(reqs))}
        infos = srv.wait(tickets.values())
        mixed = {i: srv.take(tk) for i, tk in tickets.items()}
        srv.stop()
        with pytest.raises(MgoError) as e:
            Server(ctx, flags=8)
        assert e.value.status == capi.ERR_INVALID_VALUE
    assert all(infos[tk].status == capi.OK for tk in tickets.values())
    assert np.array_equal(alone, mixed[4])
    for i, (spec, cfg, tmax, tol) in enumerate(reqs):
        assert rel_err(mixed[i], _oracle_daily(oracle, spec, cfg, tmax)) < tol, i


def test_serve_queueing_and_budget(oracle):
    """max_resident = 2: requests queue and are admitted as others retire; results unchanged."""
    reqs = _requests()
    with Context() as ctx:
        srv = Server(ctx, max_resident=2)
        tickets = [srv.submit(*r[:3]) for r in reqs]
        infos = srv.wait(tickets)
        outs = [srv.take(tk) for tk in tickets]
        srv.stop()
    admits = sorted(infos[tk].admit_ms for tk in tickets)
    retires = sorted(infos[tk].retire_ms for tk in tickets)
    # the third admission waits for the first retirement
    assert admits[2] >= retires[0]
    assert all(infos[tk].done_ms >= infos[tk].retire_ms >= infos[tk].admit_ms for tk in tickets)
    for (spec, cfg, tmax, tol), out in zip(reqs, outs):
        assert rel_err(out, _oracle_daily(oracle, spec, cfg, tmax)) < tol


def test_serve_full_arena_churn(oracle):
    """A 12 MB serving arena that holds only a few requests at a time: admissions keep failing for memory and
    waiting for retirements while others run (a failed preparation waits for its queued uploads before its
    memory goes back to the arena). Every request finishes, equal to the oracle."""
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(24):
        n, G = int(rng.integers(60, 220)), int(rng.integers(2, 7))
        reqs.append((synthetic.graph(num_nodes=n, G=G, seed=300 + k, mean_degree=6.0), synthetic.ck_cfg(1e-8),
                     float(rng.integers(10, 25))))
    with Context() as ctx:
        srv = Server(ctx, hbm_budget_bytes=12 << 20)
        tickets = [srv.submit(*r) for r in reqs]
        infos = srv.wait(tickets, timeout_s=120.0)
        outs = [srv.take(tk) for tk in tickets]
        st = srv.stats()
        srv.stop()
    assert all(infos[tk].status == capi.OK for tk in tickets)
    assert st["max_resident_seen"] < len(reqs)       # the arena, not the slots, limited admission
    for (spec, cfg, tmax), out in zip(reqs, outs):
        assert rel_err(out, _oracle_daily(oracle, spec, cfg, tmax)) < 1e-6


def test_serve_damping_update(oracle):
    """Two scenario updates queued before admission (the request waits behind a long one with
    max_resident = 1) equal the oracle with the same dampings on every node from the start."""
    blocker = synthetic.graph(num_nodes=400, G=6, seed=21)
    spec = synthetic.graph(num_nodes=50, G=3, seed=22, mean_degree=6.0)
    cfg, tmax = synthetic.fixed_cfg(0.1), 60.0
    ups = [(0.4, 20.0, 0, 0), (0.8, 35.0, 1, 0)]
    with Context() as ctx:
        srv = Server(ctx, max_resident=1)
        tb = srv.submit(blocker, synthetic.ck_cfg(1e-10), 200.0)
        tk = srv.submit(spec, cfg, tmax)
        for s, ts, lev, typ in ups:
            srv.update(tk, s, ts, lev, typ)
        infos = srv.wait([tb, tk], timeout_s=120.0)
        assert infos[tk].admit_ms >= infos[tb].retire_ms   # the slot frees at the blocker's retirement
        assert infos[tb].done_ms >= infos[tb].retire_ms
        srv.take(tb)
        got = srv.take(tk)
        srv.stop()
    damped = synthetic.graph(num_nodes=50, G=3, seed=22, mean_degree=6.0)
    damped.set_dampings([[(ts, lev, typ, 1.0 - s) for s, ts, lev, typ in ups]] * damped.num_nodes)
    ref = _oracle_daily(oracle, damped, cfg, tmax)
    undamped = _oracle_daily(oracle, spec, cfg, tmax)
    assert rel_err(got, ref) < 1e-12
    assert rel_err(got, undamped) > 1e-3   # the update changed the course


def test_serve_errors():
    with Context() as ctx:
        srv = Server(ctx)
        spec = synthetic.single_model(seed=1, G=2)
        tk = srv.submit(spec, synthetic.ck_cfg(1e-8), 5.0)
        with pytest.raises(MgoError) as e:
            srv.update(tk, 1.5)
        assert e.value.status == capi.ERR_INVALID_VALUE
        with pytest.raises(MgoError) as e:
            srv.update(tk + 1000, 0.5)
        assert e.value.status == capi.ERR_NOT_FOUND
        bad = synthetic.single_model(seed=2, G=2)
        bad.age_params[0, capi.A_TIME_EXPOSED, 0] = -1.0
        with pytest.raises(MgoError) as e:
            srv.submit(bad, synthetic.ck_cfg(1e-8), 5.0)
        assert e.value.status == capi.ERR_INVALID_VALUE
        assert "stay times must be > 0" in str(e.value), str(e.value)   # the serving loop's message
        srv.wait([tk])
        out = srv.take(tk)
        assert out.shape == (6, 1, 20)
        with pytest.raises(MgoError) as e:
            Server(ctx)   # already serving
        assert e.value.status == capi.ERR_INVALID_VALUE
        srv.stop()


def test_serve_poisson_load_with_updates(oracle):
    """Serving under load (C4 shape): Poisson arrivals at 200 requests/s for 1.5 s of solves drawn from the
    uneven-batch distribution (1..2000 nodes, 1..8 age groups, 30..180 days, rel tol 1e-6/1e-8/1e-10), 10 % of
    the arrivals being contact-scaling updates U[0.3, 1] (from the next day) on a resident request. Every
    result finishes OK and keeps each node's total every day; every result equals, bit for bit, the same
    request served alone with the updates the loop reports it ran with (mgo_serve_updates_applied); the four
    smallest results with updates, and four without, equal the oracle run with those updates as dampings of
    every node's contact matrix (graph_simulation.h:101-120, contact_matrix.h:190-194)."""
    import threading
    import time
    pool = synthetic.uneven(48, seed=3)
    rate, seconds = 200.0, 1.5
    rng = np.random.default_rng([7, 200])
    n_arr = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n_arr)
    is_update = rng.random(n_arr) < 0.1
    picks = rng.integers(0, len(pool), n_arr)
    scales = rng.uniform(0.3, 1.0, n_arr)
    with Context() as ctx:
        srv = Server(ctx)
        tickets, live, req_of = [], [], {}
        t_next = time.perf_counter()
        for k in range(n_arr):
            t_next += gaps[k]
            while time.perf_counter() < t_next:
                time.sleep(0.0005)
            for r in srv.poll():
                srv.finished[r.ticket] = r   # Server.wait reads them there
            if is_update[k]:
                live = [tk for tk in live if tk not in srv.finished]
                if live:
                    tk = live[int(rng.integers(0, len(live)))]
                    try:
                        srv.update(tk, float(scales[k]))
                    except MgoError as e:
                        assert e.status == capi.ERR_OUT_OF_RANGE   # finished meanwhile
                continue
            spec, horizon, tol = pool[picks[k]]
            cfg = synthetic.ck_cfg(tol)
            tk = srv.submit(spec, cfg, horizon)
            tickets.append(tk)
            live.append(tk)
            req_of[tk] = (spec, cfg, horizon)
        infos = srv.wait(tickets, timeout_s=120.0)
        applied = {tk: srv.updates_applied(tk) for tk in tickets}
        served = {tk: srv.take(tk) for tk in tickets}
        # each request again, alone, with the updates it ran with
        alone = {}
        for tk in tickets:
            spec, cfg, horizon = req_of[tk]
            t2 = srv.submit(spec, cfg, horizon)
            for (s, ts, lev, typ) in applied[tk]:
                srv.update(t2, s, ts, lev, typ)
            srv.wait([t2], timeout_s=60.0)
            assert srv.updates_applied(t2) == applied[tk]
            alone[tk] = srv.take(t2)
        srv.stop()
    n_updates = sum(len(v) for v in applied.values())
    assert len(tickets) > 200 and n_updates > 10
    for tk in tickets:
        assert infos[tk].status == capi.OK, tk
        got = served[tk]
        spec, cfg, horizon = req_of[tk]
        assert got.shape == (spec.num_days(cfg.t0, horizon), spec.num_nodes, spec.state_size())
        tot = got.sum(axis=2)
        assert np.max(np.abs(tot - tot[0]) / tot[0]) < 1e-10, tk
        assert np.array_equal(got, alone[tk]), f"request {tk} differs served in the mix vs alone"
        for (s, ts, lev, typ) in applied[tk]:
            assert 0.3 <= s <= 1.0 and cfg.t0 + 1.0 <= ts <= horizon + 1.0
    by_size = sorted(tickets, key=lambda tk: req_of[tk][0].num_nodes * req_of[tk][2])
    sample = [tk for tk in by_size if applied[tk]][:4] + [tk for tk in by_size if not applied[tk]][:4]
    for tk in sample:
        spec, cfg, horizon = req_of[tk]
        if applied[tk]:
            spec = copy_spec(spec)
            spec.set_dampings([[(ts, lev, typ, 1.0 - s) for (s, ts, lev, typ) in applied[tk]]] * spec.num_nodes)
        assert rel_err(served[tk], _oracle_daily(oracle, spec, cfg, horizon)) < 1e-6, tk


def copy_spec(spec):
    """A GraphSpec with its own copies of every array (dampings are set on the copy)."""
    import copy
    c = copy.copy(spec)
    for name, v in vars(spec).items():
        if isinstance(v, np.ndarray):
            setattr(c, name, v.copy())
    c._desc = None
    return c
