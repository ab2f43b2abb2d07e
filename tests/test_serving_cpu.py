"""CPU tests of the continuous-batching ROM server with the numpy engine:
FIFO admission gated by free slots and by the HBM footprint budget, requests
joining a running batch, uneven horizons, latency accounting, rejection of
requests that can never fit, and alone-vs-mixed equality of results."""
import numpy as np
import pytest

from shaw_amd import serving


def stable_ops(k, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((k, k)) * 0.05
    return A, -A.T.copy()  # skew pair: the reduced leapfrog stays bounded


def make_engine(k, cap, seed=0, r=3):
    arv, ars = stable_ops(k, seed)
    C = np.random.default_rng(seed + 1).standard_normal((r, k))
    return serving.NumpyEngine(arv, ars, 0.5, C, cap)


def req(rid, k, m, h, arrival, seed, every=5):
    rng = np.random.default_rng(seed)
    return serving.Request(rid, k, rng.standard_normal((k, m)), rng.standard_normal((h, m)), h, every, arrival)


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def test_results_alone_equal_results_in_mix():
    engines = {8: make_engine(8, 16), 12: make_engine(12, 16, seed=3)}
    mix = [req(0, 8, 3, 40, 0.0, 10), req(1, 12, 2, 25, 0.0, 11), req(2, 8, 5, 60, 0.0, 12),
           req(3, 8, 4, 20, 0.0, 13)]
    srv = serving.Server(engines, hbm_budget=1e12, chunk=7, clock=FakeClock())
    rep = srv.run(mix)
    assert rep["requests"] == 4
    for r in mix:
        alone_engines = {8: make_engine(8, 16), 12: make_engine(12, 16, seed=3)}
        a = req(r.rid, r.k, r.m, r.horizon, 0.0, 10 + r.rid)
        serving.Server(alone_engines, 1e12, chunk=7, clock=FakeClock()).run([a])
        for x, y in zip(r.seismograms, a.seismograms):
            np.testing.assert_array_equal(x, y)
        assert all(s.shape == (r.horizon // r.every, 3) for s in r.seismograms)


def test_admission_by_slots_is_fifo_and_joins_running_batch():
    engines = {8: make_engine(8, 6)}
    reqs = [req(0, 8, 4, 30, 0.0, 1), req(1, 8, 4, 30, 0.0, 2), req(2, 8, 1, 30, 0.0, 3)]
    srv = serving.Server(engines, hbm_budget=1e12, chunk=10, clock=FakeClock())
    for r in reqs:
        srv.submit(r)
    srv.step()
    # request 1 does not fit (4 + 4 > 6 slots); request 2 would, but FIFO keeps it queued
    assert [r.rid for r in srv.resident] == [0]
    assert [r.rid for r in srv.queue] == [1, 2]
    while srv.resident or srv.queue:
        srv.step()
    assert [r.rid for r in srv.finished] == [0, 1, 2]
    # 1 and 2 were admitted in the same scheduler step (one clock tick apart), after 0 finished
    assert abs(reqs[1].admitted - reqs[2].admitted) < 0.0015 and reqs[1].admitted >= reqs[0].done


def test_admission_by_hbm_footprint():
    e = make_engine(8, 64)
    engines = {8: e}
    one = serving.column_bytes(e.kv, e.ks, 30, 5, e.R)
    srv = serving.Server(engines, hbm_budget=5 * one, chunk=10, clock=FakeClock())
    a, b = req(0, 8, 3, 30, 0.0, 1), req(1, 8, 3, 30, 0.0, 2)
    srv.submit(a)
    srv.submit(b)
    srv.step()
    assert len(srv.resident) == 1 and srv.resident_bytes == 3 * one
    with pytest.raises(ValueError, match="footprint"):
        srv.submit(req(2, 8, 6, 30, 0.0, 3))
    with pytest.raises(ValueError, match="capacity"):
        serving.Server({8: make_engine(8, 4)}, 1e12).submit(req(3, 8, 5, 30, 0.0, 4))
    with pytest.raises(ValueError, match="no resident ROM"):
        srv.submit(req(4, 16, 1, 30, 0.0, 5))
    while srv.resident or srv.queue:
        srv.step()
    assert srv.resident_bytes == 0


def test_uneven_horizons_and_latency_accounting():
    engines = {8: make_engine(8, 32)}
    reqs = [req(i, 8, 2, h, 0.001 * i, 20 + i) for i, h in enumerate([15, 45, 30, 60])]
    srv = serving.Server(engines, hbm_budget=1e12, chunk=15, clock=FakeClock())
    rep = srv.run(reqs)
    assert rep["forcing_steps"] == sum(r.horizon * r.m for r in reqs)
    assert rep["p50_latency_s"] > 0 and rep["p99_latency_s"] >= rep["p50_latency_s"]
    assert all(r.done >= r.arrival for r in reqs)
    # submitted together, the shorter horizons finish in earlier scheduler steps
    engines = {8: make_engine(8, 32)}
    reqs = [req(i, 8, 2, h, 0.0, 30 + i) for i, h in enumerate([15, 45, 30, 60])]
    srv = serving.Server(engines, hbm_budget=1e12, chunk=15, clock=FakeClock())
    for r in reqs:
        srv.submit(r)
    order, steps = [], 0
    while srv.resident or srv.queue:
        srv.step()
        steps += 1
        order += [(r.rid, steps) for r in srv.finished if r.rid not in dict(order)]
    assert dict(order) == {0: 1, 2: 2, 1: 3, 3: 4}
    assert all(len(r.seismograms[0]) == r.horizon // r.every for r in reqs)


def test_poisson_request_generator():
    rng = np.random.default_rng(0)
    mk = lambda k, m, h, g: (np.zeros((k, m)), np.zeros((h, m)))
    reqs = serving.poisson_requests(rng, 200.0, 2.0, [128, 256, 512], (1, 256), (5000, 20000), mk)
    assert 300 < len(reqs) < 500
    assert all(1 <= r.m <= 256 and 5000 <= r.horizon <= 20000 and r.horizon % 10 == 0 for r in reqs)
    assert {r.k for r in reqs} == {128, 256, 512}
    assert all(np.diff([r.arrival for r in reqs]) > 0)


def test_sjf_admits_least_work_first_among_those_that_fit():
    """Policy "sjf": the least work (horizon x columns) first, not blocked by
    a queued request that does not fit; ties in arrival order."""
    engines = {8: make_engine(8, 6)}
    reqs = [req(0, 8, 4, 30, 0.0, 1), req(1, 8, 4, 30, 0.0, 2), req(2, 8, 1, 10, 0.0, 3)]
    srv = serving.Server(engines, hbm_budget=1e12, chunk=10, clock=FakeClock(), policy="sjf")
    for r in reqs:
        srv.submit(r)
    srv._admit()
    assert [r.rid for r in srv.resident] == [2, 0] and [r.rid for r in srv.queue] == [1]
    rep = srv.run([])
    assert rep["requests"] == 3 and not srv.queue and not srv.resident


def test_sjf_and_fifo_give_the_same_results():
    def mix():
        return [req(i, 8 if i % 3 else 12, 1 + i % 4, 15 + 7 * (i % 5), 0.0, 40 + i) for i in range(9)]

    out = {}
    for policy in ("fifo", "sjf"):
        engines = {8: make_engine(8, 6), 12: make_engine(12, 5, seed=3)}
        rs = mix()
        rep = serving.Server(engines, hbm_budget=1e12, chunk=6, clock=FakeClock(), policy=policy).run(rs)
        assert rep["requests"] == 9
        out[policy] = {r.rid: r.seismograms for r in rs}
    for rid in out["fifo"]:
        for x, y in zip(out["fifo"][rid], out["sjf"][rid]):
            np.testing.assert_array_equal(x, y)


def test_unknown_policy_is_refused():
    with pytest.raises(ValueError, match="admission policy"):
        serving.Server({8: make_engine(8, 6)}, hbm_budget=1e12, policy="lifo")
