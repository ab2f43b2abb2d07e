"""CPU tests of the continuous-batching serving loop (shaw_serve, C++) driving
numpy engines through the engine table, on a simulated clock: FIFO admission
gated by free slots and by the HBM footprint budget, requests joining a
running batch, uneven horizons, latency accounting, refusal of requests that
can never fit, least-work-first admission, its wait bound under sustained
overload, and alone-vs-mixed equality of results."""
import numpy as np
import pytest

import shaw_amd
from shaw_amd import serving


def stable_ops(k, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((k, k)) * 0.05
    return A, -A.T.copy()  # skew pair: the reduced leapfrog stays bounded


def make_engine(k, cap, seed=0, r=3, clock=None):
    arv, ars = stable_ops(k, seed)
    C = np.random.default_rng(seed + 1).standard_normal((r, k))
    return serving.NumpyEngine(arv, ars, 0.5, C, cap, tick=clock.work if clock else None)


def req(rid, k, m, h, arrival, seed, every=5):
    rng = np.random.default_rng(seed)
    return serving.Request(rid, k, rng.standard_normal((k, m)), rng.standard_normal((h, m)), h, every, arrival)


def serve(engines, reqs, clock, **kw):
    return serving.Server(engines, clock=clock, **kw).run(reqs)


def test_results_alone_equal_results_in_mix():
    clk = serving.SimClock()
    engines = {8: make_engine(8, 16, clock=clk), 12: make_engine(12, 16, seed=3, clock=clk)}
    mix = [req(0, 8, 3, 40, 0.0, 10), req(1, 12, 2, 25, 0.0, 11), req(2, 8, 5, 60, 0.0, 12),
           req(3, 8, 4, 20, 0.0, 13)]
    rep = serve(engines, mix, clk, hbm_budget=1e12, chunk=7)
    assert rep["requests"] == 4
    for r in mix:
        c2 = serving.SimClock()
        alone_engines = {8: make_engine(8, 16, clock=c2), 12: make_engine(12, 16, seed=3, clock=c2)}
        a = req(r.rid, r.k, r.m, r.horizon, 0.0, 10 + r.rid)
        serve(alone_engines, [a], c2, hbm_budget=1e12, chunk=7)
        assert len(a.seismograms) == r.m
        for x, y in zip(r.seismograms, a.seismograms):
            np.testing.assert_array_equal(x, y)
        assert all(s.shape == (r.horizon // r.every, 3) for s in r.seismograms)


def test_seismograms_equal_the_engine_run_directly():
    """The loop's seismograms are the engine's own, column by column."""
    clk = serving.SimClock()
    e = make_engine(8, 8, clock=clk)
    r = req(0, 8, 3, 35, 0.0, 5)
    serve({8: e}, [r], clk, hbm_budget=1e12, chunk=6)
    d = make_engine(8, 8)
    slots = d.admit(r.b, r.ftab, [r.horizon] * r.m, r.every)
    d.advance(r.horizon)
    for j, s in enumerate(slots):
        np.testing.assert_array_equal(r.seismograms[j], d.retire(s))


def test_admission_by_slots_is_fifo_and_joins_running_batch():
    clk = serving.SimClock()
    engines = {8: make_engine(8, 6, clock=clk)}
    reqs = [req(0, 8, 4, 30, 0.0, 1), req(1, 8, 4, 30, 0.0, 2), req(2, 8, 1, 30, 0.0, 3)]
    serve(engines, reqs, clk, hbm_budget=1e12, chunk=10)
    a, b, c = reqs
    # request 1 does not fit beside 0 (4 + 4 > 6 slots); request 2 would, but FIFO keeps it behind 1
    assert b.admitted >= a.done and c.admitted >= a.done and c.admitted >= b.admitted
    # 1 and 2 join in the same admission pass (no engine work between them)
    assert c.admitted - b.admitted < 1e-4


def test_late_arrival_joins_the_running_batch():
    clk = serving.SimClock(per_step=1e-3)
    engines = {8: make_engine(8, 16, clock=clk)}
    reqs = [req(0, 8, 2, 100, 0.0, 1), req(1, 8, 2, 20, 0.015, 2)]
    serve(engines, reqs, clk, hbm_budget=1e12, chunk=10)
    a, b = reqs
    assert a.admitted < b.admitted < a.done  # admitted while request 0 was mid-flight
    assert b.done < a.done


def test_admission_by_hbm_footprint():
    clk = serving.SimClock()
    e = make_engine(8, 64, clock=clk)
    one = serving.column_bytes(e.kv, e.ks, 30, 5, e.R)
    a, b = req(0, 8, 3, 30, 0.0, 1), req(1, 8, 3, 30, 0.0, 2)
    serve({8: e}, [a, b], clk, hbm_budget=5 * one, chunk=10)
    assert b.admitted >= a.done  # 3 + 3 columns exceed the 5-column budget
    with pytest.raises(shaw_amd.ShawError, match="footprint"):
        serve({8: make_engine(8, 64)}, [req(2, 8, 6, 30, 0.0, 3)], serving.SimClock(), hbm_budget=5 * one)
    with pytest.raises(shaw_amd.ShawError, match="capacity"):
        serve({8: make_engine(8, 4)}, [req(3, 8, 5, 30, 0.0, 4)], serving.SimClock(), hbm_budget=1e12)
    with pytest.raises(ValueError, match="no resident ROM"):
        serve({8: make_engine(8, 4)}, [req(4, 16, 1, 30, 0.0, 5)], serving.SimClock(), hbm_budget=1e12)


def test_uneven_horizons_and_latency_accounting():
    clk = serving.SimClock()
    engines = {8: make_engine(8, 32, clock=clk)}
    reqs = [req(i, 8, 2, h, 0.001 * i, 20 + i) for i, h in enumerate([15, 45, 30, 60])]
    rep = serve(engines, reqs, clk, hbm_budget=1e12, chunk=15)
    assert rep["forcing_steps"] == sum(r.horizon * r.m for r in reqs)
    lat = np.array([r.latency for r in reqs])
    assert rep["p50_latency_s"] == pytest.approx(np.percentile(lat, 50))
    assert rep["p99_latency_s"] == pytest.approx(np.percentile(lat, 99))
    assert rep["max_latency_s"] == pytest.approx(lat.max())
    assert all(r.done >= r.admitted >= r.arrival for r in reqs)
    # submitted together, the shorter horizons finish first
    clk = serving.SimClock()
    engines = {8: make_engine(8, 32, clock=clk)}
    reqs = [req(i, 8, 2, h, 0.0, 30 + i) for i, h in enumerate([15, 45, 30, 60])]
    serve(engines, reqs, clk, hbm_budget=1e12, chunk=15)
    assert [r.rid for r in sorted(reqs, key=lambda r: r.done)] == [0, 2, 1, 3]
    assert all(len(r.seismograms[0]) == r.horizon // r.every for r in reqs)


def test_poisson_request_generator():
    rng = np.random.default_rng(0)
    mk = lambda k, m, h, g: (np.zeros((k, m)), np.zeros((h, m)))
    reqs = serving.poisson_requests(rng, 200.0, 2.0, [128, 256, 512], (1, 256), (5000, 20000), mk)
    assert 300 < len(reqs) < 500
    assert all(1 <= r.m <= 256 and 5000 <= r.horizon <= 20000 and r.horizon % 10 == 0 for r in reqs)
    assert {r.k for r in reqs} == {128, 256, 512}
    assert all(np.diff([r.arrival for r in reqs]) > 0)


def test_least_work_first_among_those_that_fit():
    """Least work (horizon x columns) first, not blocked by a queued request
    that does not fit."""
    clk = serving.SimClock()
    engines = {8: make_engine(8, 6, clock=clk)}
    reqs = [req(0, 8, 4, 30, 0.0, 1), req(1, 8, 4, 30, 0.0, 2), req(2, 8, 1, 10, 0.0, 3)]
    rep = serve(engines, reqs, clk, hbm_budget=1e12, chunk=10, policy="least_work")
    a, b, c = reqs
    assert rep["requests"] == 3
    assert c.admitted <= a.admitted < b.admitted and b.admitted >= a.done


def test_least_work_and_fifo_give_the_same_results():
    def mix():
        return [req(i, 8 if i % 3 else 12, 1 + i % 4, 15 + 7 * (i % 5), 0.0, 40 + i) for i in range(9)]

    out = {}
    for policy in ("fifo", "least_work"):
        clk = serving.SimClock()
        engines = {8: make_engine(8, 6, clock=clk), 12: make_engine(12, 5, seed=3, clock=clk)}
        rs = mix()
        rep = serve(engines, rs, clk, hbm_budget=1e12, chunk=6, policy=policy)
        assert rep["requests"] == 9
        out[policy] = {r.rid: r.seismograms for r in rs}
    for rid in out["fifo"]:
        for x, y in zip(out["fifo"][rid], out["least_work"][rid]):
            np.testing.assert_array_equal(x, y)


def _overload(max_wait):
    """A stream of small requests (2 columns, 10 steps) arriving faster than
    they drain keeps an 8-slot engine full; one large request (all 8 slots)
    arrives early."""
    clk = serving.SimClock(per_call=1e-6, per_step=1e-4)
    e = make_engine(4, 8, r=1, clock=clk)
    small = [req(i + 1, 4, 2, 10, 2e-4 * i, 100 + i, every=10) for i in range(600)]
    big = req(0, 4, 8, 10, 1e-3, 7, every=10)
    rep = serve({4: e}, [big] + small, clk, hbm_budget=1e12, chunk=10, policy="least_work", max_wait=max_wait)
    return big, small, rep


def test_least_work_wait_is_bounded_under_sustained_overload():
    """Without a bound the large request waits until the stream ends (it
    never fits while smaller work keeps arriving); with max_wait it is
    admitted within max_wait plus the time the resident work needs to
    drain (one chunk here), and every request still completes."""
    big, small, rep = _overload(0.0)
    stream_end = max(r.arrival for r in small)
    assert big.admitted - big.arrival > 0.5 * stream_end
    W = 0.005
    big, small, rep = _overload(W)
    assert big.admitted - big.arrival <= W + 0.003
    assert rep["requests"] == len(small) + 1
    assert rep["max_wait_s"] == pytest.approx(max(r.admitted - r.arrival for r in small + [big]))


def test_overdue_request_holds_only_its_engine():
    """A request past its wait bound that waits for its engine's slots
    keeps the requests of that engine from passing it, but not those of
    another engine: they are admitted as they arrive (within about one
    chunk of either engine), although the overdue one waits 20 chunks for
    its engine's resident requests to finish."""
    clk = serving.SimClock(per_call=1e-6, per_step=1e-4)
    a, b = make_engine(4, 8, r=1, clock=clk), make_engine(6, 8, seed=5, r=1, clock=clk)
    small = [req(i + 1, 4, 2, 200, 2e-4 * i, 100 + i, every=10) for i in range(40)]
    big = req(0, 4, 8, 10, 1e-3, 7, every=10)
    other = [req(1000 + i, 6, 1, 10, 5e-4 * i + 1e-4, 500 + i, every=10) for i in range(200)]
    W = 0.005
    rep = serve({4: a, 6: b}, [big] + small + other, clk, hbm_budget=1e12, chunk=10, policy="least_work",
                max_wait=W)
    assert rep["requests"] == 1 + len(small) + len(other)
    overdue = big.arrival + W
    assert big.admitted - overdue > 0.02  # its engine's resident requests take 200 steps
    waiting = [r for r in other if overdue < r.arrival < big.admitted]
    assert len(waiting) > 20
    # the other engine's requests are not held back by it
    assert max(r.admitted - r.arrival for r in waiting) < 3e-3
    # no later request of its own engine passes it while it waits
    assert not [r for r in small if r.arrival > big.arrival and overdue <= r.admitted < big.admitted]
    assert any(r.arrival > big.arrival and r.admitted > big.admitted for r in small)


def test_unknown_policy_is_refused():
    with pytest.raises(ValueError, match="admission policy"):
        serving.Server({8: make_engine(8, 6)}, hbm_budget=1e12, policy="lifo")


def test_unkept_seismograms_are_discarded():
    """keep(r) False: the loop retires into a scratch buffer (seis NULL) and
    returns no seismograms for that request; kept requests are unaffected."""
    clk = serving.SimClock()
    engines = {8: make_engine(8, 16, clock=clk)}
    reqs = [req(0, 8, 2, 20, 0.0, 1), req(1, 8, 3, 25, 0.0, 2)]
    serve_all = serving.Server(engines, clock=clk, hbm_budget=1e12, chunk=5)
    rep = serve_all.run(reqs, keep=lambda r: r.rid == 1)
    assert rep["requests"] == 2 and reqs[0].seismograms == [] and len(reqs[1].seismograms) == 3
    c2 = serving.SimClock()
    alone = req(1, 8, 3, 25, 0.0, 2)
    serving.Server({8: make_engine(8, 16, clock=c2)}, clock=c2, hbm_budget=1e12, chunk=5).run([alone])
    for x, y in zip(reqs[1].seismograms, alone.seismograms):
        np.testing.assert_array_equal(x, y)


def test_signal_requests_equal_table_requests():
    """A request that carries its forcing as signal parameters (the engine
    evaluates the series at admission) gives the seismograms of the same
    request with the host table; an engine without that entry refuses it."""
    clk = serving.SimClock()
    engines = {8: make_engine(8, 16, clock=clk)}
    rng = np.random.default_rng(40)
    kinds, delays, periods = ["ricker", "gaussDer", "sinusoid"], [4.0, 6.0, 2.0], [3.0, 5.0, 9.0]
    b = rng.standard_normal((8, 3))
    h = 45
    tab = shaw_amd.forcing_table(kinds, delays, periods, 0.5, h)
    sig = serving.Request(0, 8, b, None, h, 5, 0.0, signals=(kinds, delays, periods))
    ref = serving.Request(1, 8, b, tab, h, 5, 0.0)
    serve(engines, [sig, ref], clk, hbm_budget=1e12, chunk=10)
    assert len(sig.seismograms) == 3
    for x, y in zip(sig.seismograms, ref.seismograms):
        np.testing.assert_array_equal(x, y)
    assert any(c[0] == "admit_signals" for c in engines[8].calls)
    eng = make_engine(8, 16)
    t = eng.table()
    t.admit_signals = serving._ADMIT_SIG()  # NULL: tables only

    class TablesOnly:
        kv, ks, R = eng.kv, eng.ks, eng.R

        def table(self):
            return t

    with pytest.raises(shaw_amd.ShawError, match="forcing tables only"):
        serve({8: TablesOnly()}, [serving.Request(2, 8, b, None, h, 5, 0.0, signals=(kinds, delays, periods))],
              serving.SimClock(), hbm_budget=1e12)


def quiet_req(rid, k, m, h, arrival, seed, forced, every=5):
    """A request whose forcing is zero after its first `forced` steps."""
    r = req(rid, k, m, h, arrival, seed, every)
    r.ftab[forced:] = 0.0
    return r


def test_time_budget_engine_reports_its_steps():
    """An engine whose columns advance by different amounts per chunk (as
    GPU engines under a time budget do, quiet columns further) reports its
    step counts; the loop retires a request only when its slowest column is
    at the horizon, counts every step once, and the results are the engine's
    own one-step results."""
    clk = serving.SimClock()
    arv, ars = stable_ops(8, 0)
    C = np.random.default_rng(1).standard_normal((3, 8))
    e = serving.NumpyEngine(arv, ars, 0.5, C, 32, tick=clk.work, stretch=4)
    reqs = [quiet_req(0, 8, 3, 90, 0.0, 40, forced=10), quiet_req(1, 8, 2, 60, 0.0, 41, forced=60),
            quiet_req(2, 8, 4, 75, 0.001, 42, forced=5), quiet_req(3, 8, 1, 30, 0.002, 43, forced=0)]
    # one column of request 0 keeps its forcing to the end: that request
    # advances at the one-step pace of its slowest column
    reqs[0].ftab[:, 1] = np.random.default_rng(7).standard_normal(90)
    rep = serve({8: e}, reqs, clk, hbm_budget=1e12, chunk=10)
    assert rep["forcing_steps"] == sum(r.horizon * r.m for r in reqs)
    assert ("advance", 10) in e.calls
    for r in reqs:
        assert all(s.shape == (r.horizon // r.every, 3) for s in r.seismograms)
        d = make_engine(8, 8)
        slots = d.admit(r.b, r.ftab, [r.horizon] * r.m, r.every)
        d.advance(r.horizon)
        for j, s in enumerate(slots):
            np.testing.assert_array_equal(r.seismograms[j], d.retire(s))
    # request 2 (75 steps, quiet after 5) takes 3 chunks and finishes before
    # request 1 (60 steps, forced throughout, 6 chunks); one step at a time
    # it would take 8
    assert reqs[2].done < reqs[1].done
