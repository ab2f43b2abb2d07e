"""GPU tests of the serving path: batched retire against per-slot retire,
and the continuous-batching server on resident GPU engines, where every
request's seismograms are bit-identical alone and in a mix with requests of
other sizes, widths and horizons that join and leave the running batch."""
import sys

import numpy as np
import pytest

import shaw_amd
from shaw_amd import serving

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def skew_rom(k, seed, prec, capacity=64, samples=40):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((k, k)) * (0.5 / np.sqrt(k))
    C = rng.standard_normal((7, k))
    return shaw_amd.Rom(A, -A.T.copy(), 0.5, prec, C, max_cols=capacity, max_samples=samples)


def request(rid, k, m, h, seed, every=5):
    rng = np.random.default_rng(seed)
    return serving.Request(rid, k, rng.standard_normal((k, m)), rng.standard_normal((h, m)), h, every, 0.0)


@pytest.mark.parametrize("prec", [shaw_amd.ROM_FP64, shaw_amd.ROM_BF16X4])
def test_retire_many_equals_retire(prec):
    a, b = skew_rom(48, 1, prec), skew_rom(48, 1, prec)
    r = request(0, 48, 9, 120, 2)
    sa = a.admit(r.b, r.ftab, [100, 120, 60, 120, 80, 120, 120, 35, 120], 5)
    sb = b.admit(r.b, r.ftab, [100, 120, 60, 120, 80, 120, 120, 35, 120], 5)
    a.advance(150)
    b.advance(150)
    one = [a.retire(s, 30)[0] for s in sa]
    many = b.retire_many(sb, 30)
    for x, y in zip(one, many):
        np.testing.assert_array_equal(x, y)
    assert [x.shape[0] for x in many] == [20, 24, 12, 24, 16, 24, 24, 7, 24]
    assert b.free_slots() == 64
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [shaw_amd.ROM_FP64, shaw_amd.ROM_BF16X3, shaw_amd.ROM_BF16X4])
def test_gpu_server_alone_equals_mix(prec):
    say(f"[gpu] server alone==mix precision {prec}")

    def engines():
        return {32: serving.GpuEngine(skew_rom(32, 10, prec), 64), 80: serving.GpuEngine(skew_rom(80, 11, prec), 64)}

    specs = [(0, 32, 5, 120), (1, 80, 17, 60), (2, 32, 40, 200), (3, 80, 3, 95), (4, 32, 20, 40), (5, 80, 33, 180),
             (6, 32, 30, 75)]
    mix = [request(rid, k, m, h, 100 + rid) for rid, k, m, h in specs]
    e = engines()
    srv = serving.Server(e, hbm_budget=1e12, chunk=23)
    rep = srv.run(mix)
    assert rep["requests"] == len(specs)
    for r in mix:
        ea = engines()
        alone = request(r.rid, r.k, r.m, r.horizon, 100 + r.rid)
        serving.Server(ea, hbm_budget=1e12, chunk=23).run([alone])
        assert len(alone.seismograms) == r.m
        for x, y in zip(r.seismograms, alone.seismograms):
            assert x.shape == (r.horizon // 5, 7)
            np.testing.assert_array_equal(x, y)
        for eng in ea.values():
            eng.rom.close()
    for eng in e.values():
        eng.rom.close()


def test_gpu_server_macro_steps_vs_fp64():
    """The server on BF16x4 engines with macro steps (forcing that ends after
    30 steps, sampling every 10 steps, chunks of 20 so that admissions stay
    aligned): every request's seismograms stay within the BF16 bound of the
    FP64 engines' (which are bit-identical alone and in a mix), and the macro
    steps did run (the results differ from one-step BF16x4). The same under a
    time budget (quiet tiles run ahead; the loop reads the engines' step
    counts), with the same seismogram lengths."""
    say("[gpu] server with macro steps")

    def engines(prec, macro, budget=0):
        e = {32: serving.GpuEngine(skew_rom(32, 10, prec), 64), 80: serving.GpuEngine(skew_rom(80, 11, prec), 64)}
        for eng in e.values():
            if macro:
                eng.rom.set_macro_steps(macro)
            if budget:
                eng.rom.set_time_budget(budget)
        return e

    specs = [(0, 32, 5, 120), (1, 80, 17, 60), (2, 32, 40, 200), (3, 80, 3, 95), (4, 32, 20, 40), (5, 80, 33, 180),
             (6, 32, 30, 75), (7, 80, 30, 400)]

    def requests():
        out = []
        for rid, k, m, h in specs:
            r = request(rid, k, m, h, 200 + rid, every=10)
            r.ftab[30:] = 0.0
            out.append(r)
        return out

    res, launches = {}, {}
    for name, prec, macro, budget in (("fp64", shaw_amd.ROM_FP64, 0, 0), ("one", shaw_amd.ROM_BF16X4, 0, 0),
                                      ("macro", shaw_amd.ROM_BF16X4, 10, 0),
                                      ("budget", shaw_amd.ROM_BF16X4, 10, 2)):
        e = engines(prec, macro, budget)
        reqs = requests()
        rep = serving.Server(e, hbm_budget=1e12, chunk=20).run(reqs)
        assert rep["requests"] == len(specs)
        assert rep["forcing_steps"] == sum(r.horizon * r.m for r in reqs)
        res[name], launches[name] = [r.seismograms for r in reqs], rep["launches"]
        for eng in e.values():
            eng.rom.close()
    for name in ("macro", "budget"):
        worst = 0.0
        for ref, mac in zip(res["fp64"], res[name]):
            for x, y in zip(mac, ref):
                assert x.shape == y.shape
                worst = max(worst, float(np.linalg.norm(x - y) / np.linalg.norm(y)))
        say(f"[gpu] {name} serving vs FP64: worst column {worst:.2e}, {launches[name]} launches")
        assert worst <= 2e-3
        assert any(np.abs(x - y).max() > 0 for a, b in zip(res["one"], res[name]) for x, y in zip(a, b))
    # quiet tiles ran ahead: fewer chunks than one step at a time
    assert launches["budget"] < launches["macro"]
