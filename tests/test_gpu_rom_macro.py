"""Macro steps of the BF16 ROM kernels (shaw_rom_set_macro_steps): tiles
whose columns are quiet and aligned advance m steps per product by
G = (F^m - I) / dt. Checked against the FP64 ROM (the one-step reference)
at the bounds the one-step BF16 runs meet, on every row-tile layout (K = 128,
256, 384, 512), with horizons that are not multiples of m, sampling, a tile
whose columns sit at different residues mod m (no macro steps), and chunked
launches; and the step counters and seismogram lengths equal the one-step
run's exactly. Under a time budget (shaw_rom_set_time_budget) quiet tiles
run ahead by the stated rule, with the same bounds; without macro steps a
budget changes nothing."""
import sys

import numpy as np
import pytest

import shaw_amd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def problem(K, M, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((K, K)) / np.sqrt(K)
    B = rng.standard_normal((K, M))
    C = rng.standard_normal((12, K))
    nf = 40 + 7 * (np.arange(M) % 5)  # forcing ends at different steps per column
    ftab = np.zeros((n, M))
    for j in range(M):
        ftab[: nf[j], j] = rng.standard_normal(nf[j])
    hz = [n - (j % 3) * 7 for j in range(M)]  # some horizons off the multiples of m
    return A, B, C, ftab, hz


def run(A, B, C, ftab, hz, precision, macro, every, chunks):
    K, M = B.shape
    n = ftab.shape[0]
    rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, precision, C, max_cols=M, max_samples=n // every)
    if macro:
        rom.set_macro_steps(macro)
    slots = rom.admit(B, ftab, hz, every)
    for c in chunks:
        rom.advance(c)
    rom.sync()
    steps = [rom.state(s)[2] for s in slots]
    out = [rom.retire(s, n // every) for s in slots]
    rom.close()
    return [o[0] for o in out], np.array([np.concatenate([o[1], o[2]]) for o in out]), steps


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", [shaw_amd.ROM_BF16X3, shaw_amd.ROM_BF16X4])
@pytest.mark.parametrize("K", [128, 512])
def test_time_budget_vs_fp64(K, precision):
    """Under a time budget (shaw_rom_set_time_budget, a macro step counting
    as two one-steps) the first launch of 200 gives a tile k0 one-steps, k0
    the first multiple of 10 at which its columns' forcing has ended, and
    (200 - k0) / 2 macro steps, then every launch 100 macro steps (1000
    steps); the step counters say so, every column ends at its horizon in 4
    launches instead of 15, and the results stay within the macro-step
    bound of the FP64 ROM."""
    say(f"[gpu] ROM time budget K={K} precision={precision}")
    M, n, every = 48, 3000, 10
    A, B, C, ftab, hz = problem(K, M, n, 7 * K + precision)
    ref_seis, ref_q, ref_steps = run(A, B, C, ftab, hz, shaw_amd.ROM_FP64, 0, every, [n])
    rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, precision, C, max_cols=M, max_samples=n // every)
    rom.set_macro_steps(10)
    rom.set_time_budget(2)
    slots = rom.admit(B, ftab, hz, every)
    seen = []
    for _ in range(15):
        rom.advance(200)
        seen.append(rom.steps(slots))
        if all(s == h for s, h in zip(seen[-1], hz)):
            break
    say(f"[gpu] steps after each launch (column 0): {[s[0] for s in seen]}")
    nf = 40 + 7 * (np.arange(M) % 5)  # where each column's forcing ends (problem())
    last = {}
    for j, sl in enumerate(slots):
        last[sl // 16] = max(last.get(sl // 16, 0), int(nf[j]))
    k0 = [-(-last[sl // 16] // 10) * 10 for sl in slots]
    assert seen[0] == [k + 10 * ((200 - k) // 2) for k in k0]
    assert [b - a for a, b in zip(seen[0], seen[1])] == [1000] * M
    assert len(seen) == 4 and seen[-1] == hz
    assert [rom.state(s)[2] for s in slots] == hz
    out = [rom.retire(s, n // every) for s in slots]
    rom.close()
    seis = [o[0] for o in out]
    q = np.array([np.concatenate([o[1], o[2]]) for o in out])
    for a, b in zip(seis, ref_seis):
        assert a.shape == b.shape
    err = max(rel(a, b) for a, b in zip(seis, ref_seis))
    say(f"[gpu] K={K} precision={precision}: time budget vs FP64 seismograms {err:.2e}, states {rel(q, ref_q):.2e}")
    bound = 2e-3 if precision == shaw_amd.ROM_BF16X3 else 5e-4
    assert err <= bound and rel(q, ref_q) <= bound


@pytest.mark.parametrize("precision", [shaw_amd.ROM_BF16X3, shaw_amd.ROM_BF16X4])
@pytest.mark.parametrize("K", [128, 256, 384, 512])
def test_macro_steps_vs_fp64(K, precision):
    say(f"[gpu] ROM macro K={K} precision={precision}")
    M, n, every = 48, 3000, 10
    A, B, C, ftab, hz = problem(K, M, n, K + precision)
    chunks = [1000, 1000, 1000]
    ref_seis, ref_q, ref_steps = run(A, B, C, ftab, hz, shaw_amd.ROM_FP64, 0, every, chunks)
    one_seis, one_q, one_steps = run(A, B, C, ftab, hz, precision, 0, every, chunks)
    mac_seis, mac_q, mac_steps = run(A, B, C, ftab, hz, precision, 10, every, chunks)
    assert mac_steps == one_steps == ref_steps == hz
    err_one = max(rel(a, b) for a, b in zip(one_seis, ref_seis))
    err_mac = max(rel(a, b) for a, b in zip(mac_seis, ref_seis))
    q_one, q_mac = rel(one_q, ref_q), rel(mac_q, ref_q)
    say(f"[gpu] K={K} precision={precision}: seismogram vs FP64 one-step {err_one:.2e} macro {err_mac:.2e}; "
        f"states {q_one:.2e} / {q_mac:.2e}")
    for a, b in zip(mac_seis, ref_seis):
        assert a.shape == b.shape
    bound = 2e-3 if precision == shaw_amd.ROM_BF16X3 else 5e-4
    assert err_mac <= bound and q_mac <= bound
    # the macro run changes the result (its tiles did take macro steps)
    assert max(np.abs(a - b).max() for a, b in zip(mac_seis, one_seis)) > 0


def test_macro_steps_skip_unaligned_tiles():
    """Two admissions five steps apart in one 16-column tile: the columns sit
    at different residues mod m, the tile takes no macro steps, and its
    results equal the one-step run bit for bit; a second tile admitted
    aligned takes them."""
    say("[gpu] ROM macro unaligned tile")
    K, n, every = 128, 2000, 10
    A, B, C, ftab, hz = problem(K, 32, n, 3)
    hz = [n] * 32

    def go(macro):
        rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, shaw_amd.ROM_BF16X4, C, max_cols=32, max_samples=n // every)
        if macro:
            rom.set_macro_steps(macro)
        s1 = rom.admit(B[:, :8], ftab[:, :8], hz[:8], every)     # 8 slots of a free tile
        rom.advance(5)
        s2 = rom.admit(B[:, 8:], ftab[:, 8:], [n - 5] * 24, every)  # the other tile, then the 8 left beside s1
        rom.advance(n)
        rom.sync()
        slots = s1 + s2
        res = [rom.retire(s, n // every) for s in slots]
        rom.close()
        return res, slots

    (one, _), (mac, slots) = go(0), go(10)
    shared = slots[0] // 16  # the tile of s1, shared with 8 columns of s2
    assert sum(1 for s in slots if s // 16 == shared) == 16
    for j, s in enumerate(slots):
        if s // 16 == shared:  # no macro steps there: identical to the one-step run
            for x, y in zip(one[j], mac[j]):
                np.testing.assert_array_equal(x, y)
    # the tile of s2 alone took macro steps
    assert any(np.abs(one[j][0] - mac[j][0]).max() > 0 for j, s in enumerate(slots) if s // 16 != shared)


def test_macro_admission_groups_columns_by_quiet_step():
    """With macro steps on, admission gives the free slots (in tile order)
    to the columns in the order their forcing ends, so each tile holds
    columns that turn quiet together."""
    say("[gpu] ROM macro admission order")
    K, n = 128, 400
    rng = np.random.default_rng(9)
    A = rng.standard_normal((K, K)) / np.sqrt(K)
    M = 48
    ends = rng.permutation(np.arange(20, 20 + 5 * M, 5))  # distinct last forcing steps
    ftab = np.zeros((n, M))
    for j in range(M):
        ftab[: ends[j], j] = 1.0 + rng.random(ends[j])
    rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, shaw_amd.ROM_BF16X4, None, max_cols=M)
    rom.set_macro_steps(10)
    slots = rom.admit(rng.standard_normal((K, M)), ftab, [n] * M, 0)
    rom.close()
    assert sorted(slots) == list(range(M))
    assert [int(ends[j]) for j in np.argsort(slots)] == sorted(int(e) for e in ends)


def test_macro_signal_admission_orders_by_signal_end():
    """Signal requests are ordered on the host by when their signals end
    (no wait for the device); the order agrees with the quiet steps of the
    forcing series themselves (the first step from which |f| stays at or
    below 2^-64 of its peak), here from the host forcing table: Ricker
    wavelets and gaussDer pulses of mixed delays and periods, and a
    sinusoid (never quiet) last."""
    say("[gpu] ROM macro signal admission order")
    K, n, dt = 128, 6000, 0.05
    rng = np.random.default_rng(11)
    A = rng.standard_normal((K, K)) / np.sqrt(K)
    M = 40
    kinds = ["ricker"] * 30 + ["gaussDer"] * 9 + ["sinusoid"]
    delays = list(rng.uniform(20.0, 120.0, M))
    periods = list(rng.uniform(5.0, 30.0, M))
    f = shaw_amd.forcing_table(kinds, delays, periods, dt, n)
    peak = np.abs(f).max(axis=0)
    above = np.abs(f) > peak * 2.0 ** -64
    quiet = np.array([n - int(np.argmax(above[::-1, j])) if above[-1, j] == 0 else n + 1 for j in range(M)])
    rom = shaw_amd.Rom(A, -A.T.copy(), dt, shaw_amd.ROM_BF16X4, None, max_cols=M)
    rom.set_macro_steps(10)
    slots = rom.admit_signals(rng.standard_normal((K, M)), kinds, delays, periods, n, None, 0)
    rom.close()
    assert sorted(slots) == list(range(M))
    by_slot = [int(quiet[j]) for j in np.argsort(slots)]
    say(f"[gpu] quiet steps in slot order: {by_slot}")
    assert by_slot == sorted(by_slot)
    assert by_slot[-1] == n + 1  # the sinusoid


def test_macro_steps_only_bf16():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((32, 32)) / 6.0
    rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, shaw_amd.ROM_FP64, None, max_cols=16)
    with pytest.raises(shaw_amd.ShawError):
        rom.set_macro_steps(10)
    rom.set_macro_steps(0)
    rom.close()
    rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, shaw_amd.ROM_BF16X3, None, max_cols=16)
    with pytest.raises(shaw_amd.ShawError):
        rom.set_macro_steps(1)
    with pytest.raises(shaw_amd.ShawError):
        rom.set_time_budget(65)
    rom.close()


def test_time_budget_without_macro_steps_keeps_step_counts():
    """A time budget on an engine without macro steps (FP64, or BF16 with
    m = 0) changes nothing: every launch advances exactly its steps."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((32, 32)) / 6.0
    B = rng.standard_normal((32, 20))
    f = np.zeros((500, 20))
    f[:5] = rng.standard_normal((5, 20))
    for prec in (shaw_amd.ROM_FP64, shaw_amd.ROM_BF16X4):
        rom = shaw_amd.Rom(A, -A.T.copy(), 0.5, prec, None, max_cols=32)
        rom.set_time_budget(2)
        slots = rom.admit(B, f, [500] * 20, 0)
        rom.advance(120)
        assert rom.steps(slots) == [120] * 20
        rom.advance(1000)
        assert rom.steps(slots) == [500] * 20
        rom.close()
