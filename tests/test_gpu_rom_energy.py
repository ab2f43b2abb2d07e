"""The reduced leapfrog's discrete energy over 20000 steps at K = 256, 384
and 512 (the three row-tile layouts of the BF16 kernels), a size-independent
property of the ROM kernels at full horizon.

For an energy-consistent pair Ar_s = -Ar_v^T (an energy-orthonormal basis),
the reduced leapfrog q_v' = q_v + dt Ar_v q_s, q_s' = q_s + dt Ar_s q_v'
conserves Q = |q_v|^2 + |q_s|^2 + dt q_v^T Ar_v q_s exactly in exact
arithmetic once the forcing has ended (the FOM's invariant with unit
weights). FP64 keeps it to rounding; the BF16 kernels, whose operator parts
are exactly skew too (the split of -x is the negated split of x), drift by
the rounding of the state's bf16 split and the FP32 accumulation only. A
wrong tile, part or sign breaks it at O(1)."""
import sys

import numpy as np
import pytest

import shaw_amd

pytestmark = pytest.mark.gpu

# measured over 20000 steps (K = 256, 384, 512; four columns each): FP64 at
# most 5.3e-15, BF16x3 4.6e-5, BF16x4 3.7e-5; with macro steps of 10 BF16x3
# 2.1e-4 and BF16x4 3.9e-5; without the cross term the energy moves by 1e-2
# to 1.6e-1
TOL = {(shaw_amd.ROM_FP64, 0): 1e-12, (shaw_amd.ROM_BF16X3, 0): 2e-4, (shaw_amd.ROM_BF16X4, 0): 2e-4,
       (shaw_amd.ROM_BF16X3, 10): 1e-3, (shaw_amd.ROM_BF16X4, 10): 2e-4}


def q_energy(qv, qs, A, dt):
    return float(qv @ qv + qs @ qs + dt * qv @ (A @ qs))


@pytest.mark.parametrize("precision,macro", [(shaw_amd.ROM_FP64, 0), (shaw_amd.ROM_BF16X3, 0), (shaw_amd.ROM_BF16X4, 0),
                                             (shaw_amd.ROM_BF16X3, 10), (shaw_amd.ROM_BF16X4, 10)])
@pytest.mark.parametrize("K", [256, 384, 512])
def test_rom_energy_conserved(K, precision, macro):
    """(macro = 10: after the forcing every tile advances ten steps per
    product by (F^10 - I) / dt, whose bf16 parts are not exactly structured;
    the same bound holds.)"""
    print(f"[gpu] ROM energy K={K} precision={precision} macro={macro}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(K)
    M, n, nf, dt = 64, 20000, 50, 0.5
    A = rng.standard_normal((K, K)) / np.sqrt(K)  # singular values up to about 2: dt * 2 < 2, stable
    B = rng.standard_normal((K, M))
    ftab = np.zeros((n, M))
    ftab[:nf] = rng.standard_normal((nf, M))
    rom = shaw_amd.Rom(A, -A.T.copy(), dt, precision, None, max_cols=M)
    if macro:
        rom.set_macro_steps(macro)
    slots = rom.admit(B, ftab, [n] * M, 0)
    rom.advance(nf + 10)
    rom.sync()
    probe = slots[:: M // 4]
    q0 = [q_energy(*rom.state(s)[:2], A, dt) for s in probe]
    rom.advance(n - nf - 10)
    rom.sync()
    q1 = [q_energy(*rom.state(s)[:2], A, dt) for s in probe]
    e1 = [float(np.sum(np.square(np.concatenate(rom.state(s)[:2])))) for s in probe]
    rom.close()
    dev = np.abs(np.array(q1) / np.array(q0) - 1.0)
    plain = np.abs(np.array(e1) / np.array(q0) - 1.0)
    print(f"[gpu] K={K} precision={precision} macro={macro}: |Q_end / Q_start - 1| = {dev}, "
          f"without the cross term {plain}", flush=True)
    assert np.all(np.array(q0) > 0)
    assert dev.max() <= TOL[(precision, macro)]
    assert plain.min() > 10 * TOL[(precision, macro)]  # the property needs its cross term
