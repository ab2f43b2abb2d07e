"""The FOM's discrete energy at the full config-2 grid (1024 x 4096) over
thousands of steps: a size-independent parity property.

With the energy weights Mv, Ms of the operators (Mv J_vp = -(Ms J_sp)^T on
every coupling of a point off the axis, whose velocity stays zero), the
leapfrog v' = v + dt J_vp s, s' = s + dt J_sp v' conserves

    Q(v, s) = v^T Mv v + s^T Ms s + dt v^T Mv J_vp s

exactly in exact arithmetic once the forcing has ended (expand Q(v', s') -
Q(v, s): the dt^2 terms of the two squares cancel those of the cross term).
A wrong coefficient, neighbour, sign or step order in any kernel breaks it
at O(1); the oracle's own run conserves it to a few ulp
(tests/test_energy_cpu.py::test_fom_discrete_energy_conserved)."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import shaw_amd
from shaw_amd import workloads

pytestmark = pytest.mark.gpu


def energy_parts(v, s, mv, ms, Jv, dt):
    e = float(v @ (mv * v) + s @ (ms * s))
    x = float(dt * ((mv * v) @ (Jv @ s)))
    return e, x


# FP32 states: Q drifts by single-precision rounding (about 1e-7 over these
# steps), which is also the size of the cross term, so only the bound is checked
@pytest.mark.parametrize("precision,tol", [(shaw_amd.FP64, 1e-12), (shaw_amd.FP32, 2e-6)])
def test_fom_energy_conserved_config2_grid(precision, tol):
    w = workloads.config2()
    nr, nth, dt = w.nr, w.nth, w.dt
    print(f"[gpu] FOM energy {nr}x{nth} dt={dt} precision={precision}", flush=True)
    sys.stdout.flush()
    mat = workloads.layered_material()
    app = shaw_amd.App.structured(nr, nth, mat)
    mv, ms = app.energy_weights()
    ora = oracle.App("", mat, structured=(nr, nth, 0.0, 180.0, 3480.0, 6371.0))
    ptr, ind, val = ora.csr("vp")
    Jv = sp.csr_matrix((val, ind.astype(np.int64), ptr.astype(np.int64)), shape=(ora.nVp, ora.nSp))
    # eight columns (four column blocks: the production kernel shape), short
    # Ricker pulses off the axis that have ended (below 1e-18 of their peak)
    # by step 1800; snapshots every 600 steps (whole six-step launches)
    M, n, every = 8, 3600, 600
    src = [(200.0 + 300.0 * k, 20.0 + 18.0 * k, 16.0 + 2.0 * k, 25.0) for k in range(M)]
    fom = shaw_amd.Fom(app, M, dt, precision=precision, rank1_forcing=False)
    fom.set_sources([app.map_source(d, a) for d, a, p, dl in src])
    fom.set_forcing(shaw_amd.forcing_table(["ricker"] * M, [s[3] for s in src], [s[2] for s in src], dt, n))
    cols = [0, M - 1]
    fom.set_snapshots("vp", every, cols, n // every)
    fom.set_snapshots("sp", every, cols, n // every)
    fom.advance(n)
    fom.sync()
    assert fom.config()["steps_per_launch"] == 6
    V, S = fom.snapshots("vp"), fom.snapshots("sp")
    fom.close()
    assert V.shape == (2, n // every, app.nVp) and S.shape == (2, n // every, app.nSp)
    first = 1800 // every - 1  # sample index of step 1800
    for c in range(2):
        parts = [energy_parts(V[c, k], S[c, k], mv, ms, Jv, dt) for k in range(V.shape[1])]
        q = np.array([e + x for e, x in parts])
        e = np.array([e for e, _ in parts])
        ref = q[-1]
        dev = np.abs(q[first:] / ref - 1.0)
        print(f"[gpu] column {cols[c]}: Q/Q_end - 1 after the forcing {dev}, E spread {np.ptp(e[first:]) / ref:.2e}",
              flush=True)
        assert ref > 0 and np.isfinite(q).all()
        assert dev.max() <= tol
        if precision == shaw_amd.FP64:
            # the field is moving: without its cross term (O(dt^2) of Q at
            # this dt) the energy varies far above the tolerance
            assert np.ptp(e[first:]) / ref > 100 * tol
