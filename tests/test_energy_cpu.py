"""CPU tests of the FOM energy weights behind the energy-consistent ROM: with
them the oracle's CSR operators (the reference's J_vp, J_sp) satisfy
Mv J_vp = -(Ms J_sp)^T on every coupling of a live point; the only
asymmetric entries are couplings of the axis points, whose velocity rows are
identically zero (the field vanishes on the axis)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import shaw_amd
from shaw_amd import rom_workflow as rw
from tests.helpers import MESHES

LAYERED = {"kind": "layered", "radius": [3480.0, 4100.0, 4700.0, 5200.0, 5600.0, 5900.0, 6150.0, 6371.0],
           "density": [5600.0, 5300.0, 4950.0, 4600.0, 4200.0, 3900.0, 3600.0, 3300.0],
           "velocity": [7300.0, 7000.0, 6600.0, 6200.0, 5700.0, 5200.0, 4800.0, 4400.0]}
PREM = {"kind": "prem"}
BILAYER = {"kind": "bilayer", "layer1": {"density": [2500.0, 0.001], "velocity": [4500.0, 0.0001]},
           "layer2": {"depth": 700.0, "density": [4000.0, 0.0002], "velocity": [6000.0, 0.0003]}}


def csr(ora, dof, shape):
    ptr, ind, val = ora.csr(dof)
    return sp.csr_matrix((val, ind.astype(np.int64), ptr.astype(np.int64)), shape=shape)


def numpy_energy_weights(app: shaw_amd.App):
    """A numpy restatement of the energy weights, which the C++
    shaw_app_energy_weights is compared with: (Mv [nVp], Ms [nSp]) of the FOM operators, from
    the ratios Ms[q] / Mv[p] = -J_vp[p, q] / J_sp[q, p] of the unmerged
    couplings, chained along row 1 and up every column (exact products of the
    coefficients). Points whose couplings are all merged (rows 0 and nr-1,
    the two axis columns) take the weight of the adjacent point, the
    stress points the ratio through an unmerged coupling where one exists."""
    cv, cs, _, _ = app.operators()
    nr, nth = app.nr, app.nth
    rs = 2 * nth - 1
    cN = cv[:, 0].reshape(nr, nth); cS = cv[:, 1].reshape(nr, nth)
    cW = cv[:, 2].reshape(nr, nth); cE = cv[:, 3].reshape(nr, nth)
    # stress coefficient rows: srp(i, j) = j*rs + 2i, stp(i, j) = j*rs + 2i + 1 (j <= nr-2),
    # stp(i, nr-1) = (nr-1)*rs + i
    body = cs[:(nr - 1) * rs].reshape(nr - 1, rs, 2)
    srp0, srp1 = body[:, 0::2, 0], body[:, 0::2, 1]          # [nr-1, nth]: north v, south v
    stp0 = np.vstack([body[:, 1::2, 0], cs[(nr - 1) * rs:, 0][None, :]])  # [nr, nth-1]: west v
    stp1 = np.vstack([body[:, 1::2, 1], cs[(nr - 1) * rs:, 1][None, :]])  # east v
    with np.errstate(divide="ignore", invalid="ignore"):
        # Mv[j+1]/Mv[j] through srp(i, j): (Ms/Mv[j]) / (Ms/Mv[j+1]) = (-cN[j]/srp1) / (-cS[j+1]/srp0)
        up = (cN[:-1, :] / srp1) / (cS[1:, :] / srp0)         # valid for 1 <= j <= nr-3
        # Mv[i+1]/Mv[i] through stp(i, j): (-cE[i]/stp0) / (-cW[i+1]/stp1)
        east = (cE[:, :-1] / stp0) / (cW[:, 1:] / stp1)       # valid for 1 <= i <= nth-3
    lv = np.zeros((nr, nth))
    lv[1, 2:nth - 1] = np.cumsum(np.log(east[1, 1:nth - 2]))
    lv[2:nr - 1, 1:nth - 1] = lv[1, 1:nth - 1][None, :] + np.cumsum(np.log(up[1:nr - 2, 1:nth - 1]), axis=0)
    lv -= lv[nr // 2, nth // 2]
    mv = np.exp(lv)
    with np.errstate(divide="ignore", invalid="ignore"):
        # radial stress from the unmerged side: the south point for rows j >= 1, the north one for row 0
        srp = np.where(np.arange(nr - 1)[:, None] >= 1, -mv[:-1, :] * cN[:-1, :] / srp1, -mv[1:, :] * cS[1:, :] / srp0)
        # CMB and surface rows: N and S of a point there are one stress point, so its merged
        # coefficient (in the N slot) and that stress point's coefficient fix the weight
        mv[0, :] = -srp[0, :] * srp1[0, :] / cN[0, :]
        mv[nr - 1, :] = -srp[nr - 2, :] * srp0[nr - 2, :] / cN[nr - 1, :]
        # axis points: all coefficients vanish (the field is zero there); take the neighbour's weight
        mv[:, 0] = mv[:, 1]; mv[:, nth - 1] = mv[:, nth - 2]
        srp[:, 0] = srp[:, 1]; srp[:, nth - 1] = srp[:, nth - 2]
        # angular stress from the point off the axis
        stp = np.where(np.arange(nth - 1)[None, :] >= 1, -mv[:, :-1] * cE[:, :-1] / stp0, -mv[:, 1:] * cW[:, 1:] / stp1)
    ms = np.empty(app.nSp)
    msb = np.empty((nr - 1, rs))
    msb[:, 0::2] = srp
    msb[:, 1::2] = stp[:nr - 1, :]
    ms[:(nr - 1) * rs] = msb.ravel()
    ms[(nr - 1) * rs:] = stp[nr - 1, :]
    mvf = mv.ravel()
    bad = int((~np.isfinite(ms) | (ms <= 0)).sum() + (~np.isfinite(mvf) | (mvf <= 0)).sum())
    if bad:
        raise RuntimeError(f"energy weights: {bad} points without a usable coupling")
    return mvf, ms


@pytest.mark.parametrize("nr,nth,mat", [(21, 51, PREM), (40, 96, LAYERED), (200, 1000, LAYERED), (5, 7, BILAYER)])
def test_energy_weights_cpp_equal_numpy(nr, nth, mat):
    """shaw_app_energy_weights (C++) against the numpy restatement, within
    1e-13 relative (the two use different log/exp implementations)."""
    app = shaw_amd.App.structured(nr, nth, mat)
    for got, want in zip(app.energy_weights(), numpy_energy_weights(app)):
        np.testing.assert_allclose(got, want, rtol=1e-13)


@pytest.mark.parametrize("nr,nth,mat", [(21, 51, PREM), (40, 96, LAYERED), (33, 70, BILAYER), (5, 7, LAYERED)])
def test_energy_weights_make_operators_adjoint(nr, nth, mat):
    app = shaw_amd.App.structured(nr, nth, mat)
    mv, ms = rw.energy_weights(app)
    assert np.all(mv > 0) and np.all(ms > 0) and np.isfinite(mv).all() and np.isfinite(ms).all()
    ora = oracle.App("", mat, structured=(nr, nth, 0.0, 180.0, 3480.0, 6371.0))
    Jv = csr(ora, "vp", (ora.nVp, ora.nSp))
    Js = csr(ora, "sp", (ora.nSp, ora.nVp))
    # axis velocity rows are zero: those points never change
    axis = np.zeros(ora.nVp, bool)
    axis[0::nth] = True
    axis[nth - 1::nth] = True
    assert abs(Jv[axis]).max() == 0
    L = sp.diags(mv) @ Jv
    R = -(sp.diags(ms) @ Js).T
    D = (L - R).tocoo()
    scale = abs(L).max()
    live = ~axis[D.row]
    assert np.abs(D.data[live]).max() <= 1e-13 * scale
    # and the live block has the same couplings on both sides
    assert abs((L - R)[~axis]).max() <= 1e-13 * scale


def test_energy_weights_from_mesh_files_match_structured():
    mesh = os.path.join(MESHES, "fullMesh21x51")
    a = shaw_amd.App.from_mesh(mesh, PREM)
    b = shaw_amd.App.structured(21, 51, PREM)
    for x, y in zip(rw.energy_weights(a), rw.energy_weights(b)):
        np.testing.assert_allclose(x, y, rtol=1e-13)


@pytest.mark.parametrize("a,s,dt", [(1.0, -1.0, 0.1), (3.0, -0.5, 0.2), (1.0, 1.0, 0.1), (2.0, 0.3, 0.05)])
def test_amplification_radius_closed_form(a, s, dt):
    """1 x 1 operators: the one-step matrix [[1, dt a], [dt s, 1 + dt^2 s a]]
    has determinant 1 and trace 2 + dt^2 s a, so its eigenvalues solve
    l^2 - (2 + dt^2 s a) l + 1 = 0: on the unit circle when -4 < dt^2 s a < 0
    (a stable leapfrog), real with radius (t + sqrt(t^2 - 4)) / 2 when s a > 0."""
    t = 2.0 + dt * dt * s * a
    want = 1.0 if t * t < 4.0 else (abs(t) + np.sqrt(t * t - 4.0)) / 2.0
    got = rw.amplification_radius(np.array([[a]]), np.array([[s]]), dt)
    assert got == pytest.approx(want, rel=1e-14)


def test_fom_discrete_energy_conserved():
    """The oracle's FOM (the reference's leapfrog) conserves
    Q = v^T Mv v + s^T Ms s + dt v^T Mv J_vp s once its forcing has ended, to
    a few ulp; the energy without the cross term, or with its sign flipped,
    is not conserved (so the property pins the cross term, which the GPU
    test of the full config-2 grid relies on)."""
    nr, nth, dt, n, every = 64, 256, 0.5, 1200, 100
    app = shaw_amd.App.structured(nr, nth, LAYERED)
    mv, ms = app.energy_weights()
    ora = oracle.App("", LAYERED, structured=(nr, nth, 0.0, 180.0, 3480.0, 6371.0))
    Jv = csr(ora, "vp", (ora.nVp, ora.nSp))
    # pulses ended (below 1e-30 of their peak) by step 500
    src = [(300.0, 40.0, 60.0, 60.0), (1200.0, 120.0, 50.0, 80.0)]
    out = ora.fom_run(dt, n, src, rank2=True, snap_freq_vp=every, snap_freq_sp=every)
    V, S = out["snaps_vp"], out["snaps_sp"]
    first = 500 // every  # sample of step 600
    for m in range(len(src)):
        e = np.array([V[m, k] @ (mv * V[m, k]) + S[m, k] @ (ms * S[m, k]) for k in range(V.shape[1])])
        x = np.array([dt * (mv * V[m, k]) @ (Jv @ S[m, k]) for k in range(V.shape[1])])
        q = e + x
        assert q[-1] > 0
        assert np.abs(q[first:] / q[-1] - 1).max() <= 1e-13
        assert np.ptp(e[first:]) / q[-1] > 1e-4
        assert np.ptp((e - x)[first:]) / q[-1] > 1e-4
