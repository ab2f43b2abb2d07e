"""Pins the CPU oracle against the reference's own golds and against closed forms.

The oracle (oracle/shaw_oracle.cpp) is what every GPU parity test compares with,
so it is validated here first: CSR Jacobians, graphs, coordinates, labels,
mesh info, source/receiver gids (C++ unit-test literals), every present FOM
state/seismogram gold (rank-1 and rank-2), the Ricker closed form, and ROM
invariants (dense projection, full basis == FOM, rank-2 column == rank-1 run).
"""
import math
import os

import numpy as np
import pytest

import oracle
from tests.helpers import (MESHES, as_list, fom_golds, load_input, realizations, rel_l2,
                           rel_max, unit_golds)

UNIT_MAT = {"kind": "unilayer", "layer": {"density": [1.0, 0.0], "velocity": [1.0, 0.0]}}
PHYS_MAT = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}


@pytest.fixture(scope="module")
def app45():
    return oracle.App(os.path.join(MESHES, "fullMesh4x5"), UNIT_MAT)


@pytest.fixture(scope="module")
def app21():
    return oracle.App(os.path.join(MESHES, "fullMesh21x51"), PHYS_MAT)


@pytest.mark.parametrize("dof", ["vp", "sp"])
def test_jacobian_csr_gold(app45, dof):
    g = unit_golds()["jacobian_" + dof]
    ptr, ind, val = app45.csr(dof)
    assert list(ptr) == [int(x) for x in g["ptr"]]
    assert list(ind) == [int(x) for x in g["ind"]]
    # the reference test allows 1e-13 absolute; the literals are 17 digits
    np.testing.assert_allclose(val, g["val"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(val, g["val"], rtol=4e-15, atol=1e-300)


def test_graphs_gold(app45):
    g = unit_golds()["graphs"]
    gv, _, _ = app45.graph("vp")
    gs, _, lab = app45.graph("sp")
    assert (gv[:, 0] == np.arange(app45.nVp)).all()
    assert gv[:, 1:].tolist() == [[int(x) for x in r] for r in g["vp"]]
    assert gs[:, 1:].tolist() == [[int(x) for x in r] for r in g["sp"]]
    assert lab.tolist() == [int(x) for x in unit_golds()["stress_labels"]["labels"]]


def test_coords_gold():
    app = oracle.App(os.path.join(MESHES, "nonPhysicalFullMesh5x5"), UNIT_MAT)
    g = unit_golds()["coords"]
    for dof in ("vp", "sp"):
        _, c, _ = app.graph(dof)
        gold = np.array(g[dof])
        np.testing.assert_allclose(c[:, 0], gold[:, 0], rtol=0, atol=1e-13)
        np.testing.assert_allclose(c[:, 1], 1.0 / (gold[:, 1] * 1000.0), rtol=0, atol=1e-13)


def test_mesh_info_gold(app21):
    g = unit_golds()["meshInfo"]
    assert (app21.nr, app21.nth, app21.nVp, app21.nSp) == (g["nr"], g["nth"], g["nVp"], g["nSp"])
    assert abs(app21.dth - g["dth"]) <= g["tol"]
    assert abs(app21.drr - g["drr"]) <= g["tol"]
    assert abs(1.0 / app21.dth - g["dthInv"]) <= g["tol"]
    assert abs(1.0 / app21.drr - g["drrInv"]) <= g["tol"]
    assert abs(app21.dth * app21.bounds[2] - g["minArc"]) <= g["tol"]
    assert abs(app21.dth * app21.bounds[3] - g["maxArc"]) <= g["tol"]
    # closed-form counts: nSp = (nr-1) nth + nr (nth-1); nnz = 4 nVp - 2 nr - 2 nth = 2 nSp
    assert app21.nSp == (app21.nr - 1) * app21.nth + app21.nr * (app21.nth - 1)
    assert app21.nnzVp == 4 * app21.nVp - 2 * app21.nr - 2 * app21.nth == 2 * app21.nSp


def test_receiver_and_source_gids(app21):
    s = unit_golds()["seismogram"]
    assert [app21.map_receiver(a) for a in s["receivers"]] == s["gids"]
    f = unit_golds()["forcing_rank1"]
    assert app21.map_source(f["depth"], f["angle"]) == f["gid"]
    assert 1.0 / f["period"] == f["maxFreq"]


def test_source_on_axis_is_remapped(app21):
    # a source at theta = 0 maps to theta = dth (map_point_source...:94-105)
    g = app21.map_source(640.0, 0.0)
    assert g % app21.nth == 1
    g = app21.map_source(640.0, 180.0)
    assert g % app21.nth == app21.nth - 2
    with pytest.raises(RuntimeError):
        app21.map_source(5000.0, 10.0)  # below the CMB


def test_ricker_closed_form():
    # peak 1 at t = delay; zeros at |t - delay| = 1/(pi f sqrt 2)
    T, d = 40.0, 10.0
    assert oracle.signal("ricker", d, T, d) == 1.0
    tz = d + T / (math.pi * math.sqrt(2.0))
    assert abs(oracle.signal("ricker", d, T, tz)) < 1e-15
    t = 23.0
    x = (math.pi / T) ** 2 * (t - d) ** 2
    assert abs(oracle.signal("ricker", d, T, t) - (1 - 2 * x) * math.exp(-x)) < 1e-15
    # sinusoid holds a constant after one period (signal.hpp:100-116)
    assert oracle.signal("sinusoid", d, T, 50.0) == oracle.signal("sinusoid", d, T, 75.0)


FOM_TESTS = ["fomInnerDomain", "fomNearEarthSurface",
             "multiDepthsForcingRank1", "multiDepthsForcingRank2",
             "multiPeriodsForcingRank1", "multiPeriodsForcingRank2",
             "multiDepthsAndPeriodsForcingRank1", "multiDepthsAndPeriodsForcingRank2"]


def run_oracle_like_exe(test):
    """Run the oracle the way shawExe runs the input: per realization (rank-1)
    or per set of forcingSize columns (rank-2). Yields (run_id, output)."""
    cfg = load_input(test)
    g, io, sig = cfg["general"], cfg.get("io", {}), cfg["source"]["signal"]
    app = oracle.App(os.path.join(MESHES, g["meshDir"]), cfg["material"])
    nsteps = int(g["finalTime"] / g["dt"])
    seis = io.get("seismogram", {})
    snap = io.get("snapshotMatrix", {})
    reals = realizations(sig)
    fs = sig.get("forcingSize")
    sets = [reals[i:i + fs] for i in range(0, len(reals), fs)] if fs else [[r] for r in reals]
    for rid, s in enumerate(sets):
        out = app.fom_run(g["dt"], nsteps, s, fs is not None, seis.get("receivers", []),
                          seis.get("freq", 0), snap.get("velocity", {}).get("freq", 0),
                          snap.get("stress", {}).get("freq", 0), sig["kind"])
        yield rid, len(s), fs is not None, out


def gold_view(name, arr, rank2, M):
    """Reference gold -> oracle layout ([M][T][n] snaps, [M][T][R] seismogram)."""
    if name.startswith("seismogram"):
        a = arr.reshape(M, -1, arr.shape[1]) if rank2 else arr[None]
        return a.transpose(0, 2, 1)
    return (arr if rank2 else arr[None]).transpose(0, 2, 1)


@pytest.mark.parametrize("test", FOM_TESTS)
def test_fom_golds(test):
    golds = fom_golds()
    checked = 0
    for rid, M, rank2, out in run_oracle_like_exe(test):
        for key in ("snaps_vp", "snaps_sp", "seismogram"):
            name = f"{test}/{key}_{rid}"
            if name not in golds.files:
                continue
            ref = gold_view(key, golds[name], rank2, M)
            assert out[key].shape == ref.shape, name
            for m in range(M):
                assert rel_l2(out[key][m], ref[m]) <= 1e-12, name
                assert rel_max(out[key][m], ref[m]) <= 1e-12, name
            checked += 1
    assert checked > 0


def test_rank2_column_equals_rank1_run():
    # the same (depth, period) appears as a rank-1 run and inside a rank-2 set
    r1 = dict((rid, out) for rid, _, _, out in run_oracle_like_exe("multiDepthsForcingRank1"))
    r2 = dict((rid, out) for rid, _, _, out in run_oracle_like_exe("multiDepthsForcingRank2"))
    for j in range(3):
        for key in ("snaps_vp", "snaps_sp", "seismogram"):
            a, b = r1[j][key][0], r2[0][key][j]
            # rank-2 adds the forcing as (rhoInv f) dt, rank-1 as (dt rhoInv) f: ulp-level only
            assert rel_max(a, b) <= 1e-13


def test_rom_projection_dense(app45):
    rng = np.random.default_rng(3)
    phiV = rng.standard_normal((app45.nVp, 5))
    phiS = rng.standard_normal((app45.nSp, 7))
    ArV, ArS = app45.rom_project(phiV, phiS)
    Jv, Js = app45.dense("vp"), app45.dense("sp")
    np.testing.assert_allclose(ArV, phiV.T @ (Jv @ phiS), rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(ArS, phiS.T @ (Js @ phiV), rtol=1e-12, atol=1e-18)


def test_rom_full_basis_equals_fom(app21):
    """Phi = I: the reduced leapfrog is the FOM leapfrog (rom_run_rank_one.hpp vs fom_run.hpp)."""
    nV, nS = app21.nVp, app21.nSp
    ArV, ArS = app21.rom_project(np.eye(nV), np.eye(nS))
    np.testing.assert_array_equal(ArV, app21.dense("vp"))
    src = [(434.0, 0.0, 40.0, 10.0), (720.0, 30.0, 45.0, 10.0)]
    nsteps, dt = 60, 1.0
    fom = app21.fom_run(dt, nsteps, src, True, [2, 10, 20], 10)
    ri, _ = app21.material()
    M = len(src)
    B = np.zeros((nV, M))
    ftab = np.zeros((nsteps, M))
    for m, (d, a, p, dl) in enumerate(src):
        g = app21.map_source(d, a)
        B[g, m] = ri[g]
        ftab[:, m] = [oracle.signal("ricker", dl, p, (i - 1) * dt) for i in range(1, nsteps + 1)]
    C = np.zeros((3, nV))
    for r, ang in enumerate([2, 10, 20]):
        C[r, app21.map_receiver(ang)] = 1.0
    rom = oracle.rom_run(ArV, ArS, B, ftab, dt, nsteps, C, 10)
    assert rel_max(rom["qv"], fom["final_vp"]) <= 1e-12
    assert rel_max(rom["qs"], fom["final_sp"]) <= 1e-12
    assert rel_max(rom["seismogram"], fom["seismogram"]) <= 1e-12


@pytest.mark.parametrize("mesh,shape", [("fullMesh4x5", (4, 5, 0.0, 180.0, 3480.0, 6371.0)),
                                        ("fullMesh21x51", (21, 51, 0.0, 180.0, 3480.0, 6371.0)),
                                        ("nonPhysicalFullMesh5x5", (5, 5, 0.0, 180.0, 1.0, 2.0))])
def test_structured_oracle_equals_files(mesh, shape):
    """The oracle's in-memory restatement of the mesh generator reproduces the
    reference's fixture mesh files exactly (graphs, coordinates, CSR values)."""
    a = oracle.App(os.path.join(MESHES, mesh), UNIT_MAT)
    b = oracle.App("", UNIT_MAT, structured=shape)
    for dof in ("vp", "sp"):
        for x, y in zip(a.graph(dof), b.graph(dof)):
            np.testing.assert_array_equal(x, y)
        for x, y in zip(a.csr(dof), b.csr(dof)):
            np.testing.assert_array_equal(x, y)
    assert (a.dth, a.drr) == (b.dth, b.drr)
