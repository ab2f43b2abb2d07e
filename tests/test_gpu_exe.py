"""The drop-in executable on the GPU: `shawExe input.yaml` on the reference's
own regression inputs writes the same files as the reference executable;
their contents equal the reference golds to 1e-12 and the oracle bit for bit
(17 significant digits round-trip exactly); binary output and the ROM path
through the executable are checked too."""
import os
import subprocess
import sys

import numpy as np
import pytest
import yaml

import oracle
import shaw_amd
from tests.helpers import GOLDEN, MESHES, fom_golds, load_input, realizations, rel_l2, rel_max
from tests.test_oracle import FOM_TESTS, gold_view

pytestmark = pytest.mark.gpu


def run_exe(cfg: dict, workdir) -> subprocess.CompletedProcess:
    path = os.path.join(str(workdir), "input.yaml")
    with open(path, "w") as f:
        yaml.safe_dump(cfg, f)
    print(f"[gpu] shawExe {path}", flush=True)
    return subprocess.run([shaw_amd.EXE_PATH, path], cwd=str(workdir), capture_output=True, text=True, timeout=300)


def read_ascii(path, header):
    with open(path) as f:
        lines = [ln for ln in f.read().split("\n") if ln.strip()]
    if header:
        shape = [int(x) for x in lines[0].split()]
        data = np.array([[float(x) for x in ln.split()] for ln in lines[1:]])
        return data.reshape(shape[2], shape[0], shape[1]) if len(shape) == 3 else data
    return np.array([[float(x) for x in ln.split()] for ln in lines])


def parse_perf(stdout: str) -> dict:
    """The printPerf block (src/shared/various/print_perf.hpp:57-84)."""
    out = {}
    for ln in stdout.splitlines():
        if " = " in ln:
            k, v = ln.split(" = ", 1)
            try:
                out[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    return out


def check_perf(perf: dict, nvp: int, nsp: int, m: int, rank2: bool, nsteps: int):
    """memMB and flops equal the reference's complexityFom model (nnz = 2 nsp,
    checked against the oracle CSR in test_exe_cpu.py); the per-step times
    order as min <= ave <= max, the bandwidths and GFlop/s the other way, and
    the total is the average times the steps (printed to 4 decimals)."""
    from tests.test_exe_cpu import _reference_complexity
    mb, fl = _reference_complexity(nvp, nsp, 2 * nsp, 2 * nsp, m, rank2)
    scale = 1 if rank2 else m  # rank-1 realizations run together as m columns
    assert perf["memMB"] == pytest.approx(mb * scale, rel=1e-5)
    assert perf["flops"] == pytest.approx(fl * scale, rel=1e-5)
    assert perf["minTime(ms)"] <= perf["aveTime(ms)"] <= perf["maxTime(ms)"]
    assert perf["minBandwidth(GB/s)"] <= perf["aveBandwidth(GB/s)"] <= perf["maxBandwidth(GB/s)"]
    assert perf["minGFlop"] <= perf["aveGFlop"] <= perf["maxGFlop"]
    assert perf["totTime(se)"] == pytest.approx(perf["aveTime(ms)"] * nsteps / 1e3, rel=1e-3, abs=2e-4)


@pytest.mark.parametrize("test", FOM_TESTS)
def test_exe_regression_inputs(test, tmp_path):
    cfg = load_input(test)
    cfg["general"]["meshDir"] = os.path.join(MESHES, cfg["general"]["meshDir"])
    res = run_exe(cfg, tmp_path)
    assert res.returncode == 0, res.stderr + res.stdout[-2000:]
    for key in ("loopTime =", "aveBandwidth(GB/s) =", "aveGFlop =", "aveTime(ms) ="):
        assert key in res.stdout
    golds = fom_golds()
    sig, io, g = cfg["source"]["signal"], cfg["io"], cfg["general"]
    reals = realizations(sig)
    fs = sig.get("forcingSize")
    sets = [reals[i:i + fs] for i in range(0, len(reals), fs)] if fs else [[r] for r in reals]
    ora = oracle.App(cfg["general"]["meshDir"], cfg["material"])
    nsteps = int(g["finalTime"] / g["dt"])
    check_perf(parse_perf(res.stdout), ora.nVp, ora.nSp, fs if fs else len(reals), fs is not None, nsteps)
    for rid, s in enumerate(sets):
        ref = ora.fom_run(g["dt"], nsteps, s, fs is not None, io["seismogram"]["receivers"], io["seismogram"]["freq"],
                          io["snapshotMatrix"]["velocity"]["freq"], io["snapshotMatrix"]["stress"]["freq"], sig["kind"])
        for key in ("snaps_vp", "snaps_sp", "seismogram"):
            got = read_ascii(str(tmp_path / f"{key}_{rid}"), key != "seismogram")
            mine = gold_view(key, got, fs is not None, len(s))
            np.testing.assert_array_equal(mine, ref[key], err_msg=f"{test} {key}_{rid}")
            name = f"{test}/{key}_{rid}"
            if name in golds.files:
                gref = gold_view(key, golds[name], fs is not None, len(s))
                assert rel_l2(mine, gref) <= 1e-12 and rel_max(mine, gref) <= 1e-12, name
    assert os.path.exists(tmp_path / "coords_vp.txt") and os.path.exists(tmp_path / "coords_sp.txt")
    c = np.loadtxt(tmp_path / "coords_vp.txt")
    assert c.shape == (ora.nVp, 2)


def read_binary_2d(path, header=True, shape=None):
    with open(path, "rb") as f:
        if header:
            m, n = (int(x) for x in np.fromfile(f, dtype=np.uint64, count=2))
        else:
            m, n = shape
        return np.fromfile(f, dtype=np.float64, count=m * n).reshape(n, m).T


def test_exe_binary_outputs_and_rom(tmp_path):
    """Binary snapshots/seismograms of a rank-1 run, a POD basis written as
    lsv files, and the ROM run of the executable on that basis: with all modes
    of a 4 x 5 mesh (Phi = I) the ROM seismogram equals the FOM's."""
    mesh = os.path.join(MESHES, "fullMesh4x5")
    mat = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}
    base = {"general": {"meshDir": mesh, "dt": 1.0, "finalTime": 60.0, "checkNumericalDispersion": False,
                        "checkCfl": False},
            "io": {"snapshotMatrix": {"binary": True, "velocity": {"freq": 2, "fileName": "snaps_vp"},
                                      "stress": {"freq": 3, "fileName": "snaps_sp"}},
                   "seismogram": {"binary": True, "freq": 2, "receivers": [45.0, 90.0, 135.0]}},
            "source": {"signal": {"kind": "ricker", "depth": 900.0, "angle": 45.0, "period": 40.0, "delay": 10.0}},
            "material": mat}
    fdir = tmp_path / "fom"
    fdir.mkdir()
    res = run_exe(base, fdir)
    assert res.returncode == 0, res.stderr
    ora = oracle.App(mesh, mat)
    ref = ora.fom_run(1.0, 60, [(900.0, 45.0, 40.0, 10.0)], False, [45.0, 90.0, 135.0], 2, 2, 3)
    np.testing.assert_array_equal(read_binary_2d(fdir / "snaps_vp_0").T, ref["snaps_vp"][0])
    np.testing.assert_array_equal(read_binary_2d(fdir / "snaps_sp_0").T, ref["snaps_sp"][0])
    np.testing.assert_array_equal(read_binary_2d(fdir / "seismogram_0", False, (3, 30)).T, ref["seismogram"][0])
    # identity bases as lsv files (binary, with extents), then the ROM input
    for dof, n in (("vp", ora.nVp), ("sp", ora.nSp)):
        with open(tmp_path / f"lsv_{dof}", "wb") as f:
            np.array([n, n], dtype=np.uint64).tofile(f)
            np.eye(n).T.ravel().tofile(f)
    rom = dict(base)
    rom["io"] = {"seismogram": {"binary": False, "freq": 2, "receivers": [45.0, 90.0, 135.0]}}
    rom["rom"] = {"velocity": {"numModes": ora.nVp, "modesFile": str(tmp_path / "lsv_vp")},
                  "stress": {"numModes": ora.nSp, "modesFile": str(tmp_path / "lsv_sp")}}
    rdir = tmp_path / "rom"
    rdir.mkdir()
    res = run_exe(rom, rdir)
    assert res.returncode == 0, res.stderr + res.stdout[-3000:]
    seis = read_ascii(str(rdir / "seismogram_0"), False)
    assert rel_max(seis, ref["seismogram"][0].T) <= 1e-12


def test_basis_orthonormalize_vs_numpy():
    """phi_out = phi R^-1 with R^T R = phi^T diag(w) phi: phi_out is
    w-orthonormal to 1e-12, spans phi (phi_out R == phi), and R equals
    numpy's Cholesky factor of the weighted Gram matrix."""
    rng = np.random.default_rng(3)
    n, k = 5000, 40
    phi = rng.standard_normal((n, k))
    w = rng.uniform(0.5, 2.0, n)
    out, R = shaw_amd.basis_orthonormalize(phi, w)
    assert np.abs(out.T @ (w[:, None] * out) - np.eye(k)).max() <= 1e-12
    assert np.abs(out @ R - phi).max() <= 1e-12 * np.abs(phi).max()
    L = np.linalg.cholesky(phi.T @ (w[:, None] * phi))
    assert np.abs(R - L.T).max() <= 1e-11 * np.abs(L).max()


def test_exe_energy_projection_full_basis_equals_fom(tmp_path):
    """rom.projection: energy on identity bases of the 4 x 5 mesh: the bases
    become diag(Mv)^-1/2 and diag(Ms)^-1/2, the reduced leapfrog is the FOM's
    in rescaled coordinates, and the ROM seismogram equals the FOM's within
    1e-12 (the axis couplings where the energy pair differs from J_sp act on
    velocities that stay zero)."""
    mesh = os.path.join(MESHES, "fullMesh4x5")
    mat = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}
    base = {"general": {"meshDir": mesh, "dt": 1.0, "finalTime": 60.0, "checkNumericalDispersion": False,
                        "checkCfl": False},
            "io": {"seismogram": {"binary": False, "freq": 2, "receivers": [45.0, 90.0, 135.0]}},
            "source": {"signal": {"kind": "ricker", "depth": 900.0, "angle": 45.0, "period": 40.0, "delay": 10.0}},
            "material": mat}
    ora = oracle.App(mesh, mat)
    ref = ora.fom_run(1.0, 60, [(900.0, 45.0, 40.0, 10.0)], False, [45.0, 90.0, 135.0], 2)
    for dof, n in (("vp", ora.nVp), ("sp", ora.nSp)):
        with open(tmp_path / f"lsv_{dof}", "wb") as f:
            np.array([n, n], dtype=np.uint64).tofile(f)
            np.eye(n).T.ravel().tofile(f)
    rom = dict(base)
    rom["rom"] = {"projection": "energy",
                  "velocity": {"numModes": ora.nVp, "modesFile": str(tmp_path / "lsv_vp")},
                  "stress": {"numModes": ora.nSp, "modesFile": str(tmp_path / "lsv_sp")}}
    rdir = tmp_path / "rom"
    rdir.mkdir()
    res = run_exe(rom, rdir)
    assert res.returncode == 0, res.stderr + res.stdout[-3000:]
    assert "energy-consistent" in res.stdout
    seis = read_ascii(str(rdir / "seismogram_0"), False)
    assert rel_max(seis, ref["seismogram"][0].T) <= 1e-12
