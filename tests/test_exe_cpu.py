"""CPU tests of the drop-in executable's input handling. Every case here fails
(or stops) before any GPU call: argument count, YAML parsing of the
reference's own parser-test inputs, validation messages of the reference's
parser sections, the CFL and dispersion checkers and the sampling-frequency
checks (src/shared/parser/*.hpp, src/shared/checkers/*.hpp,
src/kokkos/seismogram.hpp:163-169, state_observer.hpp:130-143)."""
import copy
import os
import subprocess

import pytest
import yaml

import oracle
import shaw_amd
from tests.helpers import MESHES, load_input

GOOD = {"general": {"meshDir": os.path.join(MESHES, "fullMesh21x51"), "dt": 1.0, "finalTime": 100.0,
                    "checkNumericalDispersion": False, "checkCfl": False},
        "io": {"seismogram": {"binary": False, "freq": 10, "receivers": [10.0, 20.0]}},
        "source": {"signal": {"kind": "ricker", "depth": 500.0, "angle": 0.0, "period": 40.0, "delay": 10.0}},
        "material": {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}}


def run(cfg_or_text, tmp_path, args=None):
    path = tmp_path / "input.yaml"
    if isinstance(cfg_or_text, str):
        path.write_text(cfg_or_text)
    else:
        path.write_text(yaml.safe_dump(cfg_or_text))
    cmd = [shaw_amd.EXE_PATH] + (args if args is not None else [str(path)])
    return subprocess.run(cmd, cwd=str(tmp_path), capture_output=True, text=True, timeout=60)


def expect_error(cfg, tmp_path, msg):
    r = run(cfg, tmp_path)
    assert r.returncode != 0, r.stdout[-500:]
    assert msg in r.stderr, r.stderr


def mutate(fn):
    c = copy.deepcopy(GOOD)
    fn(c)
    return c


def test_argument_count(tmp_path):
    r = run(GOOD, tmp_path, args=[])
    assert r.returncode != 0 and "Wrong # of cmd line args" in r.stderr


@pytest.mark.parametrize("i,steps,kind", [(1, 100, "sinusoid"), (2, 100, "ricker"), (3, 100, "ricker"),
                                          (4, 100, "ricker"), (5, 100, "ricker"), (6, 100, "ricker")])
def test_reference_parser_inputs(i, steps, kind, tmp_path):
    """tests/parser/input{1..6}.yaml parse; numSteps = floor(150 / 1.5) = 100
    (tests/parser/main1.cc:19); the run then stops at the missing mesh."""
    text = open(os.path.join(os.path.dirname(__file__), "golden", "inputs", "parser", f"input{i}.yaml")).read()
    r = run(text, tmp_path)
    assert f"numSteps = {steps}" in r.stdout
    assert "mesh_info.dat" in r.stderr
    assert r.returncode != 0


@pytest.mark.parametrize("fn,msg", [
    (lambda c: c["general"].pop("meshDir"), "Empty meshDir"),
    (lambda c: c["general"].pop("dt"), "Empty dt"),
    (lambda c: c["general"].update(dt=-1.0), "Cannot have dt <= 0"),
    (lambda c: c["general"].update(finalTime=0.0), "Cannot have finalT <= 0"),
    (lambda c: c.pop("general"), "General section in yaml input is mandatory!"),
    (lambda c: c.pop("material"), "Cannot find material section in yaml"),
    (lambda c: c["material"].update(kind="granite"), "Unknown material model kind"),
    (lambda c: c.pop("source"), "Cannot find source section"),
    (lambda c: c["source"]["signal"].update(kind="square"), "Invalid source kind"),
    (lambda c: c["source"]["signal"].update(depth=-3.0), "The depth of the source must be a positive number"),
    (lambda c: c["source"]["signal"].update(period=0.0), "Period for signal must be positive"),
    (lambda c: c["source"]["signal"].update(delay=-1.0), "Delay time of the signal must be >=0"),
    (lambda c: c["source"]["signal"].update(angle=-5.0), "The angle of source must be zero or a positive value"),
    (lambda c: c["source"]["signal"].pop("period"), "You must set the period of the signal"),
    (lambda c: c["io"]["seismogram"].pop("freq"), "You must set freq for seismogram"),
    (lambda c: c["io"]["seismogram"].update(freq=7), "Seismogram sampling frequency not a divisor of steps"),
    (lambda c: c["io"].update(snapshotMatrix={"velocity": {"freq": 3}, "stress": {"freq": 10}}),
     "Snapshot Vp frequency not a divisor of steps"),
    (lambda c: c["io"].update(snapshotMatrix={"velocity": {"freq": 10}}), "Cannot find io inputs for stress"),
    (lambda c: c["general"].update(meshDir="/nonexistent/mesh"), "mesh_info.dat"),
    (lambda c: c["general"].update(checkCfl=True, dt=10.0, finalTime=100.0), "CFL error"),
    (lambda c: c["general"].update(checkNumericalDispersion=True), "Numerical dispersion criterion violated"),
    (lambda c: c["source"]["signal"].update(depth=[100.0, 200.0, 300.0], forcingSize=2),
     "must be divisible by fSize"),
    (lambda c: c["source"]["signal"].update(depth=4000.0), "outside radius axis"),
    (lambda c: c.update(rom={"velocity": {"numModes": 0}, "stress": {"numModes": 3}}), "Cannot have vpRomSize <=0"),
])
def test_validation_errors(fn, msg, tmp_path):
    expect_error(mutate(fn), tmp_path, msg)


def test_yaml_block_sequences_and_comments(tmp_path):
    """Block-sequence receivers at the key's indentation and trailing comments
    (demos/demo1/input.yaml style) parse like flow sequences."""
    text = f"""
general:
  meshDir: /nonexistent/dir   # a comment
  dt: 0.25
  finalTime: 2000.0
  checkNumericalDispersion: true
io:
  seismogram:
    binary: false
    freq: 4
    receivers:
    - 5
    - 30
    - 80
source:
  signal:
    kind: ricker
    depth: 640.0
    period: 65.0
    delay: 180.0
material:
  kind: prem
"""
    r = run(text, tmp_path)
    assert "numSteps = 8000" in r.stdout and "mesh_info.dat" in r.stderr


def _reference_complexity(nvp, nsp, nnz_vp, nnz_sp, m, rank2):
    """complexityFom (src/kokkos/fom_complexities.hpp:57-137) over the counts
    of src/shared/complexities/spmv.hpp:56-71, spmm.hpp:58-75 and mult.hpp
    (mult_beta_one: 4 n words, 3 n flops), FP64 values, 8-byte ordinals."""
    def spmm(nnz, rows, k):
        return ((nnz * 16 + rows * 8) + nnz * k * 8 + rows * k * 8 + rows * k * 8) / 1024 / 1024, 2 * nnz * k + rows * k
    k = m if rank2 else 1
    a = spmm(nnz_vp, nvp, k)
    b = (4 * m * 8 / 1024 / 1024, 3 * m) if rank2 else (4 * nvp * 8 / 1024 / 1024, 3 * nvp)
    c = spmm(nnz_sp, nsp, k)
    return a[0] + b[0] + c[0], a[1] + b[1] + c[1]


@pytest.mark.parametrize("mesh", ["fullMesh4x5", "fullMesh21x51"])
@pytest.mark.parametrize("m,rank2", [(1, False), (3, True), (32, True)])
def test_fom_complexity_model(mesh, m, rank2):
    """The byte/flop model shawExe prints (memMB, flops) equals the
    reference's, with the nnz of the oracle's CSR Jacobians; shawExe passes
    2 nsp for both, which is that nnz on every full mesh."""
    o = oracle.App(os.path.join(MESHES, mesh), {"kind": "unilayer", "layer": {"density": [2500.0],
                                                                           "velocity": [5000.0]}})
    nnz_vp = len(o.csr("vp")[2])
    nnz_sp = len(o.csr("sp")[2])
    nvp, nsp = o.nVp, o.nSp
    assert nnz_vp == nnz_sp == 2 * nsp
    got = shaw_amd.fom_complexity(nvp, nsp, nnz_vp, nnz_sp, m, rank2)
    want = _reference_complexity(nvp, nsp, nnz_vp, nnz_sp, m, rank2)
    assert got[0] == pytest.approx(want[0], rel=1e-15) and got[1] == want[1]
    if mesh == "fullMesh4x5" and not rank2:
        # by hand: nnz 62, 20 velocity and 31 stress points
        assert got[1] == (2 * 62 + 20) + 3 * 20 + (2 * 62 + 31)
        assert got[0] * 1024 * 1024 == pytest.approx((62 * 16 + 160 + 62 * 8 + 320) + 640 + (62 * 16 + 248 + 62 * 8 + 496))
