"""Regenerate the golden fixtures under tests/golden/ from the reference's own tests.

Reads (as text) the reference's test directories -- mesh fixtures, input YAML
files, ASCII gold files and the literal gold arrays of the C++ unit tests --
and stores them as data:

  meshes/<name>/            copies of tests/fullMesh4x5, fullMesh21x51 and
                            nonPhysicalFullMesh5x5
  mesh_sha256.json          SHA-256 of every file of the four fixture meshes
  inputs/<test>.yaml        the regression tests' input files
  inputs/parser/*.yaml      the parser unit tests' input files
  fom_golds.npz             every present FOM gold (snaps_vp/snaps_sp/seismogram)
                            as float64, lossless (17 significant digits parse back
                            to the exact double)
  unit_golds.json           literal arrays of tests/{jacobian_vp,jacobian_sp,graphs,
                            coords,stress_labels,meshInfo,seismogram,forcing_rank1}
  prem_table.json           the PREM table various/prem.txt (radius [km], density
                            [g/cm^3], vsv and vsh [km/s]; rows in file order)

Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FOM_TESTS = [
    "fomInnerDomain", "fomNearEarthSurface", "fomNearCmb", "fomSymmetryAxisThetaPi",
    "fomSymmetryAxisThetaZero",
    "multiDepthsForcingRank1", "multiDepthsForcingRank2",
    "multiPeriodsForcingRank1", "multiPeriodsForcingRank2",
    "multiDepthsAndPeriodsForcingRank1", "multiDepthsAndPeriodsForcingRank2",
]


def _brace_list(src: str, name: str):
    """Numbers of the first brace initializer that follows `name` in C++ source."""
    i = src.index(name)
    i = src.index("{", i)
    depth, j = 0, i
    while True:
        c = src[j]
        depth += c == "{"
        depth -= c == "}"
        j += 1
        if depth == 0:
            break
    body = src[i:j]
    # keep row structure for nested lists
    rows = re.findall(r"\{([^{}]*)\}", body[1:-1]) if body.count("{") > 1 else [body[1:-1]]
    num = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")
    return [[float(x) for x in num.findall(r)] for r in rows]


def _gold_array(path: str, header: bool) -> np.ndarray:
    with open(path) as f:
        lines = [ln for ln in f.read().split("\n") if ln.strip()]
    if header:
        shape = [int(x) for x in lines[0].split()]
        data = np.array([[float(x) for x in ln.split()] for ln in lines[1:]], dtype=np.float64)
        if len(shape) == 3:
            s0, s1, s2 = shape
            return data.reshape(s2, s0, s1)  # [forcing][dof][time]
        return data  # [dof][time]
    return np.array([[float(x) for x in ln.split()] for ln in lines], dtype=np.float64)


def main(ref: str) -> None:
    tests = os.path.join(ref, "tests")
    # meshes
    for m in ("fullMesh4x5", "fullMesh21x51", "nonPhysicalFullMesh5x5"):
        dst = os.path.join(HERE, "meshes", m)
        os.makedirs(dst, exist_ok=True)
        for fn in ("mesh_info.dat", "graph_vp.dat", "graph_sp.dat", "coeff_vp.dat"):
            shutil.copyfile(os.path.join(tests, m, fn), os.path.join(dst, fn))
    sha = {}
    for m in ("fullMesh4x5", "fullMesh21x51", "fullMesh50x200", "fullMesh101x201",
              "nonPhysicalFullMesh5x5"):
        for fn in ("mesh_info.dat", "graph_vp.dat", "graph_sp.dat", "coeff_vp.dat"):
            p = os.path.join(tests, m, fn)
            if os.path.exists(p):
                with open(p, "rb") as f:
                    sha[f"{m}/{fn}"] = hashlib.sha256(f.read()).hexdigest()
    with open(os.path.join(HERE, "mesh_sha256.json"), "w") as f:
        json.dump(sha, f, indent=1, sort_keys=True)

    # inputs + golds
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    arrays = {}
    for t in FOM_TESTS:
        shutil.copyfile(os.path.join(tests, t, "input.yaml"), os.path.join(HERE, "inputs", f"{t}.yaml"))
        for fn in sorted(os.listdir(os.path.join(tests, t))):
            if not fn.endswith("_gold"):
                continue
            arrays[f"{t}/{fn[:-5]}"] = _gold_array(os.path.join(tests, t, fn), not fn.startswith("seismogram"))
    np.savez_compressed(os.path.join(HERE, "fom_golds.npz"), **arrays)
    os.makedirs(os.path.join(HERE, "inputs", "parser"), exist_ok=True)
    for i in range(1, 7):  # tests/parser/input{1..6}.yaml
        shutil.copyfile(os.path.join(tests, "parser", f"input{i}.yaml"),
                        os.path.join(HERE, "inputs", "parser", f"input{i}.yaml"))

    # unit golds
    rd = lambda p: open(os.path.join(tests, p)).read()
    jv, js = rd("jacobian_vp/main.cc"), rd("jacobian_sp/main.cc")
    flat = lambda rows: [x for r in rows for x in r]
    unit = {
        "jacobian_vp": {"source": "tests/jacobian_vp/main.cc:12-67",
                        "ind": flat(_brace_list(jv, "goldInd")),
                        "ptr": flat(_brace_list(jv, "goldPtr")),
                        "val": flat(_brace_list(jv, "goldValues"))},
        "jacobian_sp": {"source": "tests/jacobian_sp/main.cc:12-89",
                        "ind": flat(_brace_list(js, "goldInd")),
                        "ptr": flat(_brace_list(js, "goldPtr")),
                        "val": flat(_brace_list(js, "goldValues"))},
        "graphs": {"source": "tests/graphs/main.cc:12-96",
                   "vp": _brace_list(rd("graphs/main.cc"), "goldGVp"),
                   "sp": _brace_list(rd("graphs/main.cc"), "goldGSp")},
        "coords": {"source": "tests/coords/main.cc:11-130 (theta, r[km])",
                   "vp": _brace_list(rd("coords/main.cc"), "goldVp"),
                   "sp": _brace_list(rd("coords/main.cc"), "goldSp")},
        "stress_labels": {"source": "tests/stress_labels/main.cc:25",
                          "labels": flat(_brace_list(rd("stress_labels/main.cc"), "goldLabels"))},
    }
    mi = rd("meshInfo/main.cc")
    num = r"([-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)"
    get = lambda key: float(re.search(re.escape(key) + r"\(\)\s*-\s*" + num, mi).group(1))
    unit["meshInfo"] = {"source": "tests/meshInfo/main.cc:17-63", "mesh": "fullMesh21x51",
                        "dth": get("getAngularSpacing"), "drr": get("getRadialSpacing"),
                        "dthInv": get("getAngularSpacingInverse"),
                        "drrInv": get("getRadialSpacingInverse"),
                        "minArc": get("getMinArc"), "maxArc": get("getMaxArc"),
                        "nr": 21, "nth": 51, "nVp": 1071, "nSp": 2070, "tol": 1e-10}
    sm = rd("seismogram/main.cc")
    unit["seismogram"] = {"source": "tests/seismogram/main.cc:26-30", "mesh": "fullMesh21x51",
                          "receivers": [25, 50, 120, 160],
                          "gids": [int(x) for x in re.findall(r"gids\(\d\) != (\d+)", sm)]}
    fr = rd("forcing_rank1/main.cc")
    unit["forcing_rank1"] = {"source": "tests/forcing_rank1/main.cc:24-25", "mesh": "fullMesh21x51",
                             "depth": 1111.0, "angle": 88.0, "period": 25.0,
                             "maxFreq": float(re.search(r"getMaxFreq\(\) != ([\d.]+)", fr).group(1)),
                             "gid": int(re.search(r"getVpGid\(\) != (\d+)", fr).group(1))}
    with open(os.path.join(HERE, "unit_golds.json"), "w") as f:
        json.dump(unit, f, indent=1)
    prem_table(ref)


def prem_table(ref: str) -> None:
    """Columns 1, 3, 6 and 7 of various/prem.txt (radius, density, vsv, vsh)."""
    rows = []
    with open(os.path.join(ref, "various", "prem.txt")) as f:
        for ln in f:
            x = ln.split()
            if len(x) >= 7:
                rows.append([float(x[0]), float(x[2]), float(x[5]), float(x[6])])
    with open(os.path.join(HERE, "prem_table.json"), "w") as f:
        f.write('{"source": "various/prem.txt (columns radius km, density g/cm^3, vsv km/s, vsh km/s)",\n')
        f.write(' "rows": [\n')
        f.write(",\n".join("  " + json.dumps(r) for r in rows))
        f.write("\n ]}\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
