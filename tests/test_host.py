"""CPU tests of the product's host layer against the oracle and the reference
fixtures: operator coefficients (bit-exact vs. the CSR rows), the in-memory
structured mesh (bit-exact vs. the mesh files), the mesh writer (byte-identical
to the reference's fixture meshes), source/receiver mapping, forcing tables,
and the C ABI export list. No GPU is touched here."""
import hashlib
import json
import os
import re

import numpy as np
import pytest

import oracle
import shaw_amd
from shaw_amd import workloads
from tests.helpers import GOLDEN, MESHES, unit_golds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHYS = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}
UNIT = {"kind": "unilayer", "layer": {"density": [1.0, 0.0], "velocity": [1.0, 0.0]}}
LAYERED = {"kind": "layered", "radius": [3480.0, 4100.0, 4700.0, 5200.0, 5600.0, 5900.0, 6150.0, 6371.0],
           "density": [5600.0, 5300.0, 4950.0, 4600.0, 4200.0, 3900.0, 3600.0, 3300.0],
           "velocity": [7300.0, 7000.0, 6600.0, 6200.0, 5700.0, 5200.0, 4800.0, 4400.0]}
BILAYER = {"kind": "bilayer", "layer1": {"density": [2500.0, 0.001], "velocity": [4500.0, 0.0001]},
           "layer2": {"depth": 700.0, "density": [4000.0, 0.0002], "velocity": [6000.0, 0.0003]}}


def slots_from_csr(app_o, nr, nth):
    """Oracle CSR rows -> (N, S, W, E) slots with merged entries in N / W."""
    ptr, ind, val = app_o.csr("vp")
    out = np.zeros((app_o.nVp, 4))
    for p in range(app_o.nVp):
        i, j = p % nth, p // nth
        vals = val[int(ptr[p]):int(ptr[p + 1])]
        k = 0
        if j == 0 or j == nr - 1:
            out[p, 0] = vals[k]; k += 1
        else:
            out[p, 0], out[p, 1] = vals[k], vals[k + 1]; k += 2
        if i == 0 or i == nth - 1:
            out[p, 2] = vals[k]; k += 1
        else:
            out[p, 2], out[p, 3] = vals[k], vals[k + 1]; k += 2
        assert k == len(vals)
    ptr, ind, val = app_o.csr("sp")
    return out, val.reshape(-1, 2)


@pytest.mark.parametrize("mesh,mat", [("fullMesh4x5", UNIT), ("fullMesh21x51", PHYS),
                                      ("fullMesh21x51", LAYERED), ("fullMesh21x51", BILAYER),
                                      ("fullMesh21x51", {"kind": "prem"}),
                                      ("nonPhysicalFullMesh5x5", UNIT)])
def test_operators_bit_exact_vs_oracle_csr(mesh, mat):
    d = os.path.join(MESHES, mesh)
    a = shaw_amd.App.from_mesh(d, mat)
    o = oracle.App(d, mat)
    cv, cs, ri, g = a.operators()
    ov, osp = slots_from_csr(o, a.nr, a.nth)
    np.testing.assert_array_equal(cv, ov)
    np.testing.assert_array_equal(cs, osp)
    ori, og = o.material()
    np.testing.assert_array_equal(ri, ori)
    np.testing.assert_array_equal(g, og)
    assert (a.minVs, a.maxVs) == (o.minVs, o.maxVs)


@pytest.mark.parametrize("mat", [workloads.layered_material(), {"kind": "prem"}, BILAYER],
                         ids=["layered", "prem", "bilayer"])
def test_material_arrays_bit_exact_vs_oracle_config1_mesh(mat):
    """rho^-1 at every velocity point and the shear modulus at every stress
    point of the 200 x 1000 mantle mesh: product == oracle bit for bit (the
    oracle's materials are pinned by tests/test_oracle_pins.py)."""
    a = shaw_amd.App.structured(200, 1000, mat)
    o = oracle.App("", mat, structured=(200, 1000, 0.0, 180.0, 3480.0, 6371.0))
    _, _, ri, g = a.operators()
    ori, og = o.material()
    np.testing.assert_array_equal(ri, ori)
    np.testing.assert_array_equal(g, og)


def test_jacobian_gold_via_product():
    a = shaw_amd.App.from_mesh(os.path.join(MESHES, "fullMesh4x5"), UNIT)
    cv, cs, _, _ = a.operators()
    gold = unit_golds()["jacobian_vp"]
    # product slots, flattened in CSR order (dropping merged zeros), equal the gold values
    flat = []
    for p in range(a.nVp):
        i, j = p % a.nth, p // a.nth
        flat += [cv[p, 0]] if j in (0, a.nr - 1) else [cv[p, 0], cv[p, 1]]
        flat += [cv[p, 2]] if i in (0, a.nth - 1) else [cv[p, 2], cv[p, 3]]
    np.testing.assert_allclose(flat, gold["val"], rtol=4e-15, atol=1e-300)
    np.testing.assert_allclose(cs.ravel(), unit_golds()["jacobian_sp"]["val"], rtol=4e-15, atol=1e-300)


@pytest.mark.parametrize("mesh", ["fullMesh4x5", "fullMesh21x51", "nonPhysicalFullMesh5x5"])
def test_structured_equals_files(mesh):
    d = os.path.join(MESHES, mesh)
    mi = dict(l.split() for l in open(os.path.join(d, "mesh_info.dat")))
    mat = UNIT if mesh != "fullMesh21x51" else LAYERED
    a = shaw_amd.App.from_mesh(d, mat)
    b = shaw_amd.App.structured(int(mi["nr"]), int(mi["nth"]), mat, float(mi["thL"]), float(mi["thR"]),
                                float(mi["rCmb"]), float(mi["rSurf"]))
    assert (a.dth, a.drr) == (b.dth, b.drr)
    for dof in ("vp", "sp"):
        np.testing.assert_array_equal(a.coords(dof), b.coords(dof))
    for x, y in zip(a.operators(), b.operators()):
        np.testing.assert_array_equal(x, y)


def test_mesh_writer_byte_identical(tmp_path):
    sha = json.load(open(os.path.join(GOLDEN, "mesh_sha256.json")))
    cases = {"fullMesh4x5": (4, 5, 3480.0, 6371.0), "fullMesh21x51": (21, 51, 3480.0, 6371.0),
             "fullMesh50x200": (50, 200, 3480.0, 6371.0), "fullMesh101x201": (101, 201, 3480.0, 6371.0),
             "nonPhysicalFullMesh5x5": (5, 5, 1.0, 2.0)}
    for name, (nr, nth, rc, rs) in cases.items():
        d = str(tmp_path / name)
        shaw_amd.mesh_write(d, nr, nth, 0.0, 180.0, rc, rs)
        for fn in ("mesh_info.dat", "graph_vp.dat", "graph_sp.dat", "coeff_vp.dat"):
            key = f"{name}/{fn}"
            if key in sha:
                got = hashlib.sha256(open(os.path.join(d, fn), "rb").read()).hexdigest()
                assert got == sha[key], key


def test_mapping_matches_full_scan_oracle():
    d = os.path.join(MESHES, "fullMesh21x51")
    a = shaw_amd.App.from_mesh(d, PHYS)
    o = oracle.App(d, PHYS)
    rng = np.random.default_rng(7)
    for _ in range(300):
        depth, ang = float(rng.uniform(1, 2890)), float(rng.uniform(0, 180))
        assert a.map_source(depth, ang) == o.map_source(depth, ang)
        assert a.map_receiver(ang) == o.map_receiver(ang)
    for depth in (434.0, 720.0, 580.0, 1111.0, 2400.0, 300.0, 500.0):
        for ang in (0.0, 88.0, 180.0, 45.0, 90.0):
            assert a.map_source(depth, ang) == o.map_source(depth, ang)
    g = unit_golds()
    assert [a.map_receiver(x) for x in g["seismogram"]["receivers"]] == g["seismogram"]["gids"]
    assert a.map_source(g["forcing_rank1"]["depth"], g["forcing_rank1"]["angle"]) == g["forcing_rank1"]["gid"]


def test_mapping_structured_large_mesh():
    # exact grid points and mid-cell ties on a finer mesh; full scan by brute force
    a = shaw_amd.App.structured(101, 201, PHYS)
    c = a.coords("vp")
    th, r = c[:, 0], c[:, 1]
    rng = np.random.default_rng(11)
    for _ in range(60):
        depth, ang = float(rng.uniform(1, 2880)), float(rng.uniform(0.5, 179.5))
        R = (6371.0 - depth) * 1000.0
        T = ang * (np.pi / 180.0)
        dist = np.sqrt((r * r) + (R * R) - 2.0 * r * R * np.cos(T - th))
        best = np.flatnonzero(dist == dist.min())[-1]
        assert a.map_source(depth, ang) == best


def test_forcing_table_matches_oracle_signal():
    tab = shaw_amd.forcing_table(["ricker", "sinusoid", "gaussDer"], [10.0, 10.0, 30.0], [40.0, 25.0, 60.0], 0.25, 400)
    for s in range(0, 400, 37):
        for k, kind in enumerate(["ricker", "sinusoid", "gaussDer"]):
            exp = oracle.signal(kind, [10.0, 10.0, 30.0][k], [40.0, 25.0, 60.0][k], s * 0.25)
            assert tab[s, k] == exp


def test_errors_are_reported():
    with pytest.raises(shaw_amd.ShawError, match="mesh_info.dat"):
        shaw_amd.App.from_mesh("/nonexistent/dir", PHYS)
    with pytest.raises(shaw_amd.ShawError, match="Unknown material"):
        shaw_amd.App.from_mesh(os.path.join(MESHES, "fullMesh4x5"), {"kind": "granite"})
    with pytest.raises(shaw_amd.ShawError, match="PREM"):
        shaw_amd.App.from_mesh(os.path.join(MESHES, "nonPhysicalFullMesh5x5"), {"kind": "prem"})
    a = shaw_amd.App.from_mesh(os.path.join(MESHES, "fullMesh21x51"), PHYS)
    with pytest.raises(shaw_amd.ShawError, match="outside"):
        a.map_source(4000.0, 10.0)


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "shaw_amd.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(shaw_\w+)\s*\(", hdr, flags=re.M))
    assert len(declared) > 30
    L = shaw_amd.lib()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing


@pytest.mark.parametrize("n", [1, 2, 7, 64, 300])
def test_pod_eigensolver_against_numpy(n):
    """The host eigensolver of the POD (tridiagonal symmetric QR): eigenvalues of a
    graded Gram matrix to ~eps ||G||, orthonormal eigenvectors, small residual."""
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n + 5, n))
    X[:, n // 2:] *= 1e-5
    G = X.T @ X
    w, V = shaw_amd.debug_symeig(G)
    ref = np.linalg.eigvalsh(G)
    scale = np.abs(ref).max()
    assert np.abs(np.sort(w) - ref).max() <= 1e-13 * scale
    assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-13
    assert np.abs(G @ V - V * w).max() <= 1e-13 * scale


def test_pod_eigensolver_rank_deficient_gram():
    """A Gram matrix of rank 40 (plus rounding-level noise) of size 600: the
    cluster of noise eigenvalues must not stall the QR iteration, and the
    resolved eigenvalues stay accurate to ~eps ||G||."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((700, 40)) @ rng.standard_normal((40, 600))
    X[:, 300:] *= 1e-7
    G = X.T @ X
    w, V = shaw_amd.debug_symeig(G)
    ref = np.linalg.eigvalsh(G)
    scale = np.abs(ref).max()
    assert np.abs(np.sort(w) - ref).max() <= 1e-12 * scale
    assert np.abs(G @ V - V * w).max() <= 1e-12 * scale



@pytest.mark.parametrize("n", [3, 50, 257])
def test_pod_eigensolver_closed_form(n):
    """The second-difference matrix tridiag(-1, 2, -1), rotated by a random
    orthogonal matrix so the tridiagonalization has work to do: eigenvalues
    2 - 2 cos(k pi / (n + 1)), k = 1..n (all distinct), and the eigenvectors
    rotate back to the discrete sines sin(j k pi / (n + 1)) up to sign."""
    T = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    Qr, _ = np.linalg.qr(np.random.default_rng(n).standard_normal((n, n)))
    A = Qr @ T @ Qr.T
    A = 0.5 * (A + A.T)
    w, V = shaw_amd.debug_symeig(A)
    k = np.arange(1, n + 1)
    lam = 2.0 - 2.0 * np.cos(k * np.pi / (n + 1))
    order = np.argsort(w)
    assert np.abs(w[order] - lam).max() <= 1e-13 * 4.0
    sines = np.sin(np.outer(k, k) * np.pi / (n + 1))
    sines /= np.linalg.norm(sines, axis=0)
    U = Qr.T @ V[:, order]
    assert np.abs(np.abs(U) - np.abs(sines)).max() <= 1e-10
