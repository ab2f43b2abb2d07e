"""GPU tests of the Galerkin ROM: MFMA lane layouts, FP64 projection against
the oracle, the batched reduced leapfrog against the oracle (FP64) and its
BF16x3 split-precision variant (error stated), the full-basis invariant
(Phi = I: ROM == FOM), uneven horizons, slot reuse, and alone-vs-mixed
equality of a column."""
import os
import sys

import numpy as np
import pytest

import oracle
import shaw_amd
from tests.helpers import MESHES, rel_l2, rel_max

pytestmark = pytest.mark.gpu

PHYS = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bf16_round(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_mfma_f64_layout():
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(64), rng.standard_normal(64)
    d = shaw_amd.debug_mfma(0, a, b)
    A = np.array([[a[i + 16 * k] for k in range(4)] for i in range(16)])
    B = np.array([[b[j + 16 * k] for j in range(16)] for k in range(4)])
    D = A @ B
    exp = np.array([[D[l // 16 + 4 * r, l % 16] for r in range(4)] for l in range(64)])
    if not np.allclose(d, exp, rtol=1e-13, atol=1e-13):
        say(f"[gpu] f64 mfma raw output:\n{d}")
    np.testing.assert_allclose(d, exp, rtol=1e-13, atol=1e-13)


def test_mfma_bf16_layout():
    rng = np.random.default_rng(1)
    a = bf16_round(rng.standard_normal((64, 4)))
    b = bf16_round(rng.standard_normal((64, 4)))
    d = shaw_amd.debug_mfma(1, a, b)
    A = np.array([[a[i + 16 * (k // 4), k % 4] for k in range(16)] for i in range(16)])
    B = np.array([[b[j + 16 * (k // 4), k % 4] for j in range(16)] for k in range(16)])
    D = A @ B
    exp = np.array([[D[4 * (l // 16) + r, l % 16] for r in range(4)] for l in range(64)])
    if not np.allclose(d, exp, rtol=1e-5, atol=1e-5):
        say(f"[gpu] bf16 mfma raw output:\n{d}")
    np.testing.assert_allclose(d, exp, rtol=1e-5, atol=1e-5)


def test_mfma_bf16_k32_layout():
    """v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) + e] and
    B[8 (l / 16) + e][l % 16], e < 8; D as the 16x16x16 form."""
    rng = np.random.default_rng(2)
    a = bf16_round(rng.standard_normal((64, 8)))
    b = bf16_round(rng.standard_normal((64, 8)))
    d = shaw_amd.debug_mfma(2, a, b)
    A = np.array([[a[i + 16 * (k // 8), k % 8] for k in range(32)] for i in range(16)])
    B = np.array([[b[j + 16 * (k // 8), k % 8] for j in range(16)] for k in range(32)])
    D = A @ B
    exp = np.array([[D[4 * (l // 16) + r, l % 16] for r in range(4)] for l in range(64)])
    if not np.allclose(d, exp, rtol=1e-5, atol=1e-5):
        say(f"[gpu] bf16 k32 mfma raw output:\n{d}")
    np.testing.assert_allclose(d, exp, rtol=1e-5, atol=1e-5)


@pytest.fixture(scope="module")
def pod21():
    """A POD basis of the 21 x 51 mesh from oracle FOM snapshots (numpy SVD)."""
    mesh = os.path.join(MESHES, "fullMesh21x51")
    ora = oracle.App(mesh, PHYS)
    src = [(434.0, 0.0, 40.0, 10.0), (720.0, 0.0, 45.0, 10.0), (580.0, 0.0, 50.0, 10.0)]
    out = ora.fom_run(1.0, 200, src, True, [], 0, 2, 2)
    Sv = out["snaps_vp"].reshape(-1, ora.nVp).T
    Ss = out["snaps_sp"].reshape(-1, ora.nSp).T
    Uv = np.linalg.svd(Sv, full_matrices=False)[0]
    Us = np.linalg.svd(Ss, full_matrices=False)[0]
    return mesh, ora, Uv, Us


def rom_inputs(app_o, phi_v, src, nsteps, dt, receivers):
    ri, _ = app_o.material()
    M = len(src)
    B = np.zeros((phi_v.shape[1], M))
    ftab = np.zeros((nsteps, M))
    for m, (d, a, p, dl) in enumerate(src):
        g = app_o.map_source(d, a)
        B[:, m] = ri[g] * phi_v[g, :]
        ftab[:, m] = [oracle.signal("ricker", dl, p, (i - 1) * dt) for i in range(1, nsteps + 1)]
    C = np.array([phi_v[app_o.map_receiver(r), :] for r in receivers])
    return B, ftab, C


def test_projection_vs_oracle(pod21):
    mesh, ora, Uv, Us = pod21
    app = shaw_amd.App.from_mesh(mesh, PHYS)
    pv, ps = Uv[:, :40], Us[:, :48]
    a_v, a_s = shaw_amd.rom_project(app, pv, ps)
    o_v, o_s = ora.rom_project(pv, ps)
    assert rel_max(a_v, o_v) <= 1e-12 and rel_max(a_s, o_s) <= 1e-12


@pytest.mark.parametrize("K", [40, 64])
def test_rom_fp64_vs_oracle_and_bf16x3(pod21, K):
    say(f"[gpu] ROM K={K}")
    mesh, ora, Uv, Us = pod21
    app = shaw_amd.App.from_mesh(mesh, PHYS)
    pv, ps = Uv[:, :K], Us[:, :K]
    arv, ars = shaw_amd.rom_project(app, pv, ps)
    rng = np.random.default_rng(K)
    src = [(float(rng.uniform(300, 900)), 0.0, float(rng.uniform(35, 60)), 10.0) for _ in range(21)]
    rec = [2.0, 10.0, 20.0, 35.0]
    n, dt = 300, 1.0
    B, ftab, C = rom_inputs(ora, pv, src, n, dt, rec)
    ref = oracle.rom_run(arv, ars, B, ftab, dt, n, C, 10)
    got = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, 10, shaw_amd.ROM_FP64)
    for key in ("qv", "qs", "seismogram"):
        assert rel_l2(got[key], ref[key]) <= 1e-11, key
    bf = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, 10, shaw_amd.ROM_BF16X3)
    err = max(rel_l2(bf["seismogram"][m], ref["seismogram"][m]) for m in range(len(src)))
    say(f"[gpu] ROM K={K}: BF16x3 seismogram max rel L2 error vs FP64 ROM = {err:.3e}")
    assert err <= 1e-4


@pytest.mark.parametrize("kv,ks", [(128, 128), (256, 256), (512, 512), (96, 200), (33, 17)])
def test_rom_sizes_vs_oracle(kv, ks):
    """Sizes that exercise 1..4 row tiles per wave, padding of the operator
    groups, kv != ks and receivers spanning several tiles: FP64 within 1e-11
    of the oracle, BF16x3 within 1e-4 (error printed)."""
    say(f"[gpu] ROM sizes kv={kv} ks={ks}")
    rng = np.random.default_rng(kv * 1000 + ks)
    A = rng.standard_normal((kv, ks)) * (0.5 / np.sqrt(max(kv, ks)))
    arv, ars = A, -A.T.copy()  # skew pair: the reduced leapfrog stays bounded
    M, n, dt, every = 37, 120, 0.5, 7
    B = rng.standard_normal((kv, M))
    ftab = rng.standard_normal((n, M))
    C = rng.standard_normal((19, kv))
    ref = oracle.rom_run(arv, ars, B, ftab, dt, n, C, every)
    got = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, every, shaw_amd.ROM_FP64)
    for key in ("qv", "qs", "seismogram"):
        assert rel_l2(got[key], ref[key]) <= 1e-11, key
    bf = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, every, shaw_amd.ROM_BF16X3)
    err = max(rel_l2(bf["seismogram"][m], ref["seismogram"][m]) for m in range(M))
    b4 = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, every, shaw_amd.ROM_BF16X4)
    err4 = max(rel_l2(b4["seismogram"][m], ref["seismogram"][m]) for m in range(M))
    say(f"[gpu] ROM kv={kv} ks={ks}: seismogram max rel L2 error vs FP64: BF16x3 {err:.3e}, BF16x4 {err4:.3e}")
    assert err <= 1e-4
    assert err4 <= err and err4 <= 2e-5


def test_rom_size_limit():
    A = np.zeros((513, 8))
    with pytest.raises(shaw_amd.ShawError, match="512"):
        shaw_amd.Rom(A, A.T.copy(), 0.5, shaw_amd.ROM_FP64, None, max_cols=16)


def test_full_basis_rom_equals_fom():
    """Phi = I on the 4 x 5 mesh: the reduced leapfrog is the FOM leapfrog."""
    mesh = os.path.join(MESHES, "fullMesh4x5")
    mat = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}
    app = shaw_amd.App.from_mesh(mesh, mat)
    ora = oracle.App(mesh, mat)
    arv, ars = shaw_amd.rom_project(app, np.eye(app.nVp), np.eye(app.nSp))
    np.testing.assert_array_equal(arv, ora.dense("vp"))
    np.testing.assert_array_equal(ars, ora.dense("sp"))
    src = [(900.0, 45.0, 40.0, 10.0), (1800.0, 90.0, 30.0, 5.0)]
    n, dt = 80, 1.0
    B, ftab, C = rom_inputs(ora, np.eye(app.nVp), src, n, dt, [45.0, 90.0, 135.0])
    fom = ora.fom_run(dt, n, src, True, [45.0, 90.0, 135.0], 4)
    rom = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, 4)
    assert rel_max(rom["qv"], fom["final_vp"]) <= 1e-12
    assert rel_max(rom["qs"], fom["final_sp"]) <= 1e-12
    assert rel_max(rom["seismogram"], fom["seismogram"]) <= 1e-12


def test_uneven_horizons_slot_reuse_and_mix(pod21):
    mesh, ora, Uv, Us = pod21
    app = shaw_amd.App.from_mesh(mesh, PHYS)
    pv, ps = Uv[:, :32], Us[:, :32]
    arv, ars = shaw_amd.rom_project(app, pv, ps)
    src = [(434.0, 0.0, 40.0, 10.0), (720.0, 0.0, 45.0, 10.0), (580.0, 0.0, 50.0, 10.0)]
    n = 160
    B, ftab, C = rom_inputs(ora, pv, src, n, 1.0, [5.0, 15.0])
    for prec in (shaw_amd.ROM_FP64, shaw_amd.ROM_BF16X3):
        rom = shaw_amd.Rom(arv, ars, 1.0, prec, C, max_cols=40, max_samples=n // 5)
        horizons = [60, 110, 160]
        slots = rom.admit(B, ftab, horizons, 5)
        rom.advance(70)
        _, _, st = rom.state(slots[0])
        assert st == 60
        rom.advance(100)
        outs = [rom.retire(s, n // 5) for s in slots]
        assert rom.free_slots() == 48
        for j, h in enumerate(horizons):
            alone = shaw_amd.rom_run(arv, ars, B[:, j:j + 1], ftab[:, j:j + 1], 1.0, h, C, 5, prec)
            np.testing.assert_array_equal(outs[j][1], alone["qv"][:, 0])
            np.testing.assert_array_equal(outs[j][0], alone["seismogram"][0])
        # slots are reused: admit again and get the same answer as the first time
        slots2 = rom.admit(B[:, :1], ftab[:, :1], [60], 5)
        rom.advance(60)
        again = rom.retire(slots2[0], n // 5)
        np.testing.assert_array_equal(again[1], outs[0][1])
        rom.close()


REF_SVD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "computeThinSVD")


def write_snaps_binary(path, A):
    """2D binary snapshot file: [u64 rows][u64 cols] + column-major doubles
    (src/shared/io/matrix_write.hpp:60-76)."""
    with open(path, "wb") as f:
        np.array(A.shape, dtype=np.uint64).tofile(f)
        np.asfortranarray(A).T.ravel().tofile(f)


def read_matrix_binary(path):
    with open(path, "rb") as f:
        m, n = np.fromfile(f, dtype=np.uint64, count=2)
        return np.fromfile(f, dtype=np.float64, count=int(m * n)).reshape(int(n), int(m)).T


def check_pod(phi, sig, U, s, k):
    """Singular values within 1e-12 sigma_1 (the scale of a backward-stable
    SVD's absolute error) and 1e-10 relative where sigma_k > 1e-3 sigma_1;
    orthonormal modes; every mode separated from its neighbours by > 1 %
    (relative gap) spans the same direction as the reference's,
    |cos| > 1 - 1e-6 (SVDs agree on a mode to ~ eps sigma_1 / (gap sigma_k))."""
    err = np.abs(sig[:k] - s[:k])
    np.testing.assert_array_less(err, 1e-12 * s[0])
    big = s[:k] > 1e-3 * s[0]
    np.testing.assert_array_less(err[big], 1e-10 * s[:k][big])
    np.testing.assert_allclose(phi.T @ phi, np.eye(k), atol=1e-10)
    for i in range(k):
        gap = min([abs(s[i] - s[j]) / s[i] for j in (i - 1, i + 1) if 0 <= j < len(s)])
        if gap > 1e-2 and s[i] > 1e-8 * s[0]:
            assert abs(abs(phi[:, i] @ U[:, i]) - 1.0) < 1e-6, (i, s[i] / s[0], gap)


def test_pod_basis_vs_numpy_and_reference_tool(tmp_path):
    """GPU POD (method of snapshots, FP64) against numpy's SVD and against the
    reference's own computeThinSVD (Eigen BDCSVD, built from its sources into
    oracle/_ref) on snapshot files of two FOM runs, read as that tool reads them."""
    mesh = os.path.join(MESHES, "fullMesh21x51")
    ora = oracle.App(mesh, PHYS)
    dirs = []
    for i, (depth, period) in enumerate([(434.0, 40.0), (720.0, 45.0)]):
        out = ora.fom_run(1.0, 200, [(depth, 0.0, period, 10.0)], False, [], 0, 4, 4)
        d = tmp_path / f"run{i}"
        d.mkdir()
        write_snaps_binary(str(d / "snaps_vp_0"), out["snaps_vp"][0].T)
        write_snaps_binary(str(d / "snaps_sp_0"), out["snaps_sp"][0].T)
        dirs.append(d)
    for dof in ("vp", "sp"):
        S = np.hstack([read_matrix_binary(str(d / f"snaps_{dof}_0")) for d in dirs])
        k = 18  # below the numerical rank (~22) of these 100 snapshots
        phi, sig = shaw_amd.pod_basis(S, k)
        U, s, _ = np.linalg.svd(S, full_matrices=False)
        check_pod(phi, sig, U, s, k)
        if os.path.exists(REF_SVD):
            import subprocess
            subprocess.run([REF_SVD, "--dirs"] + [str(d) for d in dirs] +
                           ["--informat", "binary", "--outformat", "binary", "--method", "2"],
                           cwd=str(tmp_path), check=True, capture_output=True)
            Ur = read_matrix_binary(str(tmp_path / f"lsv_{dof}"))
            sr = np.loadtxt(str(tmp_path / f"sva_{dof}"))
            check_pod(phi, sig, Ur, sr, k)
        else:
            say("[gpu] oracle/_ref/computeThinSVD not built: numpy SVD only")
    # asking for more modes than the snapshots resolve is an error, not garbage
    with pytest.raises(shaw_amd.ShawError, match="rank"):
        shaw_amd.pod_basis(S, 60)
