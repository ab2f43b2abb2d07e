"""The offline tools on the GPU: computeThinSVD (POD of snapshot files, 2D and
rank-2 3D) against numpy's SVD and the reference's own tool (built from its
sources into oracle/_ref when present); reconstructFomState and
reconstructSeismogram (rank-1 and rank-2) against the oracle's restatement of
the reference's reconstruction loop, bit for bit; shawExe reading an ASCII
basis file."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import shaw_amd
from tests.helpers import MESHES
from tests.test_gpu_exe import read_ascii, run_exe
from tests.test_gpu_rom import REF_SVD, check_pod, read_matrix_binary

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOLS = os.path.join(ROOT, "shaw_amd")
PHYS = {"kind": "unilayer", "layer": {"density": [2000.0, 0.0], "velocity": [5000.0, 0.0]}}


def say(msg):
    print(msg, flush=True)


def tool(name, args, cwd):
    say(f"[gpu] {name} {' '.join(str(a) for a in args)}")
    r = subprocess.run([os.path.join(TOOLS, name)] + [str(a) for a in args], cwd=str(cwd), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    return r


def write_bin(path, a):
    a = np.asarray(a, dtype=np.float64)
    with open(path, "wb") as f:
        np.asarray(a.shape, dtype=np.uint64).tofile(f)
        a.ravel(order="F").tofile(f)


def read_vec_bin(path):
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:8].view(np.uint64)[0])
    return raw[8:].view(np.float64)[:n]


def orthonormal(rng, n, k):
    q, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return q


@pytest.fixture(scope="module")
def snapshot_dirs(tmp_path_factory):
    """Two rank-1 runs (2D files) and one rank-2 set of two forcings (3D file)
    on the 21 x 51 mesh, written as shawExe writes them."""
    base = tmp_path_factory.mktemp("snaps")
    ora = oracle.App(os.path.join(MESHES, "fullMesh21x51"), PHYS)
    dirs = []
    for i, (depth, period) in enumerate([(434.0, 40.0), (720.0, 45.0)]):
        out = ora.fom_run(1.0, 200, [(depth, 0.0, period, 10.0)], False, [], 0, 4, 4)
        d = base / f"run{i}"
        d.mkdir()
        write_bin(d / "snaps_vp_0", out["snaps_vp"][0].T)
        write_bin(d / "snaps_sp_0", out["snaps_sp"][0].T)
        dirs.append(d)
    out = ora.fom_run(1.0, 200, [(1000.0, 0.0, 50.0, 10.0), (300.0, 0.0, 35.0, 10.0)], True, [], 0, 5, 5)
    d = base / "set0"
    d.mkdir()
    for dof in ("vp", "sp"):  # [M][T][n] -> LayoutLeft [n x T x M]
        write_bin(d / f"snaps_{dof}_0", np.transpose(out[f"snaps_{dof}"], (2, 1, 0)))
    dirs.append(d)
    return dirs


def load_snaps(path):
    raw = np.fromfile(path, dtype=np.uint8)
    size = len(raw)
    h = raw[:24].view(np.uint64)
    if size == 16 + 8 * int(h[0]) * int(h[1]):
        n, c = int(h[0]), int(h[1])
        return raw[16:].view(np.float64).reshape(c, n).T
    n, c = int(h[0]), int(h[1]) * int(h[2])
    return raw[24:].view(np.float64).reshape(c, n).T


def test_compute_thin_svd_tool(tmp_path, snapshot_dirs):
    """The POD tool on 2D and 3D snapshot files: lsv_<dof> holds the modes up
    to the numerical rank, sva_<dof> every singular value in Eigen's aligned
    column; both match numpy's SVD of the side-by-side snapshot matrix to the
    bounds of a backward-stable SVD."""
    r = tool("computeThinSVD", ["--dirs"] + snapshot_dirs + ["--informat", "binary", "--outformat", "binary",
                                                             "--method", "1"], tmp_path)
    for dof in ("vp", "sp"):
        S = np.hstack([load_snaps(d / f"snaps_{dof}_0") for d in snapshot_dirs])
        assert S.shape[1] == 2 * 50 + 2 * 40
        U, s, _ = np.linalg.svd(S, full_matrices=False)
        phi = read_matrix_binary(str(tmp_path / f"lsv_{dof}"))
        k = phi.shape[1]
        assert f"svd_matrix_{dof}_rank = {k}" in r.stdout
        assert phi.shape[0] == S.shape[0] and k >= np.count_nonzero(s > 1e-10 * s[0])
        lines = open(tmp_path / f"sva_{dof}").read().split("\n")
        assert lines[-1] == "" and len(lines) == min(S.shape) + 1
        assert len(set(len(l) for l in lines[:-1])) == 1  # right-aligned to one width
        sig = np.array([float(l) for l in lines[:-1]])
        check_pod(phi, sig, U, s, k)


def test_compute_thin_svd_tool_vs_reference_tool(tmp_path, snapshot_dirs):
    if not os.path.exists(REF_SVD):
        pytest.skip("oracle/_ref/computeThinSVD not built")
    two = snapshot_dirs[:2]  # the reference tool reads 2D files only
    (tmp_path / "ours").mkdir()
    (tmp_path / "ref").mkdir()
    tool("computeThinSVD", ["--dirs"] + two + ["--informat", "binary", "--outformat", "binary", "--method", "2"],
         tmp_path / "ours")
    subprocess.run([REF_SVD, "--dirs"] + [str(d) for d in two] + ["--informat", "binary", "--outformat", "binary",
                                                                  "--method", "2"],
                   cwd=str(tmp_path / "ref"), check=True, capture_output=True, timeout=300)
    for dof in ("vp", "sp"):
        phi = read_matrix_binary(str(tmp_path / "ours" / f"lsv_{dof}"))
        Ur = read_matrix_binary(str(tmp_path / "ref" / f"lsv_{dof}"))
        sr = np.loadtxt(str(tmp_path / "ref" / f"sva_{dof}"))
        sig = np.loadtxt(str(tmp_path / "ours" / f"sva_{dof}"))
        assert sig.shape == sr.shape
        check_pod(phi, sig, Ur, sr, phi.shape[1])


@pytest.mark.parametrize("modes_fmt", ["binary", "ascii"])
def test_reconstruct_fom_state_tool(tmp_path, modes_fmt):
    rng = np.random.default_rng(11)
    n, kfile, K, T = 1071, 40, 30, 12
    phi = orthonormal(rng, n, kfile)
    if modes_fmt == "binary":
        write_bin(tmp_path / "lsv_vp", phi)
    else:
        with open(tmp_path / "lsv_vp", "w") as f:
            f.write(f"{n} {kfile}\n")
            for row in phi:
                f.write("".join(f"{float(x)!r} " for x in row) + "\n")
    q = rng.standard_normal((K, T))
    write_bin(tmp_path / "rom_snaps", q)
    tool("reconstructFomState", ["--romsize", K, "--podmodes", "lsv_vp", modes_fmt, "--romsnaps=rom_snaps", "binary",
                                 "--samplingfreq", 5, "--fsize", 1, "--timesteps", 10, 35, 60, "--outformat", "binary",
                                 "--outfileappend", "x"], tmp_path)
    cols = [1, 6, 11]
    ref = oracle.reconstruct(phi[:, :K], q[:, cols])
    for j, step in enumerate((10, 35, 60)):
        got = read_vec_bin(tmp_path / f"fomReconstructedState_timestep_{step}_x")
        np.testing.assert_array_equal(got, ref[:, j])
        np.testing.assert_allclose(got, phi[:, :K] @ q[:, cols[j]], rtol=0, atol=1e-13)
    rs = open(tmp_path / "ROMSTATE.txt").read().split("\n")
    assert int(rs[0]) == K
    np.testing.assert_array_equal([float(x) for x in rs[1:1 + K]], q[:, 11])
    # rank-2: [K x T x F] ROM snapshots, one file per forcing and step (ascii out)
    q3 = rng.standard_normal((K, T, 3))
    write_bin(tmp_path / "rom_snaps3", q3)
    tool("reconstructFomState", ["--romsize", K, "--podmodes", "lsv_vp", modes_fmt, "--romsnaps", "rom_snaps3",
                                 "binary", "--samplingfreq", 1, "--fsize", 3, "--timesteps", 1, 12, "--outformat",
                                 "ascii"], tmp_path)
    for f in range(3):
        ref = oracle.reconstruct(phi[:, :K], q3[:, [0, 11], f])
        for j, step in enumerate((1, 12)):
            lines = open(tmp_path / f"fomReconstructedState_timestep_{step}_f_{f}").read().split("\n")
            assert int(lines[0]) == n
            np.testing.assert_array_equal([float(x) for x in lines[1:1 + n]], ref[:, j])


def test_reconstruct_seismogram_tool(tmp_path):
    rng = np.random.default_rng(12)
    n, kfile, K, T = 1071, 40, 30, 25
    phi = orthonormal(rng, n, kfile)
    write_bin(tmp_path / "lsv_vp", phi)
    gids = [1070, 0, 5, 17, 5]
    q = rng.standard_normal((K + 5, T))  # more coordinates than --romsize: the first K are summed
    write_bin(tmp_path / "rom_snaps", q)
    tool("reconstructSeismogram", ["--romsize", K, "--podmodes", "lsv_vp", "binary", "--romsnaps", "rom_snaps",
                                   "binary", "--gridids"] + gids + ["--fsize", 1, "--outformat", "ascii"], tmp_path)
    got = read_ascii(str(tmp_path / "rom_seismo.txt"), False)
    ref = oracle.reconstruct(phi[gids, :K], q[:K])
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_allclose(got, phi[gids, :K] @ q[:K], rtol=0, atol=1e-13)
    q3 = rng.standard_normal((K, T, 2))
    write_bin(tmp_path / "rom_snaps3", q3)
    tool("reconstructSeismogram", ["--romsize", K, "--podmodes", "lsv_vp", "binary", "--romsnaps", "rom_snaps3",
                                   "binary", "--gridids"] + gids + ["--fsize", 2, "--outformat", "binary",
                                                                    "--outfileappend", "d240"], tmp_path)
    for f in range(2):
        got = read_matrix_binary(str(tmp_path / f"seismo_f_{f}_d240"))
        np.testing.assert_array_equal(got, oracle.reconstruct(phi[gids, :K], q3[:, :, f]))


@pytest.mark.parametrize("n,k,ncols", [(1, 1, 1), (1000, 7, 3), (257, 64, 16), (3001, 130, 17), (600, 300, 41)])
def test_reconstruct_api_bit_exact(n, k, ncols):
    """shaw_reconstruct == the oracle's mode-order sum, across LDS tiles of
    64 modes, 16-column blocks and partial row blocks."""
    rng = np.random.default_rng(n + k + ncols)
    phi = rng.standard_normal((n, k))
    q = rng.standard_normal((k, ncols))
    np.testing.assert_array_equal(shaw_amd.reconstruct(phi, q), oracle.reconstruct(phi, q))


def test_pod_basis_upto_stops_at_rank():
    rng = np.random.default_rng(4)
    S = rng.standard_normal((500, 7)) @ rng.standard_normal((7, 40))
    phi, sig = shaw_amd.pod_basis_upto(S, 40)
    assert phi.shape == (500, 7)
    U, s, _ = np.linalg.svd(S, full_matrices=False)
    check_pod(phi, sig, U, s, 7)


def test_exe_rom_ascii_basis(tmp_path):
    """shawExe's ROM with `binary: false` basis files (ASCII with extents):
    identity bases on the 4 x 5 mesh reproduce the FOM seismogram."""
    mesh = os.path.join(MESHES, "fullMesh4x5")
    ora = oracle.App(mesh, PHYS)
    ref = ora.fom_run(1.0, 60, [(900.0, 45.0, 40.0, 10.0)], False, [45.0, 90.0, 135.0], 2)
    for dof, n in (("vp", ora.nVp), ("sp", ora.nSp)):
        with open(tmp_path / f"lsv_{dof}", "w") as f:
            f.write(f"{n} {n}\n")
            for row in np.eye(n):
                f.write("".join(f"{x!r} " for x in row.tolist()) + "\n")
    cfg = {"general": {"meshDir": mesh, "dt": 1.0, "finalTime": 60.0, "checkNumericalDispersion": False,
                       "checkCfl": False},
           "io": {"seismogram": {"binary": False, "freq": 2, "receivers": [45.0, 90.0, 135.0]}},
           "source": {"signal": {"kind": "ricker", "depth": 900.0, "angle": 45.0, "period": 40.0, "delay": 10.0}},
           "material": PHYS,
           "rom": {"velocity": {"numModes": ora.nVp, "modesFile": str(tmp_path / "lsv_vp"), "binary": False},
                   "stress": {"numModes": ora.nSp, "modesFile": str(tmp_path / "lsv_sp"), "binary": False}}}
    res = run_exe(cfg, tmp_path)
    assert res.returncode == 0, res.stderr + res.stdout[-3000:]
    seis = read_ascii(str(tmp_path / "seismogram_0"), False)
    np.testing.assert_allclose(seis, ref["seismogram"][0].T, rtol=0, atol=1e-12 * np.abs(ref["seismogram"]).max())
