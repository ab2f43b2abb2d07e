"""CPU tests of the offline tools (shaw_amd/computeThinSVD,
extractStateFromSnaps, reconstructFomState, reconstructSeismogram): command
lines in the reference's CLI11 forms, the state extractor's outputs against
direct indexing of the snapshot matrix (bit for bit, binary and 17-digit
ASCII), and the error paths that stop before any GPU work. The GPU products
(POD, reconstruction) are tested in test_gpu_tools.py."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOLS = os.path.join(ROOT, "shaw_amd")


def run(tool, args, cwd):
    return subprocess.run([os.path.join(TOOLS, tool)] + [str(a) for a in args], cwd=cwd, capture_output=True,
                          text=True, timeout=120)


def write_bin(path, a):
    """matrix_write.hpp: [u64 extents...] + column-major data."""
    a = np.asarray(a, dtype=np.float64)
    with open(path, "wb") as f:
        np.asarray(a.shape, dtype=np.uint64).tofile(f)
        a.ravel(order="F").tofile(f)


def write_ascii_matrix(path, a):
    with open(path, "w") as f:
        f.write(f"{a.shape[0]} {a.shape[1]}\n")
        for row in a:
            f.write("".join(f"{float(x)!r} " for x in row) + "\n")


def read_vec_bin(path):
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:8].view(np.uint64)[0])
    v = raw[8:].view(np.float64)
    assert len(v) == n
    return v


def read_vec_ascii(path):
    lines = open(path).read().split("\n")
    n = int(lines[0])
    body = lines[1:1 + n]
    assert all(l.endswith(" ") for l in body)  # vector_write.hpp writes value, " \n"
    return np.array([float(l) for l in body])


@pytest.fixture(scope="module")
def snaps():
    rng = np.random.default_rng(3)
    return rng.standard_normal((37, 12)), rng.standard_normal((29, 8, 3))


def test_extract_rank1_binary_and_ascii(tmp_path, snaps):
    S2, _ = snaps
    write_bin(tmp_path / "snaps_vp_0", S2)
    r = run("extractStateFromSnaps", ["--snaps=./snaps_vp_0", "binary", "--fsize=1", "--outformat=binary",
                                      "--timesteps=100", "250", "1200", "--samplingfreq=100", "--outfileappend=vp_d240"],
            tmp_path)
    assert r.returncode == 0, r.stderr
    assert "success" in r.stdout and "Target time steps = 100 250 1200" in r.stdout
    for step in (100, 250, 1200):
        v = read_vec_bin(tmp_path / f"state_timestep_{step}_vp_d240")
        np.testing.assert_array_equal(v, S2[:, step // 100 - 1])
    # ASCII snapshot in, ASCII state out (17 significant digits round-trip exactly)
    write_ascii_matrix(tmp_path / "snaps_a", S2)
    r = run("extractStateFromSnaps", ["--snaps", "snaps_a", "ascii", "--fsize", "1", "--outformat", "ascii",
                                      "--timesteps", "30", "--samplingfreq", "10"], tmp_path)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(read_vec_ascii(tmp_path / "state_timestep_30"), S2[:, 2])


def test_extract_rank2_three_dimensional(tmp_path, snaps):
    _, S3 = snaps
    write_bin(tmp_path / "snaps_sp_0", S3)
    r = run("extractStateFromSnaps", ["--snaps=snaps_sp_0", "binary", "--fsize=3", "--outformat=binary",
                                      "--timesteps=5", "40", "--samplingfreq=5"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "snap size: 29 8 3" in r.stdout
    for f in range(3):
        for step in (5, 40):
            v = read_vec_bin(tmp_path / f"state_timestep_{step}_f_{f}")
            np.testing.assert_array_equal(v, S3[:, step // 5 - 1, f])


@pytest.mark.parametrize("args,msg", [
    (["--snaps=snaps_vp_0", "binary", "--fsize=1", "--outformat=binary", "--timesteps=1300", "--samplingfreq=100"],
     "not in the snapshot matrix"),
    (["--snaps=snaps_vp_0", "binary", "--fsize=1", "--outformat=binary", "--timesteps=50", "--samplingfreq=100"],
     "not in the snapshot matrix"),
    (["--snaps=snaps_vp_0", "ascii", "--fsize=2", "--outformat=binary", "--timesteps=100", "--samplingfreq=100"],
     "Rank-2  snaps ascii not supported yet"),
    (["--snaps=snaps_vp_0", "binary", "--fsize=2", "--outformat=binary", "--timesteps=100", "--samplingfreq=100"],
     "three-dimensional"),
    (["--snaps=snaps_vp_0", "binary", "--fsize=1", "--outformat=binary", "--samplingfreq=100"],
     "--timesteps is required"),
    (["--snaps=snaps_vp_0", "--fsize=1", "--outformat=binary", "--timesteps=100", "--samplingfreq=100"],
     "--snaps is required"),
])
def test_extract_errors(tmp_path, snaps, args, msg):
    write_bin(tmp_path / "snaps_vp_0", snaps[0])
    r = run("extractStateFromSnaps", args, tmp_path)
    assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)


def test_svd_and_reconstruct_argument_errors(tmp_path, snaps):
    """Errors detected before any GPU call: the reference's "Invalid method",
    unreadable or truncated files, a ROM size beyond the modes on file."""
    r = run("computeThinSVD", ["--dirs", ".", "--informat", "binary", "--outformat", "binary"], tmp_path)
    assert r.returncode == 1 and "Invalid method" in r.stderr
    r = run("computeThinSVD", ["--dirs", "nowhere", "--informat", "binary", "--outformat", "binary", "--method", "1"],
            tmp_path)
    assert r.returncode == 1 and "cannot open nowhere/snaps_vp_0" in r.stderr
    with open(tmp_path / "snaps_vp_0", "wb") as f:  # header promises more data than the file holds
        np.array([10, 10], dtype=np.uint64).tofile(f)
        np.zeros(20).tofile(f)
    r = run("computeThinSVD", ["--dirs", ".", "--informat", "binary", "--outformat", "binary", "--method", "2"],
            tmp_path)
    assert r.returncode == 1 and "does not match a 2D or 3D binary header" in r.stderr
    write_bin(tmp_path / "lsv_vp", snaps[0])  # 37 x 12 modes
    write_bin(tmp_path / "rom_snaps", np.zeros((13, 4)))
    common = ["--podmodes", "lsv_vp", "binary", "--romsnaps", "rom_snaps", "binary", "--fsize", "1",
              "--outformat", "binary"]
    r = run("reconstructFomState", ["--romsize", "13", "--samplingfreq", "1", "--timesteps", "1"] + common, tmp_path)
    assert r.returncode == 1 and "exceeds the 12 modes" in r.stderr
    r = run("reconstructSeismogram", ["--romsize", "13", "--gridids", "0", "5"] + common, tmp_path)
    assert r.returncode == 1 and "exceeds the 12 modes" in r.stderr
    r = run("reconstructSeismogram", ["--romsize", "12", "--gridids", "0", "37"] + common, tmp_path)
    assert r.returncode == 1 and "grid id 37 out of range" in r.stderr
    r = run("reconstructFomState", ["--romsize", "12", "--samplingfreq", "1", "--timesteps", "1"] + common, tmp_path)
    assert r.returncode == 1 and "does not match --romsize" in r.stderr


def test_tool_name_dispatch(tmp_path):
    import shutil
    shutil.copy(os.path.join(TOOLS, "computeThinSVD"), tmp_path / "someTool")
    r = subprocess.run([str(tmp_path / "someTool")], capture_output=True, text=True, timeout=60,
                       env={**os.environ, "LD_LIBRARY_PATH": TOOLS})
    assert r.returncode == 2 and "invoke as computeThinSVD" in r.stderr
