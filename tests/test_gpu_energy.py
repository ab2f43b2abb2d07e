"""GPU tests of the energy-consistent ROM (an addition to the reference's
Euclidean Galerkin ROM): POD in the FOM energy weights against numpy's SVD of
the weighted snapshots, the energy projection against numpy on the oracle's
CSR operators (Ar_s = -Ar_v^T exactly), and the accuracy of the reduced
leapfrog against the FP64 FOM in FP64 and in BF16x3, where the Euclidean
ROM's operator rounding is amplified and the energy ROM's is not."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import shaw_amd
from shaw_amd import rom_workflow as rw, workloads
from tests.helpers import rel_l2, rel_max
from tests.test_gpu_rom import check_pod

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


LAYERED = workloads.layered_material()


@pytest.fixture(scope="module")
def setup128():
    """128 x 512 mesh, one source depth, four training periods, 1000 s; the
    snapshots (every 12 steps) have numerical rank 168 per dof."""
    nr, nth, dt, n, every = 128, 512, 0.4, 2500, 12
    app = shaw_amd.App.structured(nr, nth, LAYERED)
    ora = oracle.App("", LAYERED, structured=(nr, nth, 0.0, 180.0, 3480.0, 6371.0))
    train = [(640.0, 0.0, p, 105.0) for p in (20.0, 35.0, 50.0, 70.0)]
    return app, ora, dt, n, every, train


def test_weighted_pod_vs_numpy(setup128):
    app, ora, dt, n, every, train = setup128
    say("[gpu] weighted POD 128x512")
    mv, ms = rw.energy_weights(app)
    out = ora.fom_run(dt, n, train, True, [], 0, every, every)
    T = n // every
    for dof, w, nn in (("vp", mv, app.nVp), ("sp", ms, app.nSp)):
        S = out["snaps_" + dof].reshape(-1, nn).T
        sw = np.sqrt(w)
        U, s, _ = np.linalg.svd(sw[:, None] * S, full_matrices=False)
        fom = shaw_amd.Fom(app, len(train), dt)
        fom.set_sources([app.map_source(d, a) for d, a, p, dl in train])
        fom.set_forcing(shaw_amd.forcing_table(["ricker"] * 4, [105.0] * 4, [t[2] for t in train], dt, n))
        fom.set_snapshots(dof, every, list(range(4)), T)
        fom.advance(n)
        k = 120
        phi, sig = fom.pod(dof, k, 4 * T, w)
        fom.close()
        check_pod(sw[:, None] * phi, sig, U, s, k)


def _training_fom(app, dt, n, every, train):
    fom = shaw_amd.Fom(app, len(train), dt)
    fom.set_sources([app.map_source(d, a) for d, a, p, dl in train])
    fom.set_forcing(shaw_amd.forcing_table(["ricker"] * 4, [105.0] * 4, [t[2] for t in train], dt, n))
    T = n // every
    fom.set_snapshots("vp", every, list(range(4)), T)
    fom.set_snapshots("sp", every, list(range(4)), T)
    fom.advance(n)
    return fom


@pytest.mark.parametrize("weighted", [False, True])
def test_device_basis_equals_host_pod(setup128, weighted):
    """shaw_fom_pod_basis keeps the bases on the device: its modes, singular
    values, gathered rows and projected operators equal the host-returning
    POD (shaw_fom_pod / shaw_fom_pod_weighted) and projection calls bit for
    bit, Euclidean and energy-weighted."""
    app, ora, dt, n, every, train = setup128
    say(f"[gpu] device basis == host POD, weighted={weighted}")
    mv, ms = rw.energy_weights(app) if weighted else (None, None)
    k = 120
    host = {}
    for dof, w in (("vp", mv), ("sp", ms)):
        fom = _training_fom(app, dt, n, every, train)
        host[dof] = fom.pod(dof, k, 4 * (n // every), w)
        fom.close()
    fom = _training_fom(app, dt, n, every, train)
    dev = fom.pod_basis(k, k, mv, ms)
    fom.close()
    for dof in ("vp", "sp"):
        np.testing.assert_array_equal(dev.download(dof), host[dof][0], err_msg=dof)
        np.testing.assert_array_equal(dev.download(dof, 37), host[dof][0][:, :37], err_msg=dof)
        np.testing.assert_array_equal(dev.sigma(dof), host[dof][1], err_msg=dof)
    gids = [app.map_source(d, 0.0) for d in (10.0, 640.0, 2000.0)] + [app.map_receiver(a) for a in (0.0, 45.0, 180.0)]
    np.testing.assert_array_equal(dev.rows("vp", gids, 50), host["vp"][0][gids, :50])
    for kv, ks in ((k, k), (64, 90)):
        a = dev.project(app, kv, ks, mv)
        pv, ps = host["vp"][0][:, :kv], host["sp"][0][:, :ks]
        b = shaw_amd.rom_project_energy(app, pv, ps, mv) if weighted else shaw_amd.rom_project(app, pv, ps)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    with pytest.raises(shaw_amd.ShawError):
        dev.rows("vp", [app.nVp], 5)
    with pytest.raises(shaw_amd.ShawError):
        dev.project(app, k + 1, k)
    dev.close()


def test_energy_projection_vs_numpy(setup128):
    app, ora, dt, n, every, train = setup128
    mv, ms = rw.energy_weights(app)
    rng = np.random.default_rng(3)
    pv = rng.standard_normal((app.nVp, 40))
    ps = rng.standard_normal((app.nSp, 56))
    arv, ars = shaw_amd.rom_project_energy(app, pv, ps, mv)
    ptr, ind, val = ora.csr("vp")
    Jv = sp.csr_matrix((val, ind.astype(np.int64), ptr.astype(np.int64)), shape=(ora.nVp, ora.nSp))
    ref = pv.T @ (mv[:, None] * (Jv @ ps))
    assert rel_max(arv, ref) <= 1e-12
    np.testing.assert_array_equal(ars, -arv.T)


@pytest.fixture(scope="module")
def energy_rom(setup128):
    app, ora, dt, n, every, train = setup128
    say("[gpu] energy ROM training 128x512")
    basis = rw.train_basis(app, train, dt, n, every, 160, energy=True)
    euclid = rw.train_basis(app, train, dt, n, every, 160)
    test = [(640.0, 0.0, p, 105.0) for p in (22.1, 41.9, 63.1)]
    fom = rw.fom_seismograms(app, test, dt, n, workloads.RECEIVERS, 10)
    return basis, euclid, test, fom


@pytest.mark.parametrize("K", [120, 160])
def test_energy_rom_accuracy_fp64_and_bf16x3(setup128, energy_rom, K):
    app, ora, dt, n, every, train = setup128
    basis, euclid, test, fom = energy_rom
    C = rw.receiver_rows(app, basis, workloads.RECEIVERS, K)
    arv, ars = rw.operators(app, basis, K)
    B, ftab = rw.basis_forcing_inputs(app, basis, K, test, dt, n)
    errs = {}
    for prec, name in ((shaw_amd.ROM_FP64, "fp64"), (shaw_amd.ROM_BF16X3, "bf16x3"), (shaw_amd.ROM_BF16X4, "bf16x4")):
        r = shaw_amd.rom_run(arv, ars, B, ftab, dt, n, C, 10, prec)
        errs[name] = max(rel_l2(r["seismogram"][m], fom[m]) for m in range(len(test)))
    Ce = rw.receiver_rows(app, euclid, workloads.RECEIVERS, K)
    ae, se = rw.operators(app, euclid, K)
    Be, fe = rw.basis_forcing_inputs(app, euclid, K, test, dt, n)
    eu = {}
    for prec, name in ((shaw_amd.ROM_FP64, "fp64"), (shaw_amd.ROM_BF16X3, "bf16x3")):
        r = shaw_amd.rom_run(ae, se, Be, fe, dt, n, Ce, 10, prec)
        eu[name] = max(rel_l2(r["seismogram"][m], fom[m]) for m in range(len(test)))
    say(f"[gpu] K={K} max seismogram rel L2 vs FP64 FOM: energy ROM fp64 {errs['fp64']:.2e} bf16x3 {errs['bf16x3']:.2e}"
        f" bf16x4 {errs['bf16x4']:.2e} | Euclidean ROM fp64 {eu['fp64']:.2e} bf16x3 {eu['bf16x3']:.2e}")
    if K == 160:  # the basis spans the trajectories: FP64 error is the projection's rounding
        assert errs["fp64"] <= 1e-5
    # BF16x3 adds at most the operator rounding's phase error to the basis error
    assert errs["bf16x3"] <= max(1e-3, 2.0 * errs["fp64"])
    assert errs["bf16x4"] <= max(2e-5, 1.1 * errs["fp64"])
