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
