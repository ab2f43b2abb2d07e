"""GPU tests of the device-side setup: operators assembled on the GPU for
every material model equal the host assembly (itself bit-exact to the
oracle) bit for bit, a FOM whose material is replaced on the device equals a
FOM created with that material, forcing tables evaluated on the GPU stay
within a few ulp of the host tables (and FOM/ROM runs driven by them within
1e-12), and sources and receivers mapped on the GPU land on the host's grid
points, including exact ties."""
import os
import sys

import numpy as np
import pytest

import shaw_amd
from shaw_amd import workloads
from tests.helpers import MESHES

pytestmark = pytest.mark.gpu

UNILAYER = {"kind": "unilayer", "layer": {"density": [2600.0, 1.2e-4, 3e-11], "velocity": [3200.0, 2e-4, 1e-11]}}
BILAYER = {"kind": "bilayer", "layer1": {"density": [2500.0, 0.001], "velocity": [4500.0, 0.0001]},
           "layer2": {"depth": 700.0, "density": [4000.0, 0.0002], "velocity": [6000.0, 0.0003]}}
MATERIALS = {"unilayer": UNILAYER, "bilayer": BILAYER, "prem": {"kind": "prem"}, "layered": workloads.layered_material()}
EPS = np.finfo(np.float64).eps


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mesh", ["structured 96x300", "fullMesh21x51"])
@pytest.mark.parametrize("name", list(MATERIALS))
def test_device_operators_equal_host(mesh, name):
    say(f"[gpu] device assembly {mesh} {name}")
    mat = MATERIALS[name]
    if mesh.startswith("structured"):
        app = shaw_amd.App.structured(96, 300, mat)
    else:
        app = shaw_amd.App.from_mesh(os.path.join(MESHES, mesh), mat)
    cv, cs, ri, _ = app.operators()
    dcv, dcs, dri = app.operators_device(mat)
    np.testing.assert_array_equal(dcv, cv)
    np.testing.assert_array_equal(dcs, cs)
    np.testing.assert_array_equal(dri, ri)
    # another material on the same app's mesh
    other = MATERIALS["layered" if name != "layered" else "bilayer"]
    ref = (shaw_amd.App.structured(96, 300, other) if mesh.startswith("structured")
           else shaw_amd.App.from_mesh(os.path.join(MESHES, mesh), other)).operators()
    got = app.operators_device(other)
    for a, b in zip(got, ref[:3]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", [shaw_amd.FP64, shaw_amd.FP32])
def test_fom_set_material_equals_new_fom(precision):
    """A FOM created with one material, its operators replaced on the device
    (and its sources' 1/rho updated), runs bit for bit as a FOM created on an
    app with the new material."""
    say(f"[gpu] set_material precision={precision}")
    nr, nth, dt, n = 64, 256, 0.5, 90
    app_a = shaw_amd.App.structured(nr, nth, MATERIALS["prem"])
    app_b = shaw_amd.App.structured(nr, nth, MATERIALS["layered"])
    src = [(100.0 + 80.0 * k, 3.0 * k, 30.0 + 5.0 * k, 60.0) for k in range(6)]
    rec = [app_a.map_receiver(a) for a in np.arange(0.0, 180.0, 6.0)]

    def run(app, swap):
        f = shaw_amd.Fom(app, len(src), dt, precision=precision)
        f.set_sources([app.map_source(d, a) for d, a, p, dl in src])
        f.set_forcing(shaw_amd.forcing_table(["ricker"] * len(src), [s[3] for s in src], [s[2] for s in src], dt, n))
        f.set_receivers(rec, 3, n // 3)
        if swap:
            f.set_material(app, MATERIALS["layered"])
        f.advance(n)
        out = (f.seismogram(), *f.state())
        f.close()
        return out

    a = run(app_a, True)
    b = run(app_b, False)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert np.abs(a[0]).max() > 0


def test_forcing_table_device_within_ulps():
    say("[gpu] device forcing table")
    rng = np.random.default_rng(4)
    m, n, dt = 24, 4000, 0.05
    kinds = [["ricker", "sinusoid", "gaussDer"][k % 3] for k in range(m)]
    delays = rng.uniform(20.0, 120.0, m)
    periods = rng.uniform(15.0, 80.0, m)
    host = shaw_amd.forcing_table(kinds, delays, periods, dt, n)
    dev = shaw_amd.forcing_table_device(kinds, delays, periods, dt, n)
    scale = np.abs(host).max(axis=0)
    err = np.abs(dev - host).max(axis=0)
    assert np.all(err <= 8 * EPS * scale), (err / scale).max()
    assert np.mean(dev == host) > 0.5  # most values round identically
    with pytest.raises(shaw_amd.ShawError):
        shaw_amd.forcing_table_device(["ricker"], [1.0], [-1.0], dt, 4)


def test_fom_set_signals_matches_host_table():
    say("[gpu] FOM with device forcing")
    nr, nth, dt, n = 64, 256, 0.5, 120
    app = shaw_amd.App.structured(nr, nth, MATERIALS["layered"])
    src = [(150.0 + 90.0 * k, 2.0 * k, 25.0 + 6.0 * k, 70.0) for k in range(8)]
    kinds = ["ricker", "gaussDer", "sinusoid", "ricker"] * 2
    rec = [app.map_receiver(a) for a in np.arange(0.0, 180.0, 5.0)]
    out = []
    for device in (False, True):
        f = shaw_amd.Fom(app, len(src), dt)
        f.set_sources([app.map_source(d, a) for d, a, p, dl in src])
        if device:
            f.set_signals(kinds, [s[3] for s in src], [s[2] for s in src], n)
        else:
            f.set_forcing(shaw_amd.forcing_table(kinds, [s[3] for s in src], [s[2] for s in src], dt, n))
        f.set_receivers(rec, 2, n // 2)
        f.advance(n)
        out.append(f.seismogram())
        f.close()
    a, b = out
    assert np.abs(b).max() > 0
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_rom_admit_signals_matches_host_table():
    say("[gpu] ROM admission with device forcing")
    rng = np.random.default_rng(8)
    K, M, n, dt = 96, 40, 600, 0.05
    A = rng.standard_normal((K, K)) * (0.5 / np.sqrt(K))
    C = rng.standard_normal((20, K))
    B = rng.standard_normal((K, M))
    kinds = [["ricker", "gaussDer", "sinusoid"][j % 3] for j in range(M)]
    delays, periods = rng.uniform(5.0, 20.0, M), rng.uniform(4.0, 12.0, M)
    hz = [int(h) for h in rng.integers(n // 2, n + 1, M)]
    res = []
    for device in (False, True):
        rom = shaw_amd.Rom(A, -A.T.copy(), dt, shaw_amd.ROM_FP64, C, max_cols=M, max_samples=n // 3)
        if device:
            slots = rom.admit_signals(B, kinds, delays, periods, n, hz, 3)
        else:
            slots = rom.admit(B, shaw_amd.forcing_table(kinds, delays, periods, dt, n), hz, 3)
        rom.advance(n)
        res.append([rom.retire(s, n // 3)[0] for s in slots])
        rom.close()
    for a, b in zip(*res):
        assert np.abs(b).max() > 0
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("mesh", ["structured 200x1000", "fullMesh21x51"])
def test_map_points_device_equals_host(mesh):
    """Random locations, locations exactly between grid columns and rows
    (ties, resolved to the last gid), the bounds and the domain's corners."""
    say(f"[gpu] device mapping {mesh}")
    app = (shaw_amd.App.structured(200, 1000, MATERIALS["prem"]) if mesh.startswith("structured")
           else shaw_amd.App.from_mesh(os.path.join(MESHES, mesh), MATERIALS["prem"]))
    rng = np.random.default_rng(12)
    th_deg = np.degrees(app.coords("vp")[: app.nth, 0])
    depth_top = (app.bounds[3] - app.bounds[2]) / 1000.0 - 1e-9  # the CMB, inside after rounding
    r_km = app.coords("vp")[:: app.nth, 1] / 1000.0  # (theta, r [m]) per point
    angles = list(rng.uniform(0.0, 180.0, 400)) + list(0.5 * (th_deg[:-1] + th_deg[1:])) + [0.0, 180.0, 90.0]
    depths = list(rng.uniform(0.0, depth_top, 400)) + [float(x) for x in rng.uniform(0.0, depth_top, app.nth - 1)] \
        + [0.0, depth_top, depth_top / 2]
    mid_r = 0.5 * (r_km[:-1] + r_km[1:])
    depths += [app.bounds[3] / 1000.0 - x for x in mid_r]
    angles += [float(a) for a in rng.uniform(0.0, 180.0, len(mid_r))]
    got = app.map_points_device(angles, depths)
    want = [app.map_source(d, a) for d, a in zip(depths, angles)]
    np.testing.assert_array_equal(got, want)
    rec = list(rng.uniform(0.0, 180.0, 300)) + list(0.5 * (th_deg[:-1] + th_deg[1:])) + [0.0, 180.0]
    np.testing.assert_array_equal(app.map_points_device(rec), [app.map_receiver(a) for a in rec])
    with pytest.raises(shaw_amd.ShawError):
        app.map_points_device([190.0])
