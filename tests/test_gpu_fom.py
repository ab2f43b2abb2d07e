"""GPU parity of the FOM path (HIP kernels through the C ABI) against the oracle.

Bar: FP64 results bit-identical to the oracle (same CSR row order, no FMA on
either side) for every reference regression input and for random rank-2
batches, with both kernel families (split half-steps, fused step) and a
spread of tile shapes; within 1e-12 relative of the reference golds; FP32
state variant within a stated tolerance; size-independent properties at the
full config-2 shape.
"""
import os
import sys

import numpy as np
import pytest

import oracle
import shaw_amd
from tests.helpers import MESHES, fom_golds, load_input, realizations, rel_l2, rel_max
from tests.test_oracle import FOM_TESTS, gold_view

pytestmark = pytest.mark.gpu

LAYERED = {"kind": "layered", "radius": [3480.0, 4100.0, 4700.0, 5200.0, 5600.0, 5900.0, 6150.0, 6371.0],
           "density": [5600.0, 5300.0, 4950.0, 4600.0, 4200.0, 3900.0, 3600.0, 3300.0],
           "velocity": [7300.0, 7000.0, 6600.0, 6200.0, 5700.0, 5200.0, 4800.0, 4400.0]}

KERNELS = [dict(kernel=shaw_amd.KERNEL_SPLIT), dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=1),
           dict(kernel=shaw_amd.KERNEL_FUSED, strip=8, segment=3, steps_per_launch=1),
           dict(kernel=shaw_amd.KERNEL_FUSED, cols_per_thread=1, strip=5, segment=1, steps_per_launch=1),
           dict(kernel=shaw_amd.KERNEL_FUSED, prefetch_rows=2, strip=6, segment=4),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=2),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=2, strip=5, segment=3),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=3),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=3, strip=7, segment=2),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=4),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=4, cols_per_thread=1, strip=8, segment=5),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=4, col_block=2, strip=64),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=3, col_block=1, strip=64, segment=3),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=4, col_block=2, strip=64, prefetch_rows=2,
                segment=7),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=5, col_block=2, strip=64, segment=6),
           dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=6, col_block=1, strip=64)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(test, **kw):
    cfg = load_input(test)
    g, io, sig = cfg["general"], cfg.get("io", {}), cfg["source"]["signal"]
    mesh = os.path.join(MESHES, g["meshDir"])
    app = shaw_amd.App.from_mesh(mesh, cfg["material"])
    ora = oracle.App(mesh, cfg["material"])
    nsteps = int(g["finalTime"] / g["dt"])
    seis, snap = io.get("seismogram", {}), io.get("snapshotMatrix", {})
    reals = realizations(sig)
    fs = sig.get("forcingSize")
    sets = [reals[i:i + fs] for i in range(0, len(reals), fs)] if fs else [[r] for r in reals]
    args = dict(receivers=seis.get("receivers", []), seis_freq=seis.get("freq", 0),
                snap_freq_vp=snap.get("velocity", {}).get("freq", 0),
                snap_freq_sp=snap.get("stress", {}).get("freq", 0), signal=sig["kind"])
    for rid, s in enumerate(sets):
        a = shaw_amd.fom_run(app, g["dt"], nsteps, s, fs is not None, **args, **kw)
        b = ora.fom_run(g["dt"], nsteps, s, fs is not None, **args)
        yield rid, len(s), fs is not None, a, b


@pytest.mark.parametrize("kw", KERNELS, ids=["split", "fused", "fused-narrow", "fused-cpt1", "fused-pf2", "fused2",
                                             "fused2-narrow", "fused3", "fused3-narrow", "fused4", "fused4-min",
                                             "wave4", "wave3-1col", "wave4-pf2", "wave5", "wave6-1col"])
@pytest.mark.parametrize("test", FOM_TESTS)
def test_regression_inputs_bit_exact(test, kw):
    say(f"[gpu] {test} {kw}")
    golds = fom_golds()
    for rid, M, rank2, a, b in run_both(test, **kw):
        for key in b:
            np.testing.assert_array_equal(a[key], b[key], err_msg=f"{test} run {rid} {key}")
        for key in ("snaps_vp", "snaps_sp", "seismogram"):
            name = f"{test}/{key}_{rid}"
            if name in golds.files:
                ref = gold_view(key, golds[name], rank2, M)
                for m in range(M):
                    assert rel_l2(a[key][m], ref[m]) <= 1e-12 and rel_max(a[key][m], ref[m]) <= 1e-12, name


def random_sources(rng, M, depth_max=2800.0):
    return [(float(rng.uniform(10, depth_max)), float(rng.uniform(0, 180)), float(rng.uniform(20, 70)),
             float(rng.uniform(5, 60))) for _ in range(M)]


@pytest.fixture(scope="module")
def mesh64x256(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("mesh") / "mesh64x256")
    shaw_amd.mesh_write(d, 64, 256)
    return d


@pytest.mark.parametrize("M,kw", [(32, dict(kernel=shaw_amd.KERNEL_SPLIT)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=1)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=8, cols_per_thread=8, strip=32, segment=7)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, cols_per_thread=2, strip=16, segment=64)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, prefetch_rows=2, strip=32, segment=32)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, prefetch_rows=2, cols_per_thread=1,
                                            strip=16, segment=10)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=2)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=8, steps_per_launch=2, cols_per_thread=4,
                                            strip=12, segment=7)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=2, cols_per_thread=1,
                                            strip=16, segment=64)),
                                  (7, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=2, strip=4, segment=2)),
                                  (7, dict(kernel=shaw_amd.KERNEL_FUSED)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=3)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=4)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=8, steps_per_launch=4, cols_per_thread=4,
                                            strip=16, segment=7)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=3, cols_per_thread=1,
                                            strip=16, segment=64)),
                                  (7, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=4, strip=9, segment=1)),
                                  (3, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=3, strip=64, segment=2)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=32, steps_per_launch=4)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=8, steps_per_launch=4, strip=128)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=4, steps_per_launch=2)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, col_block=8)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=5)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=6, col_block=1, segment=9)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=6, segment=30)),
                                  (32, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=7)),
                                  (40, dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=7, col_block=1, segment=9)),
                                  (40, dict(kernel=shaw_amd.KERNEL_SPLIT, cols_per_thread=2))])
def test_random_rank2_bit_exact(mesh64x256, M, kw):
    say(f"[gpu] random rank-2 M={M} {kw}")
    rng = np.random.default_rng(M)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    ora = oracle.App(mesh64x256, LAYERED)
    src = random_sources(rng, M)
    rec = list(np.arange(0, 180, 7.5))
    dt, n = 0.5, 120
    a = shaw_amd.fom_run(app, dt, n, src, True, rec, 3, 40, 30, **kw)
    b = ora.fom_run(dt, n, src, True, rec, 3, 40, 30)
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert np.abs(a["seismogram"]).max() > 0


def random_case(seed):
    """A valid random FOM configuration (mesh, columns, forcing form, kernel
    shape, observers), drawn from a seed."""
    rng = np.random.default_rng(seed)
    nr, nth = int(rng.integers(5, 41)), int(rng.integers(7, 97))
    M = 1 if rng.random() < 0.25 else int(rng.integers(2, 71))
    spl = int(rng.integers(1, 5))
    col_block = int(rng.choice([0, 4, 8, 16, 32]))
    mc = 1
    while mc < M and mc < (col_block or 2):
        mc *= 2
    cpt = int(rng.choice([c for c in (1, 2, 4, 8) if mc % c == 0]))
    tpp = mc // cpt
    pmax = min(512 // tpp, 128)
    pmin = max(2 * spl, 2)
    strip = int(rng.integers(pmin, pmax + 1)) if pmax >= pmin else 0

    def lds(P):  # dynamic LDS of a fused launch (FP64), at most 160 KiB
        if spl == 1:
            return (2 * P + 1) * mc * 8
        if spl == 2:
            return (4 * P + 2) * mc * 8
        return (spl * P * 8 + 2 * (P + 1) * mc + spl * P * mc) * 8

    while strip > pmin and lds(strip) > 160 * 1024:
        strip -= 1
    kw = dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=spl, col_block=col_block, cols_per_thread=cpt,
              segment=int(rng.integers(1, 20)), graph_steps=int(rng.choice([-1, 4 * spl])))
    if strip:
        kw["strip"] = strip
    if spl == 1 and rng.random() < 0.5:
        kw["prefetch_rows"] = 2
    if rng.random() < 0.15:  # the split kernels take only the column shape
        kw = dict(kernel=shaw_amd.KERNEL_SPLIT, col_block=col_block, cols_per_thread=cpt)
    src = [(float(rng.uniform(10, 2800)), float(rng.uniform(0, 180)), float(rng.uniform(20, 70)),
            float(rng.uniform(0, 30))) for _ in range(M)]
    rec = [float(x) for x in rng.uniform(0, 180, int(rng.integers(1, 25)))]
    rank2 = M > 1 or bool(rng.integers(0, 2))
    if not rank2:
        kw["rank1_forcing"] = True
    if kw["kernel"] == shaw_amd.KERNEL_FUSED and rng.random() < 0.3:  # one wavefront per strip
        cbw = int(rng.choice([1, 2]))
        mcw = 1
        while mcw < M and mcw < cbw:
            mcw *= 2
        splw = int(rng.integers(3, 7))
        kw.update(col_block=cbw, cols_per_thread=mcw, strip=64, steps_per_launch=splw)
        kw.pop("prefetch_rows", None)
        if kw["graph_steps"] > 0:
            kw["graph_steps"] = 4 * splw
        pfw = int(rng.integers(1, 3))
        kw["prefetch_rows"] = pfw if splw <= 4 else 1  # 5 and 6 steps load one row ahead
    every = [int(rng.integers(1, 6)) for _ in range(3)]
    if rng.random() < 0.4:  # no snapshots: whole graph chunks replay
        every[1] = every[2] = 0
    return (nr, nth), src, rank2, rec, every, kw


@pytest.mark.parametrize("seed", range(48))
def test_random_configurations_bit_exact(seed):
    """Seeded random meshes, column counts, forcing forms, steps per launch,
    column blocks, tile shapes, graph chunks and observer periods: the GPU run
    equals the oracle bit for bit."""
    (nr, nth), src, rank2, rec, every, kw = random_case(seed)
    say(f"[gpu] random case {seed}: {nr}x{nth} M={len(src)} rank2={rank2} every={every} {kw}")
    app = shaw_amd.App.structured(nr, nth, LAYERED)
    ora = oracle.App("", LAYERED, structured=(nr, nth, 0.0, 180.0, 3480.0, 6371.0))
    n, dt = 60, 0.5
    a = shaw_amd.fom_run(app, dt, n, src, rank2, rec, every[0], every[1], every[2], **kw)
    b = ora.fom_run(dt, n, src, rank2, rec, every[0], every[1], every[2])
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=f"seed {seed} {key}")


@pytest.mark.parametrize("seed", range(100, 124))
def test_random_configurations_fp32_equal_single_steps(seed):
    """The FP32 state variant on seeded random configurations: every kernel
    and shape equals FP32 single steps of the fused kernel bit for bit (the
    same single-precision expressions in the same order)."""
    (nr, nth), src, rank2, rec, every, kw = random_case(seed)
    say(f"[gpu] fp32 random case {seed}: {nr}x{nth} M={len(src)} rank2={rank2} every={every} {kw}")
    app = shaw_amd.App.structured(nr, nth, LAYERED)
    n, dt = 60, 0.5
    one = dict(kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=1, col_block=kw.get("col_block", 0))
    if "rank1_forcing" in kw:
        one["rank1_forcing"] = kw["rank1_forcing"]
    a = shaw_amd.fom_run(app, dt, n, src, rank2, rec, every[0], every[1], every[2], precision=shaw_amd.FP32, **kw)
    b = shaw_amd.fom_run(app, dt, n, src, rank2, rec, every[0], every[1], every[2], precision=shaw_amd.FP32, **one)
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=f"seed {seed} {key}")


def test_autotune_keeps_state_and_parity(mesh64x256):
    """The tile autotuner times candidate shapes on the live state and must
    leave it unchanged; the run it tunes stays bit-identical to the oracle."""
    say("[gpu] autotune")
    rng = np.random.default_rng(77)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    ora = oracle.App(mesh64x256, LAYERED)
    src = random_sources(rng, 24)
    dt, n = 0.5, 90
    fom = shaw_amd.Fom(app, 24, dt)
    fom.set_sources([app.map_source(d, a) for d, a, p, dl in src])
    fom.set_forcing(shaw_amd.forcing_table(["ricker"] * 24, [s[3] for s in src], [s[2] for s in src], dt, n))
    fom.advance(40)
    before = fom.state()
    ms = fom.autotune(2)
    after = fom.state()
    for x, y in zip(before, after):
        np.testing.assert_array_equal(x, y)
    cfg = fom.config()
    assert ms > 0 and cfg["cols_per_thread"] in (1, 2, 4, 8) and cfg["prefetch_rows"] in (1, 2)
    assert cfg["steps_per_launch"] in (1, 2, 3, 4, 5, 6)
    assert cfg["segment"] in (2, 4, 8, 16, 32, 64, 128, 256) and cfg["strip"] in (8, 16, 32, 64, 128, 256)
    fom.close()
    rec = list(np.arange(0, 180, 7.5))
    a = shaw_amd.fom_run(app, dt, n, src, True, rec, 3, 40, 30, autotune=2)
    b = ora.fom_run(dt, n, src, True, rec, 3, 40, 30)
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)


@pytest.mark.parametrize("n,snaps", [(n, False) for n in (1, 2, 7, 8, 13, 14, 15, 20, 27)] +
                         [(n, True) for n in (13, 14, 20, 27)])
def test_last_steps_pack_evenly_bit_exact(mesh64x256, n, snaps):
    """Six steps per launch on one-wavefront strips: a run's last steps (up
    to 14) go in one or two evenly split launches of up to seven steps
    (20 = 6 + 7 + 7, 8 = 4 + 4); packs still never cross a snapshot. Equal
    to the oracle bit for bit."""
    say(f"[gpu] last-step packs n={n} snapshots={snaps}")
    rng = np.random.default_rng(n)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    ora = oracle.App(mesh64x256, LAYERED)
    src = random_sources(rng, 32)
    rec = list(np.arange(0, 180, 7.5))
    ev = (1, 5, 9) if snaps else (1, 0, 0)
    a = shaw_amd.fom_run(app, 0.5, n, src, True, rec, *ev, kernel=shaw_amd.KERNEL_FUSED, steps_per_launch=6,
                         graph_steps=-1)
    b = ora.fom_run(0.5, n, src, True, rec, *ev)
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)


@pytest.mark.parametrize("spl,gsteps", [(1, 8), (2, 8), (3, 12), (4, 8), (5, 10), (6, 12), (6, 0), (7, 14)])
def test_graph_replays_equal_eager(mesh64x256, spl, gsteps):
    """Chunks of gsteps steps (0: the default length) replay one captured HIP
    graph whose kernels read their step from a device counter. Checked against
    eager launches after every call: several replays, eager remainders, host
    state reads and a reset in between, and a forcing table replaced (which
    drops the graph)."""
    say(f"[gpu] graph replays spl={spl}")
    rng = np.random.default_rng(5)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    src = random_sources(rng, 20)
    dt, n = 0.5, 200
    tab = shaw_amd.forcing_table(["ricker"] * 20, [s[3] for s in src], [s[2] for s in src], dt, n)
    rec = [app.map_receiver(a) for a in np.arange(0, 180, 9.0)]

    def make(graph_steps):
        f = shaw_amd.Fom(app, 20, dt, steps_per_launch=spl, graph_steps=graph_steps)
        f.set_sources([app.map_source(d, a) for d, a, p, dl in src])
        f.set_forcing(tab)
        f.set_receivers(rec, 3, n // 3)
        return f

    g, e = make(gsteps), make(-1)
    # the default chunk is 64 steps or more; 25 and 26 leave one or two
    # steps over whole chunks of 12, which then go to the eager packs
    for chunk in (80 if gsteps == 0 else 30, 17, 8, 41, 25, 26):
        g.advance(chunk)
        e.advance(chunk)
        for x, y in zip(g.state(), e.state()):
            np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(g.seismogram(), e.seismogram())
    g.reset()
    e.reset()
    g.set_forcing(tab * 0.5)
    e.set_forcing(tab * 0.5)
    for chunk in (24, 9):
        g.advance(chunk)
        e.advance(chunk)
        for x, y in zip(g.state(), e.state()):
            np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(g.seismogram(), e.seismogram())
    assert np.abs(g.seismogram()).max() > 0
    g.close()
    e.close()


def test_config1_shape_rank1_bit_exact(tmp_path):
    """The config-1 workload shape (200 x 1000, layered mantle, Ricker 640 km /
    45 s at theta=0, dt=0.25, 180 surface receivers every step) for a
    600-step prefix: GPU == oracle bit for bit."""
    say("[gpu] config-1 shape")
    d = str(tmp_path / "mesh200x1000")
    shaw_amd.mesh_write(d, 200, 1000)
    app = shaw_amd.App.from_mesh(d, LAYERED)
    ora = oracle.App(d, LAYERED)
    src = [(640.0, 0.0, 45.0, 67.5)]
    rec = list(range(180))
    a = shaw_amd.fom_run(app, 0.25, 600, src, False, rec, 1)
    b = ora.fom_run(0.25, 600, src, False, rec, 1)
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)


def test_default_segment_fills_the_chip():
    """Default tile shapes: one-wavefront strips (2-column blocks, one thread
    per point); six steps per launch in 128-row segments where a launch fills
    the chip (1024 x 4096 with 32 columns), else four steps loaded two rows
    ahead with segments halved until a launch has at least 1536 wavefronts
    (6 per CU) but not below four rows (200 x 1000 with one column)."""
    say("[gpu] default segments")
    small = shaw_amd.Fom(shaw_amd.App.structured(200, 1000, LAYERED), 1, 0.25)
    cfg = small.config()
    assert (cfg["strip"], cfg["segment"], cfg["steps_per_launch"], cfg["prefetch_rows"], cfg["col_block"]) == \
        (64, 4, 4, 2, 1)
    small.close()
    big = shaw_amd.Fom(shaw_amd.App.structured(1024, 4096, LAYERED), 32, 0.05)
    cfg = big.config()
    assert (cfg["cols_per_thread"], cfg["strip"], cfg["segment"], cfg["steps_per_launch"], cfg["col_block"]) == \
        (2, 64, 128, 6, 2)
    big.close()


@pytest.mark.parametrize("kw", [dict(steps_per_launch=2), dict(steps_per_launch=3, col_block=32),
                                dict(steps_per_launch=4), dict(steps_per_launch=4, cols_per_thread=2, strip=32),
                                dict(steps_per_launch=6), dict(steps_per_launch=5, col_block=1, segment=9),
                                dict(steps_per_launch=7)])
def test_fp32_packs_equal_single_steps(mesh64x256, kw):
    """The FP32 state variant: k steps per launch == k single steps bit for bit
    (the same single-precision expressions in the same order)."""
    say(f"[gpu] fp32 packs {kw}")
    rng = np.random.default_rng(9)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    src = random_sources(rng, 20)
    rec = list(np.arange(0, 180, 5.0))
    a = shaw_amd.fom_run(app, 0.5, 150, src, True, rec, 3, precision=shaw_amd.FP32, **kw)
    b = shaw_amd.fom_run(app, 0.5, 150, src, True, rec, 3, precision=shaw_amd.FP32, steps_per_launch=1)
    for key in b:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert np.abs(a["seismogram"]).max() > 0


def test_fp32_state_error_bound(mesh64x256):
    """FP32 state/arithmetic variant: its seismogram error against FP64 is
    reported; the stencil differences neighbouring values, so single precision
    loses about a digit per derivative and ~1e-3 relative L2 after 300 steps
    is expected. Bound: 2e-2."""
    say("[gpu] fp32 variant")
    rng = np.random.default_rng(5)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    src = random_sources(rng, 16)
    rec = list(np.arange(0, 180, 5.0))
    a = shaw_amd.fom_run(app, 0.5, 300, src, True, rec, 2, precision=shaw_amd.FP32)
    b = shaw_amd.fom_run(app, 0.5, 300, src, True, rec, 2)
    err = max(rel_l2(a["seismogram"][m], b["seismogram"][m]) for m in range(16))
    say(f"[gpu] fp32 seismogram max rel L2 error = {err:.3e}")
    assert err < 2e-2


def test_full_config2_shape_properties():
    """1024 x 4096, M = 32 (config 2's full single-GPU shape), 10 steps:
    fused == split == 2, 3, 4 and 6 steps per launch (and 4 with 8-column blocks) bit for bit, and doubling the
    forcing doubles the state
    exactly (linearity; scaling by 2 is exact in binary floating point)."""
    say("[gpu] config-2 shape properties")
    app = shaw_amd.App.structured(1024, 4096, LAYERED)
    rng = np.random.default_rng(2)
    M, n, dt = 32, 10, 0.05
    src = [(float(rng.uniform(1, 700)), 0.0, float(rng.uniform(20, 70)), 0.0) for _ in range(M)]
    gids = [app.map_source(d, a) for d, a, p, dl in src]
    tab = shaw_amd.forcing_table(["ricker"] * M, [s[3] for s in src], [s[2] for s in src], dt, n)
    tab[:] = 1.0 + np.arange(M)  # large constant forcing so the wave reaches many points in 10 steps
    out = {}
    for name, kern, scale, spl, cb in [("fused", shaw_amd.KERNEL_FUSED, 1.0, 1, 0),
                                       ("split", shaw_amd.KERNEL_SPLIT, 1.0, 1, 0),
                                       ("fused2", shaw_amd.KERNEL_FUSED, 2.0, 1, 0),
                                       ("pairs", shaw_amd.KERNEL_FUSED, 1.0, 2, 0),
                                       ("threes", shaw_amd.KERNEL_FUSED, 1.0, 3, 0),
                                       ("fours", shaw_amd.KERNEL_FUSED, 1.0, 4, 0),
                                       ("fours8", shaw_amd.KERNEL_FUSED, 1.0, 4, 8),
                                       ("sixes", shaw_amd.KERNEL_FUSED, 1.0, 6, 0)]:
        f = shaw_amd.Fom(app, M, dt, kernel=kern, rank1_forcing=False, steps_per_launch=spl, col_block=cb)
        f.set_sources(gids)
        f.set_forcing(tab * scale)
        f.set_receivers([app.map_receiver(a) for a in range(180)], 1, n)
        f.advance(n)
        out[name] = f.state() + (f.seismogram(),)
        f.close()
    for a, b in zip(out["fused"], out["split"]):
        np.testing.assert_array_equal(a, b)
    for key in ("pairs", "threes", "fours", "fours8", "sixes"):  # several steps per launch at the full shape
        for a, b in zip(out["fused"], out[key]):
            np.testing.assert_array_equal(a, b)
    for a, b in zip(out["fused"], out["fused2"]):
        np.testing.assert_array_equal(2.0 * a, b)
    vp = out["fused"][0]
    assert np.count_nonzero(vp) > 32 * 50


def test_full_config2_shape_vs_oracle():
    """1024 x 4096 with 32 seeded sources (config 2's mesh, material, depths
    and periods; delay 0 so the Ricker wavelets are large from the start), 24
    steps = four launches of the default six-step kernel, 180 receivers every
    step: seismograms and final states bit-identical to the oracle's CSR run."""
    say("[gpu] config-2 shape vs oracle")
    from shaw_amd import workloads
    w = workloads.config2()
    mat = workloads.layered_material()
    app = shaw_amd.App.structured(w.nr, w.nth, mat)
    ora = oracle.App("", mat, structured=(w.nr, w.nth, 0.0, 180.0, 3480.0, 6371.0))
    src = [(d, a, p, 0.0) for d, a, p, dl in w.sources]
    rec = [float(x) for x in range(180)]
    n = 24
    got = shaw_amd.fom_run(app, w.dt, n, src, True, rec, 1)
    ref = ora.fom_run(w.dt, n, src, True, rec, 1)
    for key in ref:
        np.testing.assert_array_equal(got[key], ref[key], err_msg=key)
    assert np.count_nonzero(got["final_vp"]) > 32 * 100


def _config5_run(app, w, gids, tab, n, rec, cols, kernel=None, precision=None, every=1, snap=True):
    """One config-5 FOM run of all 64 columns: seismograms [64][T][180] and
    the states of the columns `cols` at step n (snapshots, so the 51 GB
    state is never copied to the host)."""
    kw = {}
    if kernel is not None:
        kw["kernel"] = kernel
    if precision is not None:
        kw["precision"] = precision
    f = shaw_amd.Fom(app, w.m, w.dt, rank1_forcing=False, **kw)
    f.set_sources(gids)
    f.set_forcing(tab)
    f.set_receivers(rec, every, n // every)
    if snap:
        f.set_snapshots("vp", n, cols, 1)
        f.set_snapshots("sp", n, cols, 1)
    f.advance(n)
    f.sync()
    out = [f.seismogram()]
    if snap:
        out += [f.snapshots("vp"), f.snapshots("sp")]
    f.close()
    return out


def test_config5_full_width_vs_oracle_split_and_linearity():
    """Config 5 at its full width: 2048 x 8192, all M = 64 seeded sources
    (delay 0, so the first steps carry signal), 12 steps = two launches of the
    default six-step kernel, 180 receivers every step.
    * Columns 0, 21, 42, 63: seismograms and the states at step 12 are
      bit-identical to the oracle's CSR run of those four sources (the columns
      are independent, so a subset run is the same computation).
    * All 64 columns: the fused run equals the reference's half-step structure
      (split kernels) bit for bit, and doubling the forcing doubles every
      seismogram sample and state value exactly (the step is linear and a
      factor of two commutes with rounding)."""
    say("[gpu] config-5 full width (M = 64) vs oracle subset, split, linearity")
    from shaw_amd import workloads
    w = workloads.config5()
    mat = workloads.layered_material()
    app = shaw_amd.App.structured(w.nr, w.nth, mat)
    src = [(d, a, p, 0.0) for d, a, p, dl in w.sources]
    assert len(src) == 64
    n = 12
    rec = [app.map_receiver(float(a)) for a in range(180)]
    gids = [app.map_source(d, a) for d, a, p, dl in src]
    tab = shaw_amd.forcing_table(["ricker"] * 64, [0.0] * 64, [s[2] for s in src], w.dt, n)
    cols = [0, 21, 42, 63]
    fused = _config5_run(app, w, gids, tab, n, rec, cols)
    ora = oracle.App("", mat, structured=(w.nr, w.nth, 0.0, 180.0, 3480.0, 6371.0))
    ref = ora.fom_run(w.dt, n, [src[c] for c in cols], True, [float(a) for a in range(180)], 1)
    for i, c in enumerate(cols):
        np.testing.assert_array_equal(fused[0][c], ref["seismogram"][i], err_msg=f"seismogram column {c}")
        np.testing.assert_array_equal(fused[1][i, 0], ref["final_vp"][:, i], err_msg=f"vp column {c}")
        np.testing.assert_array_equal(fused[2][i, 0], ref["final_sp"][:, i], err_msg=f"sp column {c}")
    assert np.count_nonzero(fused[1]) > 4 * 20
    say("[gpu] config-5 split kernels")
    split = _config5_run(app, w, gids, tab, n, rec, cols, kernel=shaw_amd.KERNEL_SPLIT)
    for a, b in zip(fused, split):
        np.testing.assert_array_equal(a, b)
    say("[gpu] config-5 doubled forcing")
    double = _config5_run(app, w, gids, 2.0 * tab, n, rec, cols)
    for a, b in zip(fused, double):
        np.testing.assert_array_equal(2.0 * a, b)


def test_config5_fp32_state_error():
    """Config 5's FP32-state variant against the FP64 run over the first 8000
    steps (200 s: every source has fired, delay 105 s, and its waves reach the
    surface receivers), all 64 columns, 180 receivers every 10 steps: the
    relative L2 error of each column's seismogram is at most 5e-4 (max) and
    5e-5 (median)."""
    say("[gpu] config-5 FP32 state vs FP64, 8000 steps")
    from shaw_amd import workloads
    w = workloads.config5()
    mat = workloads.layered_material()
    app = shaw_amd.App.structured(w.nr, w.nth, mat)
    n = 8000
    rec = [app.map_receiver(float(a)) for a in range(180)]
    gids = [app.map_source(d, a) for d, a, p, dl in w.sources]
    tab = shaw_amd.forcing_table(["ricker"] * 64, [s[3] for s in w.sources], [s[2] for s in w.sources], w.dt, n)
    s64 = _config5_run(app, w, gids, tab, n, rec, [], every=10, snap=False)[0]
    s32 = _config5_run(app, w, gids, tab, n, rec, [], precision=shaw_amd.FP32, every=10, snap=False)[0]
    err = np.array([np.linalg.norm(s32[m] - s64[m]) / np.linalg.norm(s64[m]) for m in range(64)])
    say(f"[gpu] config-5 FP32 vs FP64: max {err.max():.3e}, median {np.median(err):.3e}")
    assert np.all(np.isfinite(err)) and np.linalg.norm(s64) > 0
    assert err.max() <= 5e-4 and np.median(err) <= 5e-5


def test_two_handles_interleaved_equal_alone(mesh64x256):
    """Two FOM handles with different shapes (six steps per launch with graph
    chunks; four steps with 8-column blocks), advanced in alternating chunks
    on their own streams: each equals the same handle run alone, bit for bit
    (no state is shared between handles)."""
    say("[gpu] two handles interleaved")
    rng = np.random.default_rng(21)
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    dt, n = 0.5, 120
    shapes = [dict(M=20, kw=dict(steps_per_launch=6, graph_steps=12)),
              dict(M=9, kw=dict(steps_per_launch=4, col_block=8))]
    srcs = [random_sources(rng, sh["M"]) for sh in shapes]
    rec = [app.map_receiver(a) for a in np.arange(0, 180, 9.0)]

    def make(i):
        sh, src = shapes[i], srcs[i]
        f = shaw_amd.Fom(app, sh["M"], dt, **sh["kw"])
        f.set_sources([app.map_source(d, a) for d, a, p, dl in src])
        f.set_forcing(shaw_amd.forcing_table(["ricker"] * sh["M"], [s[3] for s in src], [s[2] for s in src], dt, n))
        f.set_receivers(rec, 3, n // 3)
        return f

    both = [make(0), make(1)]
    for chunk in (30, 25, 41, 24):
        for f in both:
            f.advance(chunk)
    for i, f in enumerate(both):
        alone = make(i)
        alone.advance(n)
        for x, y in zip(f.state(), alone.state()):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(f.seismogram(), alone.seismogram())
        alone.close()
        f.close()


def test_errors_from_gpu_entry_points(mesh64x256):
    app = shaw_amd.App.from_mesh(mesh64x256, LAYERED)
    f = shaw_amd.Fom(app, 4, 0.5)
    with pytest.raises(shaw_amd.ShawError, match="set_sources"):
        f.advance(1)
    f.set_sources([app.map_source(100.0, 10.0)] * 4)
    with pytest.raises(shaw_amd.ShawError, match="forcing table"):
        f.advance(1)
    f.set_forcing(np.zeros((5, 4)))
    f.set_receivers([3, 4], 2, 1)
    with pytest.raises(shaw_amd.ShawError, match="seismogram buffer"):
        f.advance(5)
    with pytest.raises(shaw_amd.ShawError):
        shaw_amd.Fom(app, 4, 0.5, cols_per_thread=3)
