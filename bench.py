#!/usr/bin/env python3
"""Benchmark of the SHAW time-stepping hot path on one MI355X.

Workload (one "step" = one leapfrog step of every forcing column):
  config2 -- plain FOM, FP64, rank-2: 1024 (r) x 4096 (theta) staggered grid,
  layered mantle, M = 32 seeded Ricker forcings, dt = 0.05 s, 180 surface
  receivers sampled every 10 steps (fused into the step kernel).

Prints ONE JSON line: forcing-steps/s (= M * steps / seconds of the timed
region, states resident in HBM), the roofline of the dominant kernel (HIP
events on the stream it runs on), and the CPU baseline (the oracle, OpenMP
C++ restatement of the reference's CSR algorithm, on a bounded sample of the
same workload on this host's cores).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-rom] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ROOF_STEPS = 240  # steps timed apart for the FOM kernel's launch duration (whole launches)
METRIC = "forcing-steps/s, plain FP64 vs batched ROM at err<=1e-3; p50/p99 latency"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fom_fused_config2_traffic.txt")


def measured_traffic(kernel_cfg: dict):
    """HBM bytes per launch of the FOM kernel from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the same
    counter reads for a copy kernel of known size), when it was measured for
    this kernel configuration; otherwise None."""
    try:
        kv = dict(line.split(":", 1) for line in open(TRAFFIC_FILE) if ":" in line and not line.startswith(" "))
    except OSError:
        return None
    kv = {k.strip(): v.strip() for k, v in kv.items()}
    want = (f"cols_per_thread={kernel_cfg['cols_per_thread']} strip={kernel_cfg['strip']} "
            f"segment={kernel_cfg['segment']} prefetch_rows={kernel_cfg['prefetch_rows']} "
            f"steps_per_launch={kernel_cfg['steps_per_launch']} col_block={kernel_cfg['col_block']}")
    if kv.get("kernel_config") != want:
        return None
    return float(kv["traffic_bytes_per_launch"])


def fused_kernel_name(cfg: dict) -> str:
    """The FOM kernel a launch of this configuration runs."""
    spl, cpt = cfg["steps_per_launch"], cfg["cols_per_thread"]
    if spl == 1:
        return f"k_fused<double,{cpt},{cfg['prefetch_rows']}>"
    if spl == 2:
        return f"k_fused2<double,{cpt}>"
    if cpt == cfg["col_block"] and cfg["strip"] == 64 and cpt <= 2:
        return f"k_fusedW<double,{cpt},{spl}>"
    return f"k_fusedN<double,{cpt},{spl}>"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(w, material, sample_steps: int) -> dict:
    """Oracle (C++ CSR restatement, OpenMP) on the first `sample_steps` steps."""
    import oracle

    t0 = time.time()
    app = oracle.App("", material, structured=(w.nr, w.nth, 0.0, 180.0, 3480.0, 6371.0))
    setup = time.time() - t0
    t0 = time.time()
    app.fom_run(w.dt, sample_steps, w.sources, w.m > 1)
    el = time.time() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": w.m * sample_steps / el, "unit": "forcing-steps/s", "cores": cores, "kind": "port",
            "sample": f"{w.name} mesh and forcings, first {sample_steps} of {w.nsteps} steps "
                      f"({el:.1f} s; operator setup {setup:.1f} s not timed)"}


def rom_section(app, w, fom_value, args) -> dict:
    """Config 3 as specified: POD (FP64, K = 256) of 400 + 400 snapshots
    (every 200 steps) of the first 4 config-2 forcings, 1024 seeded forcings
    (32 depths x 32 periods) advanced for --rom-steps steps in FP64 MFMA,
    BF16x3 and BF16x4; seismogram errors against the FP64 FOM on 16
    spot-checked forcings. Run twice on the same training set: with the
    reference's Euclidean Galerkin projection, and with the energy-consistent
    one (POD and projection in the FOM energy weights). The spectral radius
    of each pair's one-step matrix separates instability (radius > 1: the
    Euclidean pair's homogeneous solutions grow radius^steps) from basis
    error (the energy pair's radius is 1; what remains is the basis: 4
    training depths cannot span 32 test depths at K = 256)."""
    import shaw_amd
    from shaw_amd import rom_workflow as rw, workloads

    n = args.rom_steps
    train = [w.sources[8 * i] for i in range(4)]  # one per config-2 depth
    d, p = workloads.seeded_depths_periods(32, 32, 3)
    test = [(dd, 0.0, pp, workloads.DELAY_S) for dd in d for pp in p]
    spot = list(range(0, len(test), len(test) // 16))[:16]
    fom = rw.fom_seismograms(app, [test[j] for j in spot], w.dt, n, workloads.RECEIVERS, 10)
    out = {"workload": f"config3: K=256, POD FP64 of 400+400 snapshots of 4 config-2 forcings, M={len(test)} "
                       f"forcings (32 depths x 32 periods), {n} steps, 180 receivers every 10 steps"}
    for energy, tag in ((False, "euclidean"), (True, "energy")):
        t0 = time.time()
        basis = rw.train_basis(app, train, w.dt, n, max(1, n // 100), 256, energy=energy)
        t1 = time.time()
        arv, ars = rw.operators(app, basis, 256)
        t2 = time.time()
        B, ftab = rw.basis_forcing_inputs(app, basis, 256, test, w.dt, n)
        C = rw.receiver_rows(app, basis, workloads.RECEIVERS)
        t3 = time.time()
        res, rates = {}, {}
        for prec, name, macro in rom_variants(args):
            res[name], rates[name] = run_rom(arv, ars, B, ftab, C, w.dt, n, prec, macro)
            log(f"config3 ROM ({tag}) K=256 {name}: {rates[name]:.3e} forcing-steps/s")
        err_fom = {k: [rw.rel_l2(res[k][j], fom[i]) for i, j in enumerate(spot)] for k in res}
        err_split = {k: summary([rw.rel_l2(res[k][j], res["fp64"][j]) for j in range(len(test))])
                     for k in res if k != "fp64"}
        rho = rw.amplification_radius(arv, ars, w.dt)
        out[tag] = {"projection": "reference (Euclidean Galerkin)" if not energy else
                    "energy-consistent (POD and projection in the FOM energy weights)",
                    "forcing_steps_per_s": rates,
                    "speedup_vs_fp64_fom": {k: v / fom_value for k, v in rates.items()},
                    "mfma_tflops": mfma_tflops(256, rates),
                    "err_vs_fp64_fom_16_spot": {k: summary(v) for k, v in err_fom.items()},
                    "err_vs_fp64_rom": err_split,
                    "amplification_radius": rho, "amplification_over_run": rho ** n,
                    "pod_sigma_ratio_256": float(basis.sigma_v[255] / basis.sigma_v[0]),
                    "setup_s": {"fom_training": basis.seconds["fom_training"], "pod": basis.seconds["pod"],
                                "projection": t2 - t1, "inputs": t3 - t2, "total": t3 - t0}}
        log(f"config3 ({tag}): radius {rho:.8f}, errors vs FOM { {k: summary(v) for k, v in err_fom.items()} }")
    return out


def run_rom(arv, ars, B, ftab, C, dt, n, prec, macro=0):
    """Seismograms [M][T][R] and forcing-steps/s of one batched ROM run (HIP
    events around the persistent kernels on the handle's stream); macro > 0:
    BF16 tiles whose forcing has ended advance `macro` steps per product."""
    import shaw_amd

    m = B.shape[1]
    rom = shaw_amd.Rom(arv, ars, dt, prec, C, max_cols=m, max_samples=n // 10)
    if macro:
        rom.set_macro_steps(macro)
    slots = rom.admit(B, ftab, [n] * m, 10)
    ms = rom.time(n)
    seis = np.stack([rom.retire(s_, n // 10)[0] for s_ in slots])
    rom.close()
    return seis, m * n / (ms / 1e3)


def rom_variants(args):
    """(precision, name, macro steps) of the ROM runs: FP64, BF16x3 and
    BF16x4 one step at a time, and the BF16 ones with macro steps."""
    import shaw_amd

    out = [(shaw_amd.ROM_FP64, "fp64", 0), (shaw_amd.ROM_BF16X3, "bf16x3", 0), (shaw_amd.ROM_BF16X4, "bf16x4", 0)]
    if args.rom_macro:
        out += [(shaw_amd.ROM_BF16X3, "bf16x3_macro", args.rom_macro),
                (shaw_amd.ROM_BF16X4, "bf16x4_macro", args.rom_macro)]
    return out


def summary(v) -> dict:
    return {"max": float(np.max(v)), "median": float(np.median(v))}


def mfma_tflops(k: int, rates: dict) -> dict:
    """MFMA work rate of the one-step runs: 4 K^2 flops per forcing-step,
    times the products per multiply-add (3 for BF16x3, 4 for BF16x4)."""
    mult = {"fp64": 1, "bf16x3": 3, "bf16x4": 4}
    return {name: 4.0 * k * k * mult[name] * r / 1e12 for name, r in rates.items() if name in mult}


def rom_err_section(app, w, fom_value, args):
    """The ROM at err <= 1e-3: the energy-consistent Galerkin ROM for one
    source depth (640 km, theta = 0), trained on 5 periods (20, 25, 35, 50,
    70 s) with snapshots every 100 steps (200 per forcing, 1000 per dof), POD
    in the FOM energy weights to 512 modes (nested), K = 384. M = 1024 forcings with
    seeded periods in [20, 70] s, 20000 steps, in FP64, BF16x3 and BF16x4;
    errors against the FP64 FOM on 16 spot-checked forcings. Returns the
    section and the basis (reused by the serving section)."""
    import shaw_amd
    from shaw_amd import rom_workflow as rw, workloads

    n, K, depth = args.rom_steps, 384, 640.0
    t0 = time.time()
    train = [(depth, 0.0, p, workloads.DELAY_S) for p in (20.0, 25.0, 35.0, 50.0, 70.0)]
    basis = rw.train_basis(app, train, w.dt, n, max(1, n // 200), 512, energy=True)
    arv, ars = rw.operators(app, basis, K)
    periods = np.random.default_rng(640).uniform(20.0, 70.0, 1024)
    test = [(depth, 0.0, float(pp), workloads.DELAY_S) for pp in periods]
    B, ftab = rw.basis_forcing_inputs(app, basis, K, test, w.dt, n)
    C = rw.receiver_rows(app, basis, workloads.RECEIVERS, K)
    setup = time.time() - t0
    res, rates = {}, {}
    for prec, name, macro in rom_variants(args):
        res[name], rates[name] = run_rom(arv, ars, B, ftab, C, w.dt, n, prec, macro)
        log(f"energy ROM K={K} {name}: {rates[name]:.3e} forcing-steps/s")
    spot = list(range(0, len(test), len(test) // 16))[:16]
    fom = rw.fom_seismograms(app, [test[j] for j in spot], w.dt, n, workloads.RECEIVERS, 10)
    err_fom = {k: [rw.rel_l2(res[k][j], fom[i]) for i, j in enumerate(spot)] for k in res}
    err_split = {k: summary([rw.rel_l2(res[k][j], res["fp64"][j]) for j in range(len(test))])
                 for k in res if k != "fp64"}
    ok = {k: bool(max(v) <= 1e-3) for k, v in err_fom.items()}
    log(f"energy ROM errors vs FP64 FOM: { {k: summary(v) for k, v in err_fom.items()} }")
    out = {"workload": f"energy-consistent Galerkin ROM K={K}, one source depth ({depth:g} km), POD in the FOM "
                       "energy weights of 1000+1000 snapshots; M=1024 forcings with seeded periods in "
                       f"[20, 70] s, {n} steps, 180 receivers every 10 steps",
           "training": "5 forcings (periods 20, 25, 35, 50, 70 s) at the same depth, snapshots every 100 steps "
                       "(200 per forcing, 1000 per dof)",
           "forcing_steps_per_s": rates,
           "speedup_vs_fp64_fom": {k: v / fom_value for k, v in rates.items()},
           "mfma_tflops": mfma_tflops(K, rates),
           "err_vs_fp64_fom_16_spot": {k: summary(v) for k, v in err_fom.items()},
           "err_vs_fp64_rom": err_split,
           "every_spot_within_1e-3": ok,
           "pod_sigma_ratio": {"384": float(basis.sigma_v[383] / basis.sigma_v[0]),
                               "512": float(basis.sigma_v[511] / basis.sigma_v[0])},
           "setup_s": setup}
    return out, basis


def serving_section(app, w, basis, args) -> dict:
    """Config 4 on the energy-consistent basis of the ROM section, served by
    the C++ loop (shaw_serve): Poisson arrivals at 200 requests/s for
    --serve-seconds, each with M ~ U{1..256} forcings at the basis's source
    depth (periods ~ U[20, 70] s), a ROM size K ~ U{128, 256, 512} (nested
    modes of one POD) and a horizon ~ U[5000, 20000] steps; one resident
    BF16x4 engine per K; admission by HBM footprint, least work first with a
    wait bound; p50/p99 latency, the throughput and the longest queue. A
    second run at 300 requests/s keeps the queue growing, so its throughput
    is the engines' drain rate. Sampled requests'
    seismograms are checked against the FP64 ROM of the same K and the FP64
    FOM; the other requests' seismograms are discarded on retirement."""
    import shaw_amd
    from shaw_amd import rom_workflow as rw, serving, workloads

    t0 = time.time()
    depth = 640.0
    _, _, rho_inv, _ = app.operators()
    g = app.map_source(depth, 0.0)
    C512 = rw.receiver_rows(app, basis, workloads.RECEIVERS)
    ops, engines = {}, {}
    for k in (128, 256, 512):
        ops[k] = rw.operators(app, basis, k)
        group = []
        for _ in range(args.serve_engines_512 if k == 512 else 1):
            rom = shaw_amd.Rom(ops[k][0], ops[k][1], w.dt, shaw_amd.ROM_BF16X4, C512[:, :k], max_cols=4096,
                               max_samples=2000)
            if args.rom_macro:
                rom.set_macro_steps(args.rom_macro)
                if args.serve_time_budget:
                    rom.set_time_budget(args.serve_time_budget)
            group.append(serving.GpuEngine(rom, 4096))
        engines[k] = group if len(group) > 1 else group[0]
    bcol = basis.mv[g] * rho_inv[g] * basis.dev.rows("vp", [g])[0]
    setup = time.time() - t0
    mean_work = 128.5 * 12500.0  # E[M] E[horizon] forcing-steps per request

    drawn = []  # periods per request, in request-id order

    def make_inputs(k, m, h, gen):
        # the request carries its forcing as signal parameters; the engine
        # evaluates the series on the GPU at admission
        ps = gen.uniform(20.0, 70.0, m)
        drawn.append(ps)
        b = np.repeat(bcol[:k, None], m, axis=1)
        return b, None, (["ricker"] * m, [workloads.DELAY_S] * m, list(ps))

    def one_run(rate, seconds, seed, keep_every):
        drawn.clear()
        reqs = serving.poisson_requests(np.random.default_rng(seed), rate, seconds, [128, 256, 512], (1, 256),
                                        (5000, 20000), make_inputs)
        # keep a few requests' seismograms; the rest are discarded on retirement
        kept = [r for r in reqs if r.rid % keep_every == 0][:6]
        for r in reqs:
            r.keep = r in kept
        srv = serving.Server(engines, hbm_budget=200e9, chunk=args.serve_chunk, policy="least_work", max_wait=2.0)
        rep = srv.run(reqs, wall_limit=900.0, keep=lambda r: r.keep)
        rep["offered_forcing_steps_per_s"] = rate * mean_work
        rep["arrival_rate_per_s"] = rate
        rep["arrival_seconds"] = seconds
        return reqs, kept, rep

    reqs, kept, rep = one_run(200.0, args.serve_seconds, 4, 37)
    log(f"serving 200 req/s: {rep}")
    # sampled per-request errors (first column of each kept request)
    e_rom, e_fom = [], []
    hmax = max(r.horizon for r in kept)
    periods = [float(drawn[r.rid][0]) for r in kept]
    fom = rw.fom_seismograms(app, [(depth, 0.0, p_, workloads.DELAY_S) for p_ in periods], w.dt, hmax,
                             workloads.RECEIVERS, 10)
    for i, r in enumerate(kept):
        arv, ars = ops[r.k]
        f1 = shaw_amd.forcing_table(["ricker"], [workloads.DELAY_S], [periods[i]], w.dt, r.horizon)
        ref = shaw_amd.rom_run(arv, ars, r.b[:, :1], f1, w.dt, r.horizon, C512[:, :r.k], r.every)
        e_rom.append(rw.rel_l2(r.seismograms[0], ref["seismogram"][0]))
        e_fom.append(rw.rel_l2(r.seismograms[0], fom[i][:r.horizon // r.every]))
    rep["sampled_requests"] = [{"rid": r.rid, "K": r.k, "M": r.m, "horizon": r.horizon,
                                "err_vs_fp64_rom": a, "err_vs_fp64_fom": b_}
                               for r, a, b_ in zip(kept, e_rom, e_fom)]
    del reqs, kept
    # the drain capacity: arrivals at 1.5x the rate, so that the queue grows
    # and the engines never wait for work
    rate2 = 1.5 * 200.0
    _, _, rep2 = one_run(rate2, args.serve_seconds, 5, 10 ** 9)
    log(f"serving {rate2:.1f} req/s: {rep2}")
    rep["overload"] = {k: rep2[k] for k in ("arrival_rate_per_s", "offered_forcing_steps_per_s", "requests",
                                             "p50_latency_s", "p99_latency_s", "max_latency_s", "max_wait_s",
                                             "throughput_forcing_steps_per_s", "wall_s", "launches", "max_queue")}
    for e in engines.values():
        for eng in (e if isinstance(e, list) else [e]):
            eng.rom.close()
    rep.update({"workload": f"config4 on the energy-consistent basis: Poisson 200 req/s for {args.serve_seconds:g} s, "
                            "M~U{1..256}, K~U{128,256,512}, horizon~U[5000,20000] steps, source depth 640 km, "
                            "periods~U[20,70] s (forcing evaluated on the GPU at admission), BF16x4 engines"
                            + (f" with macro steps of {args.rom_macro}" if args.rom_macro else "")
                            + (f" and a time budget per chunk (a macro step counted as {args.serve_time_budget} "
                               "one-steps)" if args.rom_macro and args.serve_time_budget else "") + ", "
                            "180 receivers every 10 steps, C++ loop "
                            "(least work first, 2 s wait bound); then 10 s at 300 req/s for the drain rate",
                "setup_s": setup})
    return rep


def config1_section(args) -> dict:
    """Config 1, the reference's CPU-runnable case: 200 x 1000, one Ricker
    source (640 km, 45 s) at theta = 0, rank-1 forcing form, dt = 0.25 s,
    4000 steps, 180 receivers every step. The whole run on the GPU with HIP
    graph replays (default) and with eager launches, the oracle's whole run
    on the host cores, and the seismograms compared (bit-identical)."""
    import oracle
    import shaw_amd
    from shaw_amd import workloads

    w = workloads.config1()
    mat = workloads.layered_material()
    app = shaw_amd.App.structured(w.nr, w.nth, mat)
    out, seis = {}, {}
    for name, gsteps in (("graphs", 0), ("eager", -1)):
        fom = shaw_amd.Fom(app, 1, w.dt, rank1_forcing=True, graph_steps=gsteps)
        fom.set_sources([app.map_source(d, a) for d, a, p, dl in w.sources])
        fom.set_forcing(shaw_amd.forcing_table(["ricker"], [w.sources[0][3]], [w.sources[0][2]], w.dt, w.nsteps))
        fom.set_receivers([app.map_receiver(a) for a in workloads.RECEIVERS], 1, w.nsteps)
        fom.advance(16)  # warm (and capture)
        fom.reset()
        total, _ = fom.time(w.nsteps)
        seis[name] = fom.seismogram()
        out[name] = {"seconds": total / 1e3, "forcing_steps_per_s": w.nsteps / (total / 1e3)}
        fom.close()
    t0 = time.time()
    ora = oracle.App("", mat, structured=(w.nr, w.nth, 0.0, 180.0, 3480.0, 6371.0))
    t1 = time.time()
    ref = ora.fom_run(w.dt, w.nsteps, w.sources, False, workloads.RECEIVERS, 1)
    t2 = time.time()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out["cpu_oracle"] = {"seconds": t2 - t1, "forcing_steps_per_s": w.nsteps / (t2 - t1), "cores": cores,
                         "setup_s": t1 - t0}
    out["gpu_equals_oracle"] = bool(np.array_equal(seis["graphs"][0], ref["seismogram"][0]))
    out["graphs_equal_eager"] = bool(np.array_equal(seis["graphs"], seis["eager"]))
    out["workload"] = w.describe() + ", 180 receivers every step, rank-1 forcing form"
    log(f"config1: {out}")
    return out


def config5_section(args) -> dict:
    """Config 5: the large FOM, 2048 x 8192, M = 64 seeded forcings,
    dt = 0.025 s. FP64 and FP32 states, each with the autotuned tile shape,
    timed over --steps steps after the warmup (states in HBM; the full run is
    40000 steps); the FP32 state's seismogram error against FP64 over the
    first --c5-err-steps steps (180 receivers every 10 steps)."""
    import shaw_amd
    from shaw_amd import workloads

    w = workloads.config5()
    t0 = time.time()
    app = shaw_amd.App.structured(w.nr, w.nth, workloads.layered_material())
    gids = [app.map_source(d, a) for d, a, p, dl in w.sources]
    rec = [app.map_receiver(a) for a in workloads.RECEIVERS]
    n_err = args.c5_err_steps
    n_tab = max(n_err, args.warmup + args.steps + 1)
    tab = shaw_amd.forcing_table(["ricker"] * w.m, [s[3] for s in w.sources], [s[2] for s in w.sources], w.dt, n_tab)
    log(f"config5 setup {time.time() - t0:.1f} s")
    out, seis = {}, {}
    for prec, name in ((shaw_amd.FP64, "fp64"), (shaw_amd.FP32, "fp32")):
        fom = shaw_amd.Fom(app, w.m, w.dt, precision=prec, rank1_forcing=False)
        fom.set_sources(gids)
        fom.set_forcing(tab)
        fom.set_receivers(rec, w.seis_every, n_tab // w.seis_every)
        tuned = fom.autotune(3)
        cfg = fom.config()
        fom.advance(args.warmup)
        fom.sync()
        total, per = fom.time(args.steps)
        kb, mb = fom.bytes_per_step()
        out[name] = {"ms_per_step": per, "forcing_steps_per_s": w.m / (per / 1e3),
                     "algorithmic_GB_s": kb / (per / 1e3) / 1e9, "hbm_frac": kb / (per / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "bytes_per_step": kb, "autotuned": {"config": cfg, "trial_ms_per_step": tuned}}
        log(f"config5 {name}: {per:.3f} ms/step, {kb / (per / 1e3) / 1e9:.0f} GB/s, config {cfg}")
        fom.close()
        # error run from zero state over the first n_err steps
        fom = shaw_amd.Fom(app, w.m, w.dt, precision=prec, rank1_forcing=False, cols_per_thread=cfg["cols_per_thread"],
                           strip=cfg["strip"], segment=cfg["segment"], prefetch_rows=cfg["prefetch_rows"],
                           steps_per_launch=cfg["steps_per_launch"], col_block=cfg["col_block"])
        fom.set_sources(gids)
        fom.set_forcing(tab)
        fom.set_receivers(rec, w.seis_every, n_err // w.seis_every)
        fom.advance(n_err)
        fom.sync()
        seis[name] = fom.seismogram()
        fom.close()
    err = [float(np.linalg.norm(seis["fp32"][m] - seis["fp64"][m]) / np.linalg.norm(seis["fp64"][m]))
           for m in range(w.m)]
    out["fp32_vs_fp64_seismogram_rel_l2"] = {"max": max(err), "median": float(np.median(err)),
                                             "steps": n_err, "sim_time_s": n_err * w.dt}
    out["workload"] = w.describe() + ", 180 receivers every 10 steps"
    return out


def rom_kernel_only(args) -> int:
    """The batched reduced leapfrog alone (k_rom_steps, k_rom_steps_bf), for profiling: K x K
    skew operator pairs (random, seeded), M columns, 180 receivers every 10
    steps, --rom-steps steps timed per precision with HIP events on the
    engine's stream. Prints one JSON line of forcing-steps/s and MFMA TFLOP/s.
    With --rom-kernel-quiet-after Q > 0 each column's forcing ends at a seeded
    step in [Q/2, Q] and the BF16 precisions run again with macro steps
    (--rom-macro, k_rom_steps_bfm)."""
    import shaw_amd

    K, M, n = args.rom_kernel_k, args.rom_kernel_m, args.rom_steps
    rng = np.random.default_rng(K)
    A = rng.standard_normal((K, K)) * (0.5 / np.sqrt(K))
    C = rng.standard_normal((180, K))
    B = rng.standard_normal((K, M))
    f = rng.standard_normal((n + 50, M))
    q = args.rom_kernel_quiet_after
    if q > 0:
        ends = rng.integers(q // 2, q + 1, M)
        for j in range(M):
            f[50 + ends[j]:, j] = 0.0
    out = {"workload": f"batched ROM kernel only: K={K}, M={M}, {n} steps, synthetic skew operators"
                       + (f", forcing ending at steps in [{q // 2}, {q}]" if q > 0 else "")}
    names = {"fp64": shaw_amd.ROM_FP64, "bf16x3": shaw_amd.ROM_BF16X3, "bf16x4": shaw_amd.ROM_BF16X4}
    runs = [(name, 0) for name in args.rom_kernel_prec.split(",")]
    if q > 0 and args.rom_macro:
        runs += [(name, args.rom_macro) for name in args.rom_kernel_prec.split(",") if name != "fp64"]
    for name, macro in runs:
        rom = shaw_amd.Rom(A, -A.T.copy(), 0.05, names[name], C, max_cols=M, max_samples=(n + 50) // 10)
        if macro:
            rom.set_macro_steps(macro)
        rom.admit(B, f, [n + 50] * M, 10)
        rom.advance(50)
        rom.sync()
        ms = rom.time(n)
        rom.close()
        rate = M * n / (ms / 1e3)
        key = name + (f"_macro{macro}" if macro else "")
        out[key] = {"ms": ms, "forcing_steps_per_s": rate}
        if not macro:
            out[key]["mfma_tflops"] = mfma_tflops(K, {name: rate})[name]
        log(f"ROM kernel {key}: {rate:.3e} forcing-steps/s")
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["fused", "split"], default="fused")
    ap.add_argument("--cpt", type=int, default=0)
    ap.add_argument("--strip", type=int, default=0)
    ap.add_argument("--segment", type=int, default=0)
    ap.add_argument("--prefetch", type=int, default=0)
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--col-block", type=int, default=0, help="FOM columns stored together per grid point")
    ap.add_argument("--copy-bytes", type=int, default=4 << 30)
    ap.add_argument("--no-rom", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--rom-steps", type=int, default=20000)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--no-config5", action="store_true", help="skip the large FOM (2048 x 8192, M = 64)")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--c5-err-steps", type=int, default=8000)
    ap.add_argument("--rom-kernel", action="store_true", help="time only the batched ROM kernel (profiling)")
    ap.add_argument("--rom-kernel-k", type=int, default=256)
    ap.add_argument("--rom-kernel-m", type=int, default=1024)
    ap.add_argument("--rom-kernel-prec", default="fp64,bf16x3,bf16x4")
    ap.add_argument("--serve-engines-512", type=int, default=1,
                    help="serving: K = 512 engines (each on its own stream) taking the K = 512 requests in turn")
    ap.add_argument("--rom-kernel-quiet-after", type=int, default=0,
                    help="--rom-kernel: forcing ends at seeded steps up to this (0: forcing throughout)")
    ap.add_argument("--rom-macro", type=int, default=10,
                    help="macro steps of the BF16 ROM runs and serving engines (0: one step at a time only)")
    ap.add_argument("--serve-chunk", type=int, default=1500, help="serving: steps (or one-step times) per chunk")
    ap.add_argument("--serve-time-budget", type=int, default=2,
                    help="serving engines with macro steps: a chunk is the time of chunk one-steps, a macro step "
                         "counting as this many (0: chunks of steps)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (this build runs the hot path on a single MI355X; "
              f"--gpus {args.gpus} / WORLD_SIZE {world} is not supported)", file=sys.stderr)
        return 2

    if args.rom_kernel:
        return rom_kernel_only(args)

    import shaw_amd
    from shaw_amd import workloads

    w = workloads.config2()
    material = workloads.layered_material()
    # after the timed region: ROOF_STEPS more steps, timed apart, in whole
    # launches of the default kernel, give that kernel's launch duration; then
    # the timed steps once more (steady_repeat)
    n_total = args.warmup + 2 * args.steps + ROOF_STEPS
    log(f"workload {w.describe()}; warmup {args.warmup}, timed {args.steps}")
    t0 = time.time()
    app = shaw_amd.App.structured(w.nr, w.nth, material)
    gids = [app.map_source(d, a) for d, a, p, dl in w.sources]
    tab = shaw_amd.forcing_table(["ricker"] * w.m, [s[3] for s in w.sources], [s[2] for s in w.sources],
                                 w.dt, n_total)
    kern = shaw_amd.KERNEL_FUSED if args.kernel == "fused" else shaw_amd.KERNEL_SPLIT
    fom = shaw_amd.Fom(app, w.m, w.dt, kernel=kern, rank1_forcing=False, cols_per_thread=args.cpt,
                       strip=args.strip, segment=args.segment, prefetch_rows=args.prefetch,
                       steps_per_launch=args.steps_per_launch, col_block=args.col_block)
    fom.set_sources(gids)
    fom.set_forcing(tab)
    fom.set_receivers([app.map_receiver(a) for a in workloads.RECEIVERS], w.seis_every,
                      n_total // w.seis_every)
    log(f"setup {time.time() - t0:.1f} s; config {fom.config()}")
    fom.advance(args.warmup)
    fom.sync()
    total_ms, ms_per_step = fom.time(args.steps)
    fom.sync()
    value = w.m * args.steps / (total_ms / 1000.0)
    log(f"{total_ms:.2f} ms for {args.steps} steps -> {ms_per_step:.4f} ms/step, {value:.0f} forcing-steps/s")
    kernel_bytes, model_bytes = fom.bytes_per_step()
    cfg = fom.config()
    spl = cfg["steps_per_launch"]  # a launch of the k-step kernel advances k steps
    # the dominant kernel alone: whole launches, HIP events on the FOM stream
    roof_launches = ROOF_STEPS // spl
    roof_ms, _ = fom.time(roof_launches * spl)
    launch_ms = roof_ms / roof_launches
    bytes_per_launch = kernel_bytes * spl
    achieved = bytes_per_launch / (launch_ms / 1000.0) / 1e9
    log(f"{roof_launches} launches of {spl} steps: {launch_ms:.4f} ms per launch, {achieved:.0f} GB/s algorithmic "
        f"({bytes_per_launch / 1e9:.3f} GB per launch), {roof_launches * spl / (roof_ms / 1e3) * w.m:.0f} "
        "forcing-steps/s over that region")
    # the same run's steady state: the timed region again, right after the
    # roofline launches (the first timed region starts from an idle GPU)
    steady_ms, _ = fom.time(args.steps)
    log(f"steady: {steady_ms:.2f} ms for {args.steps} steps ({args.steps / steady_ms * w.m * 1e3:.0f} forcing-steps/s)")
    fom.sync()
    seis = fom.seismogram()
    assert np.isfinite(seis).all(), "seismogram is not finite"

    traffic = measured_traffic(fom.config())
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "forcing-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded layered mantle and Ricker sources)",
        "config": {"workload": w.describe() + ", 180 receivers every 10 steps",
                   "kernel": cfg, "fom_variant": "plain FP64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": (fused_kernel_name(cfg) if args.kernel == "fused" else "k_velocity+k_stress"),
                     "steps_per_launch": spl,
                     "bytes_per_launch": bytes_per_launch,
                     "launch_ms": launch_ms,
                     "launches_timed": roof_launches,
                     # the counter-measured HBM bytes per launch over this run's launch time
                     "traffic_GB_s": traffic / (launch_ms / 1e3) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "reference_model_bytes_per_step": model_bytes},
        # the timed steps repeated once the GPU is busy: value's figure
        # without the ramp from idle of its first launch
        "steady_repeat": {"ms_per_step": steady_ms / args.steps,
                          "forcing_steps_per_s": w.m * args.steps / (steady_ms / 1e3),
                          "value_over_steady": value / (w.m * args.steps / (steady_ms / 1e3))},
    }
    fom.close()
    if args.copy_bytes:
        ms = shaw_amd.debug_copy(args.copy_bytes, 10)
        out["copy_kernel"] = {"bytes": args.copy_bytes, "ms": ms,
                              "GB_s": 2 * args.copy_bytes / (ms / 1e3) / 1e9,
                              "note": "plain read+write streaming copy (k_copy), achievable-bandwidth reference"}
        log(f"copy kernel: {out['copy_kernel']['GB_s']:.0f} GB/s")
    if not args.no_rom:
        out["rom_config3"] = rom_section(app, w, value, args)
    basis = None
    if not args.no_rom or not args.no_serve:
        out["rom_at_err_1e-3"], basis = rom_err_section(app, w, value, args)
    if not args.no_serve:
        out["serving"] = serving_section(app, w, basis, args)
    if not args.no_config1:
        out["fom_config1"] = config1_section(args)
    if not args.no_config5:
        out["fom_config5"] = config5_section(args)
    if not args.no_cpu:
        log(f"cpu baseline: oracle, {args.cpu_steps} steps")
        out["cpu_baseline"] = cpu_baseline(w, material, args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
