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
METRIC = "forcing-steps/s, plain FP64 vs batched ROM at err<=1e-3; p50/p99 latency"


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


def rom_section(app, w, material, fom_value, args) -> dict:
    """Config 3: POD (FP64, K = 256) from 400 + 400 snapshots of 4 config-2
    forcings, Galerkin operators, 1024 seeded forcings advanced for
    --rom-steps steps in FP64 MFMA and in BF16x3; seismogram errors against
    the FP64 FOM on 16 spot-checked forcings."""
    import shaw_amd
    from shaw_amd import rom_workflow as rw, workloads

    n = args.rom_steps
    train = [w.sources[8 * i] for i in range(4)]  # one per config-2 depth
    t0 = time.time()
    basis = rw.train_basis(app, train, w.dt, n, max(1, n // 100), 256)
    arv, ars = rw.operators(app, basis, 256)
    d, p = workloads.seeded_depths_periods(32, 32, 3)
    test = [(dd, 0.0, pp, workloads.DELAY_S) for dd in d for pp in p]
    pv = basis.phi_v
    B, ftab = rw.forcing_inputs(app, pv, test, w.dt, n)
    C = rw.receiver_rows(app, pv, workloads.RECEIVERS)
    setup = time.time() - t0
    res, rates = {}, {}
    for prec, name in ((shaw_amd.ROM_FP64, "fp64"), (shaw_amd.ROM_BF16X3, "bf16x3")):
        rom = shaw_amd.Rom(arv, ars, w.dt, prec, C, max_cols=len(test), max_samples=n // 10)
        slots = rom.admit(B, ftab, [n] * len(test), 10)
        ms = rom.time(n)
        res[name] = np.stack([rom.retire(s_, n // 10)[0] for s_ in slots])
        rom.close()
        rates[name] = len(test) * n / (ms / 1e3)
        log(f"ROM K=256 {name}: {ms:.1f} ms -> {rates[name]:.3e} forcing-steps/s")
    spot = list(range(0, len(test), len(test) // 16))[:16]
    fom = rw.fom_seismograms(app, [test[j] for j in spot], w.dt, n, workloads.RECEIVERS, 10)
    err_fom = {k: [rw.rel_l2(res[k][j], fom[i]) for i, j in enumerate(spot)] for k in res}
    err_split = [rw.rel_l2(res["bf16x3"][j], res["fp64"][j]) for j in range(len(test))]
    flops = 4.0 * 256 * 256 * len(test) * n
    return {"workload": f"config3: Galerkin ROM K=256 (POD FP64 of 400+400 snapshots of 4 config-2 forcings), "
                        f"M={len(test)} forcings, {n} steps, 180 receivers every 10 steps",
            "fp64_forcing_steps_per_s": rates["fp64"], "bf16x3_forcing_steps_per_s": rates["bf16x3"],
            "speedup_vs_fp64_fom": {k: v / fom_value for k, v in rates.items()},
            "mfma_tflops": {k: flops * (3 if k == "bf16x3" else 1) / (len(test) * n / v) / 1e12
                            for k, v in rates.items()},
            "err_vs_fp64_fom_16_spot": {k: {"max": max(v), "median": float(np.median(v))} for k, v in err_fom.items()},
            "err_bf16x3_vs_fp64_rom": {"max": max(err_split), "median": float(np.median(err_split))},
            "meets_err_1e-3": bool(max(err_fom["bf16x3"]) <= 1e-3),
            "pod_sigma_ratio_256": float(basis.sigma_v[255] / basis.sigma_v[0]),
            "setup_s": setup}


def serving_section(app, w, args) -> dict:
    """Config 4: Poisson arrivals at 200 requests/s for --serve-seconds, each
    with M ~ U{1..256} forcings, K ~ U{128, 256, 512} (nested POD basis of 800
    snapshots of 8 forcings), horizon ~ U[5000, 20000] steps; BF16x3 engines
    resident per K; admission by HBM footprint; p50/p99 latency."""
    import shaw_amd
    from shaw_amd import rom_workflow as rw, serving, workloads

    t0 = time.time()
    d, p = workloads.seeded_depths_periods(4, 2, 17)
    train = [(dd, 0.0, pp, workloads.DELAY_S) for dd in d for pp in p]
    basis = rw.train_basis(app, train, w.dt, 20000, 200, 512)
    _, _, rho_inv, _ = app.operators()
    C512 = rw.receiver_rows(app, basis.phi_v, workloads.RECEIVERS)
    engines = {}
    for k in (128, 256, 512):
        arv, ars = rw.operators(app, basis, k)
        rom = shaw_amd.Rom(arv, ars, w.dt, shaw_amd.ROM_BF16X3, C512[:, :k], max_cols=2048, max_samples=2000)
        engines[k] = serving.GpuEngine(rom, 2048)
    # forcing inputs drawn per request from a seeded pool of source depths/periods
    dd, pp = workloads.seeded_depths_periods(64, 64, 23)
    gid_of = {x: app.map_source(x, 0.0) for x in dd}
    rng = np.random.default_rng(4)
    drawn = []  # (depths, periods) per request, in request-id order

    def make_inputs(k, m, h, g):
        ds = g.choice(dd, m)
        ps = g.choice(pp, m)
        drawn.append((ds, ps))
        b = np.stack([rho_inv[gid_of[x]] * basis.phi_v[gid_of[x], :k] for x in ds], axis=1)
        f = shaw_amd.forcing_table(["ricker"] * m, [workloads.DELAY_S] * m, list(ps), w.dt, h)
        return b, f

    reqs = serving.poisson_requests(rng, 200.0, args.serve_seconds, [128, 256, 512], (1, 256), (5000, 20000),
                                    make_inputs)
    setup = time.time() - t0
    srv = serving.Server(engines, hbm_budget=96e9, chunk=1000)
    rep = srv.run(reqs, wall_limit=600.0)
    log(f"serving: {rep}")
    for e in engines.values():
        e.rom.close()
    # per-request error of sampled requests (first column): against the FP64
    # ROM of the same K and against the FP64 FOM
    sample = reqs[:: max(1, len(reqs) // 4)][:4]
    e_rom, e_fom = [], []
    for r in sample:
        arv, ars = rw.operators(app, basis, r.k)
        ref = shaw_amd.rom_run(arv, ars, r.b[:, :1], r.ftab[:, :1], w.dt, r.horizon, C512[:, :r.k], r.every)
        e_rom.append(rw.rel_l2(r.seismograms[0], ref["seismogram"][0]))
        ds, ps = drawn[r.rid]
        fom = rw.fom_seismograms(app, [(ds[0], 0.0, ps[0], workloads.DELAY_S)], w.dt, r.horizon,
                                 workloads.RECEIVERS, r.every)
        e_fom.append(rw.rel_l2(r.seismograms[0], fom[0]))
    rep.update({"workload": f"config4: Poisson 200 req/s for {args.serve_seconds:g} s ({len(reqs)} requests), "
                            "M~U{1..256}, K~U{128,256,512}, horizon~U[5000,20000], BF16x3, 180 receivers every 10 steps",
                "sampled_requests": [(r.rid, r.k, r.m, r.horizon) for r in sample],
                "err_vs_fp64_rom": e_rom, "err_vs_fp64_fom": e_fom, "setup_s": setup})
    return rep


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
    ap.add_argument("--copy-bytes", type=int, default=4 << 30)
    ap.add_argument("--no-rom", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--rom-steps", type=int, default=20000)
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (this build runs the hot path on a single MI355X; "
              f"--gpus {args.gpus} / WORLD_SIZE {world} is not supported)", file=sys.stderr)
        return 2

    import shaw_amd
    from shaw_amd import workloads

    w = workloads.config2()
    material = workloads.layered_material()
    n_total = args.warmup + args.steps
    log(f"workload {w.describe()}; warmup {args.warmup}, timed {args.steps}")
    t0 = time.time()
    app = shaw_amd.App.structured(w.nr, w.nth, material)
    gids = [app.map_source(d, a) for d, a, p, dl in w.sources]
    tab = shaw_amd.forcing_table(["ricker"] * w.m, [s[3] for s in w.sources], [s[2] for s in w.sources],
                                 w.dt, n_total)
    kern = shaw_amd.KERNEL_FUSED if args.kernel == "fused" else shaw_amd.KERNEL_SPLIT
    fom = shaw_amd.Fom(app, w.m, w.dt, kernel=kern, rank1_forcing=False, cols_per_thread=args.cpt,
                       strip=args.strip, segment=args.segment)
    fom.set_sources(gids)
    fom.set_forcing(tab)
    fom.set_receivers([app.map_receiver(a) for a in workloads.RECEIVERS], w.seis_every,
                      n_total // w.seis_every)
    log(f"setup {time.time() - t0:.1f} s; config {fom.config()}")
    fom.advance(args.warmup)
    fom.sync()
    total_ms, ms_per_step = fom.time(args.steps)
    fom.sync()
    seis = fom.seismogram()
    assert np.isfinite(seis).all(), "seismogram is not finite"
    kernel_bytes, model_bytes = fom.bytes_per_step()
    value = w.m * args.steps / (total_ms / 1000.0)
    achieved = kernel_bytes / (ms_per_step / 1000.0) / 1e9
    log(f"{total_ms:.1f} ms for {args.steps} steps -> {ms_per_step:.4f} ms/step, {value:.0f} forcing-steps/s, "
        f"{achieved:.0f} GB/s algorithmic ({kernel_bytes / 1e9:.3f} GB/step)")

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
                   "kernel": fom.config(), "fom_variant": "plain FP64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused<double,4>" if args.kernel == "fused" else "k_velocity+k_stress",
                     "bytes_per_launch": kernel_bytes,
                     "reference_model_bytes_per_step": model_bytes},
    }
    fom.close()
    if args.copy_bytes:
        ms = shaw_amd.debug_copy(args.copy_bytes, 10)
        out["copy_kernel"] = {"bytes": args.copy_bytes, "ms": ms,
                              "GB_s": 2 * args.copy_bytes / (ms / 1e3) / 1e9,
                              "note": "plain read+write streaming copy (k_copy), achievable-bandwidth reference"}
        log(f"copy kernel: {out['copy_kernel']['GB_s']:.0f} GB/s")
    if not args.no_rom:
        out["rom"] = rom_section(app, w, material, value, args)
    if not args.no_serve:
        out["serving"] = serving_section(app, w, args)
    if not args.no_cpu:
        log(f"cpu baseline: oracle, {args.cpu_steps} steps")
        out["cpu_baseline"] = cpu_baseline(w, material, args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
