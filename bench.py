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

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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
    assert np.isfinite(seis).all() and np.abs(seis).max() > 0, "seismogram is empty or not finite"
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
    if not args.no_cpu:
        log(f"cpu baseline: oracle, {args.cpu_steps} steps")
        out["cpu_baseline"] = cpu_baseline(w, material, args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
