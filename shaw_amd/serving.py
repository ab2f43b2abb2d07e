"""Continuous-batching ROM server on one GPU.

Requests carry m forcing realizations (columns), a ROM size K, a horizon and
a seismogram sampling period. Each K has one resident engine (reduced
operators, receiver rows, column slots with their states, forcing series and
seismogram buffers in HBM: shaw_rom_*). The loop

  1. moves arrived requests to a queue,
  2. admits queued requests while their K's engine has free slots and the
     resident footprint stays within the HBM budget (requests admitted while
     others are mid-flight join the running batch: continuous batching).
     Policy "fifo" (the default) admits in arrival order and does not skip a
     request that does not fit; policy "sjf" admits, among the queued
     requests that fit, the one with the least work (horizon x columns)
     first, which lowers the median latency under load at the cost of the
     largest requests,
  3. advances every engine that has active columns by one chunk of steps --
     each engine on its own HIP stream, all launched before any wait, so the
     engines' work overlaps on the device -- and
  4. retires requests whose columns all reached their horizon, recording
     latency = completion time - arrival time.

A column's arithmetic does not depend on the other columns in the batch, so
a request's result is the same alone and in the mix (tested bit for bit).
The engine interface is small (admit / advance / sync / retire / free_slots),
so the scheduler is tested on the CPU with a numpy engine.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    k: int
    b: np.ndarray            # [K x m]
    ftab: np.ndarray         # [horizon x m]
    horizon: int
    every: int
    arrival: float           # seconds, on the server clock
    admitted: float = -1.0
    done: float = -1.0
    slots: list = field(default_factory=list)
    seismograms: list = field(default_factory=list)  # per column [T x R]

    @property
    def m(self) -> int:
        return self.b.shape[1]

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def column_bytes(kv: int, ks: int, horizon: int, every: int, r: int) -> int:
    """HBM bytes one resident column holds: states, forcing projection,
    forcing series and seismogram buffer (FP64)."""
    samples = horizon // every if every else 0
    return 8 * (kv + ks + kv + horizon + r * samples)


class Server:
    def __init__(self, engines: dict, hbm_budget: float, chunk: int = 500, clock=time.perf_counter,
                 policy: str = "fifo"):
        """engines: K -> engine with .admit(b, ftab, horizons, every) -> slots,
        .advance(n), .sync(), .retire_many(slots, max_samples) -> [seis],
        .free_slots(), .kv, .ks, .R. policy: "fifo" or "sjf"."""
        if policy not in ("fifo", "sjf"):
            raise ValueError(f"unknown admission policy {policy!r}")
        self.policy = policy
        self.engines = engines
        self.budget = hbm_budget
        self.chunk = chunk
        self.clock = clock
        self.t0 = clock()
        self.queue: collections.deque = collections.deque()
        self.resident: list = []
        self.finished: list = []
        self.resident_bytes = 0
        self.forcing_steps = 0
        self.launches = 0

    def now(self) -> float:
        return self.clock() - self.t0

    def footprint(self, req: Request) -> int:
        e = self.engines[req.k]
        return req.m * column_bytes(e.kv, e.ks, req.horizon, req.every, e.R)

    def submit(self, req: Request) -> None:
        if req.k not in self.engines:
            raise ValueError(f"no resident ROM of size {req.k}")
        e = self.engines[req.k]
        if req.m > e.capacity:
            raise ValueError(f"request {req.rid}: {req.m} columns exceed the engine capacity {e.capacity}")
        if self.footprint(req) > self.budget:
            raise ValueError(f"request {req.rid}: footprint exceeds the HBM budget")
        self.queue.append(req)

    def _fits(self, req: Request) -> bool:
        return (self.engines[req.k].free_slots() >= req.m and
                self.resident_bytes + self.footprint(req) <= self.budget)

    def _admit(self) -> None:
        while self.queue:
            if self.policy == "fifo":
                req = self.queue[0]
                if not self._fits(req):
                    break  # FIFO: wait for slots / memory to free up
            else:  # least work first among the requests that fit (ties: arrival order)
                fit = [r for r in self.queue if self._fits(r)]
                if not fit:
                    break
                req = min(fit, key=lambda r: r.horizon * r.m)
            self.queue.remove(req)
            e = self.engines[req.k]
            fp = self.footprint(req)
            req.slots = e.admit(req.b, req.ftab, [req.horizon] * req.m, req.every)
            req.admitted = self.now()
            req.steps_left = req.horizon
            self.resident_bytes += fp
            self.resident.append(req)

    def _advance(self) -> None:
        active = {}
        for req in self.resident:
            active[req.k] = max(active.get(req.k, 0), req.steps_left)
        for k, left in active.items():  # launch all engines before waiting on any
            self.engines[k].advance(min(self.chunk, left))
            self.launches += 1
        for k in active:
            self.engines[k].sync()
        for req in self.resident:
            n = min(self.chunk, active[req.k])
            taken = min(n, req.steps_left)
            req.steps_left -= taken
            self.forcing_steps += taken * req.m

    def _retire(self) -> None:
        keep = []
        for req in self.resident:
            if req.steps_left > 0:
                keep.append(req)
                continue
            e = self.engines[req.k]
            T = req.horizon // req.every if req.every else 0
            req.seismograms = e.retire_many(req.slots, T)
            req.done = self.now()
            self.resident_bytes -= self.footprint(req)
            self.finished.append(req)
        self.resident = keep

    def step(self) -> None:
        self._admit()
        if self.resident:
            self._advance()
            self._retire()

    def run(self, requests: list, wall_limit: float = 600.0) -> dict:
        """Serve requests arriving at their `arrival` times (server clock)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        i = 0
        self.t0 = self.clock()
        while i < len(pending) or self.queue or self.resident:
            now = self.now()
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            if not self.queue and not self.resident:
                if i < len(pending):
                    time.sleep(max(0.0, min(0.01, pending[i].arrival - now)))
                continue
            self.step()
            if now > wall_limit:
                raise RuntimeError("serving run exceeded its wall-clock limit")
        return self.report()

    def report(self) -> dict:
        lat = np.array([r.latency for r in self.finished])
        span = max(r.done for r in self.finished) - min(r.arrival for r in self.finished) if self.finished else 0.0
        return {"requests": len(self.finished),
                "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
                "forcing_steps": int(self.forcing_steps),
                "throughput_forcing_steps_per_s": self.forcing_steps / span if span > 0 else None,
                "wall_s": span, "launches": self.launches}


def poisson_requests(rng, rate: float, duration: float, k_choices, m_range, horizon_range, make_inputs,
                     every: int = 10) -> list:
    """Requests with Poisson arrivals; make_inputs(k, m, horizon, rng) -> (b, ftab)."""
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        k = int(rng.choice(k_choices))
        m = int(rng.integers(m_range[0], m_range[1] + 1))
        h = int(rng.integers(horizon_range[0], horizon_range[1] + 1))
        h -= h % every
        b, ftab = make_inputs(k, m, h, rng)
        out.append(Request(rid, k, b, ftab, h, every, t))
        rid += 1


class GpuEngine:
    """shaw_amd.Rom with the attributes the server needs."""

    def __init__(self, rom, capacity: int):
        self.rom = rom
        self.kv, self.ks, self.R = rom.kv, rom.ks, rom.R
        self.capacity = capacity

    def admit(self, b, ftab, horizons, every):
        return self.rom.admit(b, ftab, horizons, every)

    def advance(self, n):
        self.rom.advance(n)

    def sync(self):
        self.rom.sync()

    def retire(self, slot, max_samples):
        return self.rom.retire(slot, max_samples)

    def retire_many(self, slots, max_samples):
        return self.rom.retire_many(slots, max_samples)

    def free_slots(self):
        return self.rom.free_slots()


class NumpyEngine:
    """CPU engine with the same contract as GpuEngine (FP64 reduced leapfrog
    in numpy), for testing the scheduler without a GPU."""

    def __init__(self, ar_v, ar_s, dt, C, capacity):
        self.ar_v, self.ar_s, self.dt, self.C = ar_v, ar_s, dt, C
        self.kv, self.ks = ar_v.shape
        self.R = 0 if C is None else C.shape[0]
        self.capacity = capacity
        self.cols = {}

    def free_slots(self):
        return self.capacity - len(self.cols)

    def admit(self, b, ftab, horizons, every):
        slots = []
        free = [s for s in range(self.capacity) if s not in self.cols]
        for j in range(b.shape[1]):
            s = free[j]
            self.cols[s] = {"b": b[:, j].copy(), "f": ftab[:, j].copy(), "h": horizons[j], "every": every,
                            "step": 0, "qv": np.zeros(self.kv), "qs": np.zeros(self.ks), "seis": []}
            slots.append(s)
        return slots

    def advance(self, n):
        for c in self.cols.values():
            for _ in range(n):
                if c["step"] >= c["h"]:
                    break
                fdt = c["f"][c["step"]] * self.dt
                c["qv"] = c["qv"] + self.dt * (self.ar_v @ c["qs"])
                c["qv"] = c["qv"] + fdt * c["b"]
                c["step"] += 1
                if self.R and c["every"] and c["step"] % c["every"] == 0:
                    c["seis"].append(self.C @ c["qv"])
                c["qs"] = c["qs"] + self.dt * (self.ar_s @ c["qv"])

    def sync(self):
        pass

    def retire(self, slot, max_samples):
        c = self.cols.pop(slot)
        seis = np.array(c["seis"]).reshape(-1, self.R) if self.R else np.zeros((0, 0))
        return seis, c["qv"], c["qs"]

    def retire_many(self, slots, max_samples):
        return [self.retire(s, max_samples)[0] for s in slots]
