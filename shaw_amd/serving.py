"""Continuous-batching ROM serving on one GPU: Python glue over shaw_serve.

The request loop itself (admission by free slots and HBM footprint in FIFO or
least-work-first order with a wait bound, chunked advances of every engine
launched before any wait, batched retirement, latency accounting) is C++
behind the C ABI (shaw_amd/csrc/serve.cpp, shaw_serve in
include/shaw_amd.h). This module builds its inputs: request objects, the
engine tables (a resident shaw_rom engine on the GPU, or the numpy engine
below through ctypes callbacks, so the loop is tested on the CPU), the
Poisson arrivals of the benchmark, and the report.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

import shaw_amd

FIFO, LEAST_WORK = 0, 1

_ADMIT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_double),
                     C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(C.c_uint64))
_ADVANCE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64)
_SYNC = C.CFUNCTYPE(C.c_int, C.c_void_p)
_RETIRE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(C.c_double),
                      C.POINTER(C.c_uint64))
_FREE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint64))
_ADMIT_SIG = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_int),
                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_uint64), C.c_uint64,
                         C.POINTER(C.c_uint64))
_QUERY = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int))
_STEPS = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
_CLOCK = C.CFUNCTYPE(C.c_double, C.c_void_p)
_WAIT = C.CFUNCTYPE(None, C.c_void_p, C.c_double)


class ShawEngine(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("kv", C.c_uint64), ("ks", C.c_uint64), ("r", C.c_uint64),
                ("capacity", C.c_uint64), ("admit", _ADMIT), ("advance", _ADVANCE), ("sync", _SYNC),
                ("retire_batch", _RETIRE), ("free_slots", _FREE), ("query", _QUERY), ("admit_signals", _ADMIT_SIG),
                ("steps", _STEPS)]


class ShawRequest(C.Structure):
    _fields_ = [("id", C.c_uint64), ("engine", C.c_int), ("m", C.c_uint64), ("horizon", C.c_uint64),
                ("every", C.c_uint64), ("arrival_s", C.c_double), ("b", C.POINTER(C.c_double)),
                ("ftab", C.POINTER(C.c_double)), ("seis", C.POINTER(C.c_double)), ("admitted_s", C.c_double),
                ("done_s", C.c_double), ("kinds", C.POINTER(C.c_int)), ("delays", C.POINTER(C.c_double)),
                ("periods", C.POINTER(C.c_double))]


class ShawServeOptions(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("chunk_steps", C.c_uint64), ("policy", C.c_int),
                ("max_wait_s", C.c_double), ("wall_limit_s", C.c_double), ("clock", _CLOCK),
                ("wait_until", _WAIT), ("clock_ctx", C.c_void_p)]


class ShawServeStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("forcing_steps", C.c_uint64), ("launches", C.c_uint64),
                ("max_queue", C.c_uint64), ("span_s", C.c_double), ("p50_latency_s", C.c_double),
                ("p99_latency_s", C.c_double), ("max_latency_s", C.c_double), ("max_wait_s", C.c_double),
                ("throughput", C.c_double)]


def _bind():
    L = shaw_amd.lib()
    if not getattr(L, "_serve_bound", False):
        L.shaw_engine_from_rom.restype = C.c_int
        L.shaw_engine_from_rom.argtypes = [C.c_void_p, C.POINTER(ShawEngine)]
        L.shaw_serve.restype = C.c_int
        L.shaw_serve.argtypes = [C.POINTER(ShawEngine), C.c_int, C.POINTER(ShawServeOptions),
                                 C.POINTER(ShawRequest), C.c_uint64, C.POINTER(ShawServeStats)]
        L._serve_bound = True
    return L


@dataclass
class Request:
    rid: int
    k: int
    b: np.ndarray            # [K x m]
    ftab: np.ndarray         # [horizon x m], or None with `signals`
    horizon: int
    every: int
    arrival: float           # seconds after the start of the run
    admitted: float = -1.0
    done: float = -1.0
    seismograms: list = field(default_factory=list)  # per column [T x R]
    signals: tuple = None    # (kinds, delays, periods) per column: the engine evaluates the forcing

    @property
    def m(self) -> int:
        return self.b.shape[1]

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def column_bytes(kv: int, ks: int, horizon: int, every: int, r: int) -> int:
    """HBM bytes one resident column holds: states, forcing projection,
    forcing series and seismogram buffer (FP64); the footprint the loop
    admits by."""
    samples = horizon // every if every else 0
    return 8 * (kv + ks + kv + horizon + r * samples)


class Server:
    """One serving run over resident engines (K -> engine, or K -> list of
    engines of that size, which take the requests of size K in turn)."""

    def __init__(self, engines: dict, hbm_budget: float, chunk: int = 500, policy: str = "fifo",
                 max_wait: float = 0.0, clock=None):
        """policy: "fifo" or "least_work" ("sjf"); max_wait [s]: least-work
        admission serves a request that waited this long before any other
        (0: no bound). clock: None (steady clock) or an object with
        now() -> s and wait_until(t) for a simulated clock."""
        pol = {"fifo": FIFO, "least_work": LEAST_WORK, "sjf": LEAST_WORK}
        if policy not in pol:
            raise ValueError(f"unknown admission policy {policy!r}")
        self.engines, self.budget, self.chunk = engines, hbm_budget, chunk
        self.policy, self.max_wait, self.clock = pol[policy], max_wait, clock
        self.stats = None

    def run(self, requests: list, wall_limit: float = 600.0, keep=None) -> dict:
        """Serve the requests; keep(r) -> bool selects the requests whose
        seismograms are returned (default: all)."""
        L = _bind()
        flat, index = [], {}  # engine tables; K -> their indices
        for k in sorted(self.engines):
            group = self.engines[k] if isinstance(self.engines[k], (list, tuple)) else [self.engines[k]]
            index[k] = list(range(len(flat), len(flat) + len(group)))
            flat += group
        for r in requests:
            if r.k not in index:
                raise ValueError(f"no resident ROM of size {r.k}")
        tables = (ShawEngine * len(flat))(*[e.table() for e in flat])
        arr = (ShawRequest * max(1, len(requests)))()
        alive = []  # arrays and callbacks the C loop points into
        outs = []
        for i, r in enumerate(requests):
            ei = index[r.k][r.rid % len(index[r.k])]
            e = flat[ei]
            b = np.ascontiguousarray(r.b, dtype=np.float64)
            T = r.horizon // r.every if r.every else 0
            seis = np.zeros((r.m, T, e.R)) if (T and e.R and (keep is None or keep(r))) else None
            outs.append(seis)
            q = arr[i]
            q.id, q.engine, q.m, q.horizon, q.every, q.arrival_s = r.rid, ei, r.m, r.horizon, r.every, r.arrival
            q.b = b.ctypes.data_as(C.POINTER(C.c_double))
            if r.ftab is not None:
                f = np.ascontiguousarray(np.asarray(r.ftab)[:r.horizon], dtype=np.float64)
                q.ftab = f.ctypes.data_as(C.POINTER(C.c_double))
                alive += [b, f]
            else:
                kinds, delays, periods = r.signals
                ka = np.ascontiguousarray([shaw_amd.SIGNAL_KINDS[x] if isinstance(x, str) else int(x) for x in kinds],
                                          dtype=np.int32)
                da = np.ascontiguousarray(delays, dtype=np.float64)
                pa = np.ascontiguousarray(periods, dtype=np.float64)
                q.kinds = ka.ctypes.data_as(C.POINTER(C.c_int))
                q.delays = da.ctypes.data_as(C.POINTER(C.c_double))
                q.periods = pa.ctypes.data_as(C.POINTER(C.c_double))
                alive += [b, ka, da, pa]
            q.seis = seis.ctypes.data_as(C.POINTER(C.c_double)) if seis is not None else None
        opts = ShawServeOptions()
        opts.hbm_budget, opts.chunk_steps, opts.policy = int(self.budget), self.chunk, self.policy
        opts.max_wait_s, opts.wall_limit_s = self.max_wait, wall_limit
        if self.clock is not None:
            cb_clock = _CLOCK(lambda ctx: self.clock.now())
            cb_wait = _WAIT(lambda ctx, t: self.clock.wait_until(t))
            opts.clock, opts.wait_until = cb_clock, cb_wait
            alive += [cb_clock, cb_wait]
        st = ShawServeStats()
        shaw_amd._check(L.shaw_serve(tables, len(flat), C.byref(opts), arr, len(requests), C.byref(st)))
        for i, r in enumerate(requests):
            r.admitted, r.done = arr[i].admitted_s, arr[i].done_s
            r.seismograms = list(outs[i]) if outs[i] is not None else []
        self.stats = st
        return {"requests": int(st.requests), "p50_latency_s": st.p50_latency_s, "p99_latency_s": st.p99_latency_s,
                "max_latency_s": st.max_latency_s, "max_wait_s": st.max_wait_s,
                "forcing_steps": int(st.forcing_steps), "throughput_forcing_steps_per_s": st.throughput,
                "wall_s": st.span_s, "launches": int(st.launches), "max_queue": int(st.max_queue)}


def poisson_requests(rng, rate: float, duration: float, k_choices, m_range, horizon_range, make_inputs,
                     every: int = 10) -> list:
    """Requests with Poisson arrivals; make_inputs(k, m, horizon, rng) -> (b,
    ftab) or (b, None, (kinds, delays, periods)) for engine-evaluated forcing."""
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        k = int(rng.choice(k_choices))
        m = int(rng.integers(m_range[0], m_range[1] + 1))
        h = int(rng.integers(horizon_range[0], horizon_range[1] + 1))
        h -= h % every
        inp = make_inputs(k, m, h, rng)
        out.append(Request(rid, k, inp[0], inp[1], h, every, t, signals=inp[2] if len(inp) > 2 else None))
        rid += 1


class GpuEngine:
    """A resident shaw_amd.Rom served through its own entry points."""

    def __init__(self, rom, capacity: int):
        self.rom = rom
        self.kv, self.ks, self.R = rom.kv, rom.ks, rom.R
        self.capacity = capacity

    def table(self) -> ShawEngine:
        t = ShawEngine()
        shaw_amd._check(_bind().shaw_engine_from_rom(self.rom.handle, C.byref(t)))
        return t


class NumpyEngine:
    """CPU engine with the engine contract (FP64 reduced leapfrog in numpy,
    the step of rom_run_rank_one.hpp:112-151 per column), for testing the
    serving loop without a GPU. `tick(steps)` is called on every advance so a
    simulated clock can charge time for the work. With stretch > 1 it acts
    like a GPU engine under a time budget: advance(n) takes a column whose
    forcing is over n * stretch steps (one-steps, so its results do not
    change), and the table reports the step counts (the `steps` call)."""

    def __init__(self, ar_v, ar_s, dt, C_, capacity, tick=None, stretch=1):
        self.ar_v, self.ar_s, self.dt, self.C = ar_v, ar_s, dt, C_
        self.stretch = stretch
        self.kv, self.ks = ar_v.shape
        self.R = 0 if C_ is None else C_.shape[0]
        self.capacity = capacity
        self.cols = {}
        self.tick = tick
        self.calls = []

    # ---- the engine contract, in numpy
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
        if self.tick:
            self.tick(n)
        for c in self.cols.values():
            quiet = self.stretch > 1 and not np.any(c["f"][c["step"]:c["h"]])
            for _ in range(n * self.stretch if quiet else n):
                if c["step"] >= c["h"]:
                    break
                fdt = c["f"][c["step"]] * self.dt
                c["qv"] = c["qv"] + self.dt * (self.ar_v @ c["qs"])
                c["qv"] = c["qv"] + fdt * c["b"]
                c["step"] += 1
                if self.R and c["every"] and c["step"] % c["every"] == 0:
                    c["seis"].append(self.C @ c["qv"])
                c["qs"] = c["qs"] + self.dt * (self.ar_s @ c["qv"])

    def retire(self, slot):
        c = self.cols.pop(slot)
        return np.array(c["seis"]).reshape(-1, self.R) if self.R else np.zeros((0, 0))

    # ---- the same contract as a C table of ctypes callbacks
    def table(self) -> ShawEngine:
        def guard(fn):
            def call(*a):
                try:
                    return fn(*a)
                except Exception as exc:  # reported as an engine failure
                    self.calls.append(("error", repr(exc)))
                    return 1
            return call

        def admit(ctx, n, b, nsteps, ftab, horizons, every, slots):
            bb = np.ctypeslib.as_array(b, (self.kv, n)).copy()
            ff = np.ctypeslib.as_array(ftab, (nsteps, n)).copy()
            hz = [int(horizons[j]) for j in range(n)]
            for j, s in enumerate(self.admit(bb, ff, hz, int(every))):
                slots[j] = s
            self.calls.append(("admit", int(n)))
            return 0

        def admit_signals(ctx, n, b, kinds, delays, periods, nsteps, horizons, every, slots):
            ks = [int(kinds[j]) for j in range(n)]
            ff = shaw_amd.forcing_table(ks, [delays[j] for j in range(n)], [periods[j] for j in range(n)], self.dt,
                                        int(nsteps))
            bb = np.ctypeslib.as_array(b, (self.kv, n)).copy()
            hz = [int(horizons[j]) for j in range(n)]
            for j, s in enumerate(self.admit(bb, ff, hz, int(every))):
                slots[j] = s
            self.calls.append(("admit_signals", int(n)))
            return 0

        def advance(ctx, n):
            self.advance(int(n))
            self.calls.append(("advance", int(n)))
            return 0

        def retire(ctx, n, slots, max_samples, seis, samples):
            width = self.R * int(max_samples)
            for j in range(n):
                y = self.retire(int(slots[j]))
                if width and seis:  # NULL: results not wanted
                    out = np.ctypeslib.as_array(seis, (int(n) * width,))
                    out[j * width:j * width + y.size] = y.ravel()
                if samples:
                    samples[j] = len(y)
            return 0

        def free(ctx, out):
            out[0] = self.free_slots()
            return 0

        def query(ctx, idle):
            idle[0] = 1  # advance() is synchronous
            return 0

        def steps(ctx, n, slots, out):
            for j in range(n):
                out[j] = self.cols[int(slots[j])]["step"]
            return 0

        t = ShawEngine()
        t.kv, t.ks, t.r, t.capacity = self.kv, self.ks, self.R, self.capacity
        t.admit, t.advance, t.sync = _ADMIT(guard(admit)), _ADVANCE(guard(advance)), _SYNC(lambda ctx: 0)
        t.retire_batch, t.free_slots, t.query = _RETIRE(guard(retire)), _FREE(guard(free)), _QUERY(guard(query))
        t.admit_signals = _ADMIT_SIG(guard(admit_signals))
        if self.stretch > 1:
            t.steps = _STEPS(guard(steps))
        self._keep = (t.admit, t.advance, t.sync, t.retire_batch, t.free_slots, t.query, t.admit_signals, t.steps)
        return t


class SimClock:
    """A simulated clock for CPU tests: every reading costs `per_call`
    seconds, engine work costs `per_step` seconds a step, and idle waits jump
    to the next arrival."""

    def __init__(self, per_call=1e-5, per_step=1e-4):
        self.t, self.per_call, self.per_step = 0.0, per_call, per_step

    def now(self) -> float:
        self.t += self.per_call
        return self.t

    def wait_until(self, t: float) -> None:
        self.t = max(self.t, t)

    def work(self, steps: int) -> None:
        self.t += self.per_step * steps
