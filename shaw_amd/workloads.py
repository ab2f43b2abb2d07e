"""Synthetic workloads of the SHAW time-stepping benchmark.

Every quantity the workload descriptions leave open is fixed here, from
seeds, so tests, the benchmark and the drop-in YAML inputs agree:

* mesh: mantle annulus r in [3480, 6371] km, theta in [0, 180] deg;
* material "layered": 6 interior radius breaks drawn uniformly in
  (3480, 6371) km, density and shear speed continuous and piecewise linear in
  r, from (5600 kg/m^3, 7300 m/s) at the CMB to (3300, 4400) at the surface,
  with monotone break values drawn from the same seed;
* Ricker sources at theta = 0 (mapped to theta = dth). The m realizations of
  a rank-2 workload are the reference's Cartesian product depth x period
  (fom_problem_rank_two.hpp:141-162) of seeded depths ~ U(0, 700] km and
  periods ~ U[20, 70] s with one delay of 105 s (1.5 x the longest period),
  so the same set is expressible in the reference's own input format;
* 180 surface receivers at 0, 1, ..., 179 degrees.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MATERIAL_SEED = 20240613
R_CMB_KM, R_SURF_KM = 3480.0, 6371.0


def layered_material(seed: int = MATERIAL_SEED) -> dict:
    rng = np.random.default_rng(seed)
    breaks = np.sort(rng.uniform(R_CMB_KM, R_SURF_KM, 6))
    radius = [R_CMB_KM] + [float(x) for x in breaks] + [R_SURF_KM]
    u_rho = np.sort(rng.uniform(0.0, 1.0, 6))
    u_vs = np.sort(rng.uniform(0.0, 1.0, 6))
    rho = [5600.0] + [float(5600.0 - 2300.0 * u) for u in u_rho] + [3300.0]
    vs = [7300.0] + [float(7300.0 - 2900.0 * u) for u in u_vs] + [4400.0]
    return {"kind": "layered", "radius": radius, "density": rho, "velocity": vs}


DELAY_S = 105.0


def seeded_depths_periods(nd: int, npd: int, seed: int, depth=(0.0, 700.0), period=(20.0, 70.0)):
    rng = np.random.default_rng(seed)
    depths = []
    while len(depths) < nd:
        d = float(rng.uniform(*depth))
        if d > 0.0:  # the reference requires depth > 0 (parser_forcing_section.hpp:221-225)
            depths.append(d)
    periods = [float(x) for x in rng.uniform(*period, npd)]
    return depths, periods


def seeded_sources(nd: int, npd: int, seed: int):
    """[(depth_km, angle_deg, period_s, delay_s)] in the reference's nested
    order: depth outer, period inner."""
    depths, periods = seeded_depths_periods(nd, npd, seed)
    return [(d, 0.0, p, DELAY_S) for d in depths for p in periods]


RECEIVERS = [float(a) for a in range(180)]


@dataclass
class FomWorkload:
    name: str
    nr: int
    nth: int
    m: int
    dt: float
    nsteps: int
    seis_every: int
    sources: list = field(default_factory=list)

    def describe(self) -> str:
        return (f"{self.name}: FOM FP64 {'rank-2' if self.m > 1 else 'rank-1'} {self.nr}x{self.nth} "
                f"M={self.m} dt={self.dt} {self.nsteps} steps")


def config1() -> FomWorkload:
    return FomWorkload("config1", 200, 1000, 1, 0.25, 4000, 1, [(640.0, 0.0, 45.0, 67.5)])


def config2(seed: int = 2) -> FomWorkload:
    return FomWorkload("config2", 1024, 4096, 32, 0.05, 20000, 10, seeded_sources(4, 8, seed))


def config5(seed: int = 5) -> FomWorkload:
    return FomWorkload("config5", 2048, 8192, 64, 0.025, 40000, 10, seeded_sources(8, 8, seed))


def input_yaml(w: FomWorkload, mesh_dir: str, binary: bool = True) -> str:
    """A drop-in shawExe input file for the workload (rank-2 when m > 1)."""
    mat = layered_material()
    fmt = lambda xs: "[" + ", ".join(repr(float(x)) for x in xs) + "]"
    depths = sorted(set(s[0] for s in w.sources), key=[s[0] for s in w.sources].index)
    periods = sorted(set(s[2] for s in w.sources), key=[s[2] for s in w.sources].index)
    assert len(depths) * len(periods) == w.m
    lines = [
        "general:",
        f"  meshDir: {mesh_dir}",
        f"  dt: {w.dt!r}",
        f"  finalTime: {w.dt * w.nsteps!r}",
        "  checkNumericalDispersion: false",
        "  checkCfl: true",
        "io:",
        "  seismogram:",
        f"    binary: {'true' if binary else 'false'}",
        f"    freq: {w.seis_every}",
        f"    receivers: {fmt(RECEIVERS)}",
        "source:",
        "  signal:",
        "    kind: ricker",
        f"    depth: {fmt(depths)}",
        "    angle: 0.0",
        *([f"    forcingSize: {w.m}"] if w.m > 1 else []),
        f"    period: {fmt(periods)}",
        f"    delay: {w.sources[0][3]!r}",
        "material:",
        "  kind: layered",
        f"  radius: {fmt(mat['radius'])}",
        f"  density: {fmt(mat['density'])}",
        f"  velocity: {fmt(mat['velocity'])}",
    ]
    return "\n".join(lines) + "\n"
