"""Galerkin ROM workflow on one GPU, from FOM snapshots to batched ROM runs.

The reference's offline chain is: shawExe (FOM with snapshot matrices) ->
computeThinSVD (POD, src/tools/main_eigen_svd.cc) -> shawExe with a `rom`
section (rom_load_basis.hpp, rom_compute_jacobians.hpp, rom_run_rank_one.hpp)
-> reconstructSeismogram (src/tools/main_reconstruct_seismo.cc). Here the same
chain stays on the device: the training FOM keeps its snapshots in HBM, POD
runs on them in FP64, the projection runs on the GPU in FP64, and the batched
reduced leapfrog runs on MFMA with receivers fused (y = Phi_v[recv, :] q_v).
"""
from __future__ import annotations

import time

import numpy as np

import shaw_amd


class RomBasis:
    """POD bases of both dofs, resident on the device (`dev`, a
    shaw_amd.Basis), with their singular values; phi_v / phi_s download the
    full bases on first use (only for callers that need host arrays)."""

    def __init__(self, dev: shaw_amd.Basis, sigma_v, sigma_s, seconds: dict, mv=None, ms=None):
        self.dev, self.sigma_v, self.sigma_s, self.seconds = dev, sigma_v, sigma_s, seconds
        self.mv, self.ms = mv, ms  # energy weights when the bases are energy-orthonormal
        self._host = {}

    @property
    def k(self) -> int:
        return self.dev.k["vp"]

    @property
    def phi_v(self) -> np.ndarray:
        if "vp" not in self._host:
            self._host["vp"] = self.dev.download("vp")
        return self._host["vp"]

    @property
    def phi_s(self) -> np.ndarray:
        if "sp" not in self._host:
            self._host["sp"] = self.dev.download("sp")
        return self._host["sp"]


def train_basis(app: shaw_amd.App, sources, dt: float, nsteps: int, snap_every: int, k: int,
                rank1_forcing: bool = False, energy: bool = False) -> RomBasis:
    """Run the FOM for the training forcings with snapshots of both dofs every
    `snap_every` steps (all in HBM), then POD each dof to k modes (FP64),
    Euclidean (the reference's) or, with `energy`, in the FOM energy weights.
    The bases stay on the device."""
    t0 = time.time()
    m = len(sources)
    fom = shaw_amd.Fom(app, m, dt, rank1_forcing=rank1_forcing)
    fom.set_sources([app.map_source(d, a) for d, a, p, dl in sources])
    fom.set_forcing(shaw_amd.forcing_table(["ricker"] * m, [s[3] for s in sources], [s[2] for s in sources],
                                           dt, nsteps))
    T = nsteps // snap_every
    cols = list(range(m))
    fom.set_snapshots("vp", snap_every, cols, T)
    fom.set_snapshots("sp", snap_every, cols, T)
    fom.advance(nsteps)
    fom.sync()
    t1 = time.time()
    mv, ms = energy_weights(app) if energy else (None, None)
    dev = fom.pod_basis(k, k, mv, ms)
    fom.close()
    t2 = time.time()
    return RomBasis(dev, dev.sigma("vp"), dev.sigma("sp"), {"fom_training": t1 - t0, "pod": t2 - t1}, mv, ms)


def operators(app: shaw_amd.App, basis, k: int):
    """Galerkin operators of the first k modes (POD bases are nested):
    the reference's Euclidean pair, or the energy-consistent pair for
    energy-orthonormal bases. `basis`: a RomBasis (projected on the device)."""
    return basis.dev.project(app, k, k, basis.mv)


def basis_forcing_inputs(app: shaw_amd.App, basis, k: int, sources, dt: float, nsteps: int):
    """Forcing inputs of the first k modes for either projection."""
    if basis.mv is not None:
        return energy_forcing_inputs(app, basis, basis.mv, sources, dt, nsteps, k)
    return forcing_inputs(app, basis, sources, dt, nsteps, k)


def _rows_at(phi_v, gids, k=None) -> np.ndarray:
    """[m x k] rows of Phi_v at gids from a RomBasis (gathered on the device)
    or a host array."""
    if isinstance(phi_v, RomBasis):
        return phi_v.dev.rows("vp", gids, k or phi_v.k)
    rows = np.stack([phi_v[g, :] for g in gids], axis=0)
    return rows if k is None else rows[:, :k]


def forcing_inputs(app: shaw_amd.App, phi_v, sources, dt: float, nsteps: int, k=None):
    """B [K x m] with column j = rho^-1[g_j] Phi_v[g_j, :] (rom_problem_rank_one.hpp:133-147)
    and the forcing table [nsteps x m]. phi_v: [nVp x K] or a RomBasis."""
    _, _, rho_inv, _ = app.operators()
    gids = [app.map_source(d, a) for d, a, p, dl in sources]
    rows = _rows_at(phi_v, gids, k)
    B = np.stack([rho_inv[g] * rows[j] for j, g in enumerate(gids)], axis=1)
    ftab = shaw_amd.forcing_table(["ricker"] * len(sources), [s[3] for s in sources], [s[2] for s in sources],
                                  dt, nsteps)
    return B, ftab


def receiver_rows(app: shaw_amd.App, phi_v, receivers, k=None):
    """C [R x K] = Phi_v[recv, :K] (main_reconstruct_seismo.cc:13-41); phi_v:
    [nVp x K] or a RomBasis."""
    return _rows_at(phi_v, [app.map_receiver(a) for a in receivers], k)


def fom_seismograms(app: shaw_amd.App, sources, dt: float, nsteps: int, receivers, every: int,
                    rank1_forcing: bool = False):
    """FP64 FOM seismograms [m][T][R] of the given forcings (the reference FOM)."""
    out = shaw_amd.fom_run(app, dt, nsteps, sources, True, receivers, every, rank1_forcing=rank1_forcing) \
        if len(sources) > 1 else shaw_amd.fom_run(app, dt, nsteps, sources, False, receivers, every)
    return out["seismogram"]


def amplification_radius(ar_v: np.ndarray, ar_s: np.ndarray, dt: float) -> float:
    """Spectral radius of the reduced leapfrog's one-step matrix. A step
    (rom_run_rank_one.hpp:112-151) maps (q_v, q_s) to
    q_v' = q_v + dt Ar_v q_s and q_s' = q_s + dt Ar_s q_v', i.e. by
    [[I, dt Ar_v], [dt Ar_s, I + dt^2 Ar_s Ar_v]]; a radius above 1 means
    homogeneous solutions grow like radius^steps (instability), at most 1
    means they stay bounded."""
    kv, ks = ar_v.shape
    A = np.zeros((kv + ks, kv + ks))
    A[:kv, :kv] = np.eye(kv)
    A[:kv, kv:] = dt * ar_v
    A[kv:, :kv] = dt * ar_s
    A[kv:, kv:] = np.eye(ks) + dt * dt * (ar_s @ ar_v)
    return float(np.abs(np.linalg.eigvals(A)).max())


def rel_l2(x, ref) -> float:
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / n) if n > 0 else float(np.linalg.norm(x))


# ---------------------------------------------------------------- energy-consistent ROM
# The FOM conserves a discrete energy E = 1/2 v^T Mv v + 1/2 s^T Ms s with
# diagonal Mv, Ms: on every coupling that is not merged at the CMB/surface
# rows or on the axis, Mv[p] J_vp[p, q] = -Ms[q] J_sp[q, p]
# (shwavepp.hpp:319-466 gives both coefficients). With bases orthonormal in
# these inner products the Galerkin pair satisfies Ar_s = -Ar_v^T up to the
# boundary couplings; enforcing Ar_s = -Ar_v^T gives a reduced leapfrog that
# conserves a reduced energy, so rounding of the operators (BF16x3) cannot
# excite growing modes. This projection is an addition: the reference's
# Galerkin ROM (rom_compute_jacobians.hpp:58-109) is the Euclidean one.

def energy_weights(app: shaw_amd.App):
    """Diagonal energy weights (Mv [nVp], Ms [nSp]) of the FOM operators
    (shaw_app_energy_weights, host C++)."""
    return app.energy_weights()


def energy_forcing_inputs(app: shaw_amd.App, phi_v, mv: np.ndarray, sources, dt: float, nsteps: int, k=None):
    """B [K x m], column j = Mv[g_j] rho^-1[g_j] Phi_v[g_j, :] (q_v = Phi_v^T Mv v), and the forcing table."""
    _, _, rho_inv, _ = app.operators()
    gids = [app.map_source(d, a) for d, a, p, dl in sources]
    rows = _rows_at(phi_v, gids, k)
    B = np.stack([mv[g] * rho_inv[g] * rows[j] for j, g in enumerate(gids)], axis=1)
    ftab = shaw_amd.forcing_table(["ricker"] * len(sources), [s[3] for s in sources], [s[2] for s in sources],
                                  dt, nsteps)
    return B, ftab
