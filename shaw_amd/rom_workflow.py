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
from dataclasses import dataclass

import numpy as np

import shaw_amd


@dataclass
class RomBasis:
    phi_v: np.ndarray   # [nVp x K]
    phi_s: np.ndarray   # [nSp x K]
    sigma_v: np.ndarray
    sigma_s: np.ndarray
    seconds: dict
    mv: np.ndarray = None  # energy weights when the bases are energy-orthonormal
    ms: np.ndarray = None


def train_basis(app: shaw_amd.App, sources, dt: float, nsteps: int, snap_every: int, k: int,
                rank1_forcing: bool = False, energy: bool = False) -> RomBasis:
    """Run the FOM for the training forcings with snapshots of both dofs every
    `snap_every` steps (all in HBM), then POD each dof to k modes (FP64),
    Euclidean (the reference's) or, with `energy`, in the FOM energy weights."""
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
    phi_v, sig_v = fom.pod("vp", k, T * m, mv)
    phi_s, sig_s = fom.pod("sp", k, T * m, ms)
    fom.close()
    t2 = time.time()
    return RomBasis(phi_v, phi_s, sig_v, sig_s, {"fom_training": t1 - t0, "pod": t2 - t1}, mv, ms)


def operators(app: shaw_amd.App, basis: RomBasis, k: int):
    """Galerkin operators of the first k modes (POD bases are nested):
    the reference's Euclidean pair, or the energy-consistent pair for
    energy-orthonormal bases."""
    if basis.mv is not None:
        return shaw_amd.rom_project_energy(app, basis.phi_v[:, :k], basis.phi_s[:, :k], basis.mv)
    return shaw_amd.rom_project(app, basis.phi_v[:, :k], basis.phi_s[:, :k])


def basis_forcing_inputs(app: shaw_amd.App, basis: RomBasis, k: int, sources, dt: float, nsteps: int):
    """Forcing inputs of the first k modes for either projection."""
    if basis.mv is not None:
        return energy_forcing_inputs(app, basis.phi_v[:, :k], basis.mv, sources, dt, nsteps)
    return forcing_inputs(app, basis.phi_v[:, :k], sources, dt, nsteps)


def forcing_inputs(app: shaw_amd.App, phi_v: np.ndarray, sources, dt: float, nsteps: int):
    """B [K x m] with column j = rho^-1[g_j] Phi_v[g_j, :] (rom_problem_rank_one.hpp:133-147)
    and the forcing table [nsteps x m]."""
    _, _, rho_inv, _ = app.operators()
    gids = [app.map_source(d, a) for d, a, p, dl in sources]
    B = np.stack([rho_inv[g] * phi_v[g, :] for g in gids], axis=1)
    ftab = shaw_amd.forcing_table(["ricker"] * len(sources), [s[3] for s in sources], [s[2] for s in sources],
                                  dt, nsteps)
    return B, ftab


def receiver_rows(app: shaw_amd.App, phi_v: np.ndarray, receivers):
    """C [R x K] = Phi_v[recv, :] (main_reconstruct_seismo.cc:13-41)."""
    return np.stack([phi_v[app.map_receiver(a), :] for a in receivers], axis=0)


def fom_seismograms(app: shaw_amd.App, sources, dt: float, nsteps: int, receivers, every: int,
                    rank1_forcing: bool = False):
    """FP64 FOM seismograms [m][T][R] of the given forcings (the reference FOM)."""
    out = shaw_amd.fom_run(app, dt, nsteps, sources, True, receivers, every, rank1_forcing=rank1_forcing) \
        if len(sources) > 1 else shaw_amd.fom_run(app, dt, nsteps, sources, False, receivers, every)
    return out["seismogram"]


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
    """Diagonal energy weights (Mv [nVp], Ms [nSp]) of the FOM operators, from
    the ratios Ms[q] / Mv[p] = -J_vp[p, q] / J_sp[q, p] of the unmerged
    couplings, chained along row 1 and up every column (exact products of the
    coefficients). Points whose couplings are all merged (rows 0 and nr-1,
    the two axis columns) take the weight of the adjacent point, the
    stress points the ratio through an unmerged coupling where one exists."""
    cv, cs, _, _ = app.operators()
    nr, nth = app.nr, app.nth
    rs = 2 * nth - 1
    cN = cv[:, 0].reshape(nr, nth); cS = cv[:, 1].reshape(nr, nth)
    cW = cv[:, 2].reshape(nr, nth); cE = cv[:, 3].reshape(nr, nth)
    # stress coefficient rows: srp(i, j) = j*rs + 2i, stp(i, j) = j*rs + 2i + 1 (j <= nr-2),
    # stp(i, nr-1) = (nr-1)*rs + i
    body = cs[:(nr - 1) * rs].reshape(nr - 1, rs, 2)
    srp0, srp1 = body[:, 0::2, 0], body[:, 0::2, 1]          # [nr-1, nth]: north v, south v
    stp0 = np.vstack([body[:, 1::2, 0], cs[(nr - 1) * rs:, 0][None, :]])  # [nr, nth-1]: west v
    stp1 = np.vstack([body[:, 1::2, 1], cs[(nr - 1) * rs:, 1][None, :]])  # east v
    with np.errstate(divide="ignore", invalid="ignore"):
        # Mv[j+1]/Mv[j] through srp(i, j): (Ms/Mv[j]) / (Ms/Mv[j+1]) = (-cN[j]/srp1) / (-cS[j+1]/srp0)
        up = (cN[:-1, :] / srp1) / (cS[1:, :] / srp0)         # valid for 1 <= j <= nr-3
        # Mv[i+1]/Mv[i] through stp(i, j): (-cE[i]/stp0) / (-cW[i+1]/stp1)
        east = (cE[:, :-1] / stp0) / (cW[:, 1:] / stp1)       # valid for 1 <= i <= nth-3
    lv = np.zeros((nr, nth))
    lv[1, 2:nth - 1] = np.cumsum(np.log(east[1, 1:nth - 2]))
    lv[2:nr - 1, 1:nth - 1] = lv[1, 1:nth - 1][None, :] + np.cumsum(np.log(up[1:nr - 2, 1:nth - 1]), axis=0)
    lv -= lv[nr // 2, nth // 2]
    mv = np.exp(lv)
    with np.errstate(divide="ignore", invalid="ignore"):
        # radial stress from the unmerged side: the south point for rows j >= 1, the north one for row 0
        srp = np.where(np.arange(nr - 1)[:, None] >= 1, -mv[:-1, :] * cN[:-1, :] / srp1, -mv[1:, :] * cS[1:, :] / srp0)
        # CMB and surface rows: N and S of a point there are one stress point, so its merged
        # coefficient (in the N slot) and that stress point's coefficient fix the weight
        mv[0, :] = -srp[0, :] * srp1[0, :] / cN[0, :]
        mv[nr - 1, :] = -srp[nr - 2, :] * srp0[nr - 2, :] / cN[nr - 1, :]
        # axis points: all coefficients vanish (the field is zero there); take the neighbour's weight
        mv[:, 0] = mv[:, 1]; mv[:, nth - 1] = mv[:, nth - 2]
        srp[:, 0] = srp[:, 1]; srp[:, nth - 1] = srp[:, nth - 2]
        # angular stress from the point off the axis
        stp = np.where(np.arange(nth - 1)[None, :] >= 1, -mv[:, :-1] * cE[:, :-1] / stp0, -mv[:, 1:] * cW[:, 1:] / stp1)
    ms = np.empty(app.nSp)
    msb = np.empty((nr - 1, rs))
    msb[:, 0::2] = srp
    msb[:, 1::2] = stp[:nr - 1, :]
    ms[:(nr - 1) * rs] = msb.ravel()
    ms[(nr - 1) * rs:] = stp[nr - 1, :]
    mvf = mv.ravel()
    bad = int((~np.isfinite(ms) | (ms <= 0)).sum() + (~np.isfinite(mvf) | (mvf <= 0)).sum())
    if bad:
        raise RuntimeError(f"energy weights: {bad} points without a usable coupling")
    return mvf, ms


def energy_forcing_inputs(app: shaw_amd.App, phi_v: np.ndarray, mv: np.ndarray, sources, dt: float, nsteps: int):
    """B [K x m], column j = Mv[g_j] rho^-1[g_j] Phi_v[g_j, :] (q_v = Phi_v^T Mv v), and the forcing table."""
    _, _, rho_inv, _ = app.operators()
    gids = [app.map_source(d, a) for d, a, p, dl in sources]
    B = np.stack([mv[g] * rho_inv[g] * phi_v[g, :] for g in gids], axis=1)
    ftab = shaw_amd.forcing_table(["ricker"] * len(sources), [s[3] for s in sources], [s[2] for s in sources],
                                  dt, nsteps)
    return B, ftab
