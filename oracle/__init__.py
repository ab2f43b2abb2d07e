"""CPU oracle for the SHAW time-stepping path (test infrastructure only).

A ctypes front end over ``oracle/liboracle.so``, the C++ restatement of the
reference's FP64 FOM/ROM (``oracle/shaw_oracle.cpp``; each function there
cites the reference file:line it follows).  Only the tests, the smoke check
and the CPU-baseline leg of the benchmark use it, and only to check results
against: the product path never imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SIGNAL_KINDS = {"ricker": 0, "sinusoid": 1, "gaussDer": 2}
MATERIAL_KINDS = {"unilayer": 0, "bilayer": 1, "prem": 2, "layered": 3}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    lib.orc_last_error.restype = C.c_char_p
    lib.orc_app_create.restype = C.c_void_p
    lib.orc_app_create.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int]
    lib.orc_app_create_structured.restype = C.c_void_p
    lib.orc_app_create_structured.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_int, C.c_void_p, C.c_int]
    lib.orc_app_destroy.argtypes = [C.c_void_p]
    lib.orc_app_info.argtypes = [C.c_void_p, _up, _dp]
    lib.orc_app_csr.argtypes = [C.c_void_p, C.c_int, _up, _up, _dp]
    lib.orc_app_graph.argtypes = [C.c_void_p, C.c_int, _up, _dp, C.c_void_p]
    lib.orc_app_material.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_map_source.restype = C.c_int64
    lib.orc_map_source.argtypes = [C.c_void_p, C.c_double, C.c_double]
    lib.orc_map_receiver.restype = C.c_int64
    lib.orc_map_receiver.argtypes = [C.c_void_p, C.c_double]
    lib.orc_material_at.restype = None
    lib.orc_material_at.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_uint64, _dp, _dp, _dp]
    lib.orc_signal.restype = C.c_double
    lib.orc_signal.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    vp = C.c_void_p
    lib.orc_fom_run.restype = C.c_int
    lib.orc_fom_run.argtypes = [vp, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                _dp, _dp, _dp, _dp, C.c_uint64, _dp, C.c_uint64, vp,
                                C.c_uint64, vp, C.c_uint64, vp, vp, vp]
    lib.orc_rom_project.restype = C.c_int
    lib.orc_rom_project.argtypes = [vp, _dp, C.c_uint64, _dp, C.c_uint64, _dp, _dp]
    lib.orc_rom_run.restype = C.c_int
    lib.orc_rom_run.argtypes = [C.c_uint64, C.c_uint64, _dp, _dp, _dp, _dp, C.c_double,
                                C.c_uint64, C.c_uint64, C.c_uint64, _dp, C.c_uint64, vp, vp, vp]
    lib.orc_reconstruct.restype = C.c_int
    lib.orc_reconstruct.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, vp]
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def material_params(mat: dict) -> tuple[int, np.ndarray]:
    """Flatten a parsed ``material`` YAML section into (kind, params)."""
    kind = mat["kind"].lower()
    if kind == "unilayer":
        lay = mat["layer"]
        rho = list(lay["density"]) + [0.0] * (3 - len(lay["density"]))
        vs = list(lay["velocity"]) + [0.0] * (3 - len(lay["velocity"]))
        return 0, np.array(rho + vs, dtype=np.float64)
    if kind == "bilayer":
        l1, l2 = mat["layer1"], mat["layer2"]
        pad = lambda v: list(v) + [0.0] * (3 - len(v))
        p = [l2["depth"]] + pad(l1["density"]) + pad(l1["velocity"]) + pad(l2["density"]) + pad(l2["velocity"])
        return 1, np.array(p, dtype=np.float64)
    if kind == "prem":
        return 2, np.zeros(1)
    if kind == "layered":
        r = [float(x) for x in mat["radius"]]
        p = [float(len(r))] + r + [float(x) for x in mat["density"]] + [float(x) for x in mat["velocity"]]
        return 3, np.array(p, dtype=np.float64)
    raise ValueError(f"unknown material kind {kind}")


class App:
    """Operators assembled by the oracle from a mesh directory (CSR, as ShWavePP)."""

    def __init__(self, mesh_dir: str, material: dict, structured=None):
        """structured = (nr, nth, thL_deg, thR_deg, rCmb_km, rSurf_km) builds the
        mesh in memory (restated generator) instead of reading mesh_dir."""
        lib = _load()
        kind, params = material_params(material)
        self._params = np.ascontiguousarray(params)
        if structured is None:
            h = lib.orc_app_create(mesh_dir.encode(), kind, _ptr(self._params), len(self._params))
        else:
            nr, nth, thl, thr, rc, rs = structured
            h = lib.orc_app_create_structured(nr, nth, thl, thr, rc, rs, kind, _ptr(self._params),
                                              len(self._params))
        if not h:
            raise RuntimeError(lib.orc_last_error().decode())
        self._h = h
        info = np.zeros(6, np.uint64)
        dinfo = np.zeros(8, np.float64)
        lib.orc_app_info(h, info, dinfo)
        self.nVp, self.nSp, self.nth, self.nr, self.nnzVp, self.nnzSp = (int(x) for x in info)
        self.bounds = dinfo[:4].copy()
        self.dth, self.drr, self.minVs, self.maxVs = (float(x) for x in dinfo[4:])

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_app_destroy(self._h)
            self._h = None

    def csr(self, dof: str):
        n, nnz = (self.nVp, self.nnzVp) if dof == "vp" else (self.nSp, self.nnzSp)
        ptr = np.zeros(n + 1, np.uint64)
        ind = np.zeros(nnz, np.uint64)
        val = np.zeros(nnz, np.float64)
        _lib.orc_app_csr(self._h, 0 if dof == "vp" else 1, ptr, ind, val)
        return ptr, ind, val

    def dense(self, dof: str) -> np.ndarray:
        ptr, ind, val = self.csr(dof)
        n, m = (self.nVp, self.nSp) if dof == "vp" else (self.nSp, self.nVp)
        A = np.zeros((n, m))
        for r in range(n):
            for k in range(int(ptr[r]), int(ptr[r + 1])):
                A[r, int(ind[k])] += val[k]
        return A

    def graph(self, dof: str):
        n, w = (self.nVp, 5) if dof == "vp" else (self.nSp, 3)
        g = np.zeros(n * w, np.uint64)
        c = np.zeros(n * 2, np.float64)
        lab = np.zeros(n, np.int32)
        _lib.orc_app_graph(self._h, 0 if dof == "vp" else 1, g, c, _ptr(lab))
        return g.reshape(n, w), c.reshape(n, 2), lab

    def material(self):
        ri = np.zeros(self.nVp)
        g = np.zeros(self.nSp)
        _lib.orc_app_material(self._h, ri, g)
        return ri, g

    def map_source(self, depth_km: float, angle_deg: float) -> int:
        r = _lib.orc_map_source(self._h, depth_km, angle_deg)
        if r < 0:
            raise RuntimeError(_lib.orc_last_error().decode())
        return int(r)

    def map_receiver(self, angle_deg: float) -> int:
        r = _lib.orc_map_receiver(self._h, angle_deg)
        if r < 0:
            raise RuntimeError(_lib.orc_last_error().decode())
        return int(r)

    def fom_run(self, dt, nsteps, sources, rank2, receivers=(), seis_freq=0,
                snap_freq_vp=0, snap_freq_sp=0, signal="ricker"):
        """sources: list of (depth_km, angle_deg, period_s, delay_s)."""
        M = len(sources)
        src = np.array(sources, dtype=np.float64).reshape(M, 4)
        d, a, p, dl = (np.ascontiguousarray(src[:, k]) for k in range(4))
        rec = np.ascontiguousarray(np.array(receivers, dtype=np.float64).reshape(-1))
        R = len(rec)
        if R == 0:
            rec = np.zeros(1)
        seis = np.zeros((M, nsteps // seis_freq, R)) if (seis_freq and R) else None
        sv = np.zeros((M, nsteps // snap_freq_vp, self.nVp)) if snap_freq_vp else None
        ss = np.zeros((M, nsteps // snap_freq_sp, self.nSp)) if snap_freq_sp else None
        fv = np.zeros((self.nVp, M))
        fs = np.zeros((self.nSp, M))
        rc = _lib.orc_fom_run(self._h, dt, nsteps, M, int(rank2), SIGNAL_KINDS[signal], d, a, p, dl,
                              R, rec, seis_freq if R else 0, _ptr(seis), snap_freq_vp, _ptr(sv),
                              snap_freq_sp, _ptr(ss), _ptr(fv), _ptr(fs))
        if rc != 0:
            raise RuntimeError(_lib.orc_last_error().decode())
        # LayoutLeft [n x T x M] == C-order [M][T][n]
        out = {"final_vp": fv, "final_sp": fs}
        if seis is not None:
            out["seismogram"] = seis
        if sv is not None:
            out["snaps_vp"] = sv
        if ss is not None:
            out["snaps_sp"] = ss
        return out

    def rom_project(self, phiV: np.ndarray, phiS: np.ndarray):
        """phiV [nVp x Kv], phiS [nSp x Ks] -> (ArV [Kv x Ks], ArS [Ks x Kv])."""
        Kv, Ks = phiV.shape[1], phiS.shape[1]
        pv = np.ascontiguousarray(phiV.T)  # column-major storage of phiV
        ps = np.ascontiguousarray(phiS.T)
        ArV = np.zeros((Kv, Ks))
        ArS = np.zeros((Ks, Kv))
        _lib.orc_rom_project(self._h, pv, Kv, ps, Ks, ArV, ArS)
        return ArV, ArS


def material_at(material: dict, radius_m, r_surf_km: float = 6371.0):
    """(rho [kg/m^3], vs [m/s]) of a parsed material section at the given
    radii [m] (Material::computeAt), for a domain whose surface is at
    r_surf_km."""
    kind, params = material_params(material)
    r = np.ascontiguousarray(np.atleast_1d(np.asarray(radius_m, dtype=np.float64)))
    rho, vs = np.empty_like(r), np.empty_like(r)
    params = np.ascontiguousarray(params, dtype=np.float64)
    _load().orc_material_at(kind, params, len(params), r_surf_km * 1000.0, len(r), r, rho, vs)
    return rho, vs


def signal(kind: str, delay: float, period: float, t: float) -> float:
    return float(_load().orc_signal(SIGNAL_KINDS[kind], delay, period, t))


def rom_run(ArV, ArS, B, ftab, dt, nsteps, C=None, seis_freq=0):
    """FP64 rank-2 reduced leapfrog. B [Kv x M], ftab [nsteps x M], C [R x Kv]."""
    lib = _load()
    Kv, Ks = ArV.shape
    M = B.shape[1]
    R = 0 if C is None else C.shape[0]
    Cc = np.zeros((1, Kv)) if C is None else np.ascontiguousarray(C, dtype=np.float64)
    seis = np.zeros((M, nsteps // seis_freq, R)) if (seis_freq and R) else None
    qv = np.zeros((Kv, M))
    qs = np.zeros((Ks, M))
    lib.orc_rom_run(Kv, Ks, np.ascontiguousarray(ArV), np.ascontiguousarray(ArS),
                    np.ascontiguousarray(B), np.ascontiguousarray(ftab), dt, nsteps, M, R, Cc,
                    seis_freq if R else 0, _ptr(seis), _ptr(qv), _ptr(qs))
    return {"seismogram": seis, "qv": qv, "qs": qs}


def reconstruct(phi: np.ndarray, q: np.ndarray) -> np.ndarray:
    """phi [n x K] q [K x ncols] as reconstructSeismogram sums it (mode order)."""
    phi = np.asfortranarray(phi, dtype=np.float64)
    q = np.asfortranarray(q, dtype=np.float64)
    n, K = phi.shape
    out = np.zeros((q.shape[1], n))
    _load().orc_reconstruct(_ptr(phi), n, K, _ptr(q), q.shape[1], _ptr(out))
    return out.T
