"""MI355X-native SHAW time-stepping path: Python front end over the C ABI.

The library (``shaw_amd/libshaw_amd.so``: gfx950 HIP kernels + C++ host
layer, declared in ``include/shaw_amd.h``) is loaded with ctypes. There is
no fallback: if the library is missing, or no GPU is present when a GPU
entry point is called, the call raises.

Mirrors of the reference's in-process interface (src/kokkos/*):
  App            ShWavePP (operator assembly from mesh files, shwavepp.hpp)
  App.map_*      source / receiver mappers (nominal_to_grid_mappers/)
  forcing_table  RankOne/RankTwoForcing signal evaluation (signal.hpp)
  Fom            runFom + updateVelocity/updateStress (fom_run.hpp,
                 fom_update_kernels.hpp), seismogram.hpp, state_observer.hpp
  Rom            rom_compute_jacobians.hpp + the reduced leapfrog
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libshaw_amd.so")
EXE_PATH = os.path.join(HERE, "shawExe")

FP64, FP32 = 0, 1
KERNEL_AUTO, KERNEL_SPLIT, KERNEL_FUSED = 0, 1, 2
ROM_FP64, ROM_BF16X3, ROM_BF16X4 = 0, 1, 2
SIGNAL_KINDS = {"ricker": 0, "sinusoid": 1, "gaussDer": 2}


class ShawError(RuntimeError):
    pass


class _Material(C.Structure):
    _fields_ = [("kind", C.c_int), ("nparams", C.c_int), ("params", C.POINTER(C.c_double))]


class _AppInfo(C.Structure):
    _fields_ = [("nr", C.c_uint64), ("nth", C.c_uint64), ("nvp", C.c_uint64), ("nsp", C.c_uint64),
                ("th_left_deg", C.c_double), ("th_right_deg", C.c_double), ("r_cmb_m", C.c_double),
                ("r_surf_m", C.c_double), ("dth", C.c_double), ("drr", C.c_double),
                ("min_vs", C.c_double), ("max_vs", C.c_double)]


class _FomOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("kernel", C.c_int), ("rank1_forcing", C.c_int),
                ("cols_per_thread", C.c_int), ("strip", C.c_int), ("segment", C.c_int),
                ("prefetch_rows", C.c_int), ("steps_per_launch", C.c_int), ("graph_steps", C.c_int),
                ("col_block", C.c_int)]


class _RomDesc(C.Structure):
    _fields_ = [("kv", C.c_uint64), ("ks", C.c_uint64), ("dt", C.c_double), ("precision", C.c_int),
                ("ar_v", C.c_void_p), ("ar_s", C.c_void_p), ("r", C.c_uint64), ("c", C.c_void_p),
                ("max_cols", C.c_uint64), ("max_samples", C.c_uint64)]


_lib = None


def build() -> None:
    """Compile libshaw_amd.so and shawExe for gfx950 (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ShawError(f"{LIB_PATH} is not built: run `make -C shaw_amd` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, u64, dbl, i32 = C.c_void_p, C.c_uint64, C.c_double, C.c_int
    P = C.POINTER
    sig = {
        "shaw_last_error": (C.c_char_p, []),
        "shaw_version": (C.c_char_p, []),
        "shaw_app_create_from_mesh": (i32, [C.c_char_p, P(_Material), P(vp)]),
        "shaw_app_create_structured": (i32, [u64, u64, dbl, dbl, dbl, dbl, P(_Material), P(vp)]),
        "shaw_mesh_write": (i32, [C.c_char_p, u64, u64, dbl, dbl, dbl, dbl]),
        "shaw_app_info_get": (i32, [vp, P(_AppInfo)]),
        "shaw_app_operators": (i32, [vp, vp, vp, vp, vp]),
        "shaw_app_coords": (i32, [vp, i32, vp]),
        "shaw_map_source": (i32, [vp, dbl, dbl, P(u64)]),
        "shaw_map_receiver": (i32, [vp, dbl, P(u64)]),
        "shaw_app_destroy": (None, [vp]),
        "shaw_forcing_table": (i32, [i32, vp, vp, vp, dbl, u64, vp]),
        "shaw_fom_create": (i32, [vp, u64, dbl, P(_FomOptions), P(vp)]),
        "shaw_fom_set_sources": (i32, [vp, vp]),
        "shaw_fom_set_forcing": (i32, [vp, u64, vp]),
        "shaw_fom_set_receivers": (i32, [vp, u64, vp, u64, u64]),
        "shaw_fom_set_snapshots": (i32, [vp, i32, u64, u64, vp, u64]),
        "shaw_fom_reset": (i32, [vp]),
        "shaw_fom_advance": (i32, [vp, u64]),
        "shaw_fom_sync": (i32, [vp]),
        "shaw_fom_step": (i32, [vp, P(u64)]),
        "shaw_fom_get_seismogram": (i32, [vp, vp, P(u64)]),
        "shaw_fom_get_snapshots": (i32, [vp, i32, vp, P(u64)]),
        "shaw_fom_get_state": (i32, [vp, vp, vp]),
        "shaw_fom_set_state": (i32, [vp, vp, vp]),
        "shaw_fom_time": (i32, [vp, u64, P(dbl), P(dbl)]),
        "shaw_fom_time_chunks": (i32, [vp, u64, u64, vp]),
        "shaw_app_energy_weights": (i32, [vp, vp, vp]),
        "shaw_basis_orthonormalize": (i32, [vp, u64, u64, vp, vp, vp]),
        "shaw_fom_complexity": (i32, [u64, u64, u64, u64, u64, i32, P(dbl), P(dbl)]),
        "shaw_fom_bytes_per_step": (i32, [vp, P(dbl), P(dbl)]),
        "shaw_fom_config": (i32, [vp, P(i32), P(i32), P(i32), P(i32), P(i32), P(i32), P(i32)]),
        "shaw_fom_autotune": (i32, [vp, u64, P(C.c_double)]),
        "shaw_fom_destroy": (None, [vp]),
        "shaw_fom_pod": (i32, [vp, i32, u64, vp, vp]),
        "shaw_fom_pod_weighted": (i32, [vp, i32, u64, vp, vp, vp]),
        "shaw_fom_pod_basis": (i32, [vp, u64, u64, vp, vp, P(vp)]),
        "shaw_app_assemble_device": (i32, [vp, P(_Material), vp, vp, vp, vp]),
        "shaw_forcing_table_device": (i32, [u64, vp, vp, vp, dbl, u64, vp, vp]),
        "shaw_map_points_device": (i32, [vp, u64, vp, vp, vp]),
        "shaw_fom_set_signals": (i32, [vp, vp, vp, vp, u64]),
        "shaw_fom_set_material": (i32, [vp, vp, P(_Material)]),
        "shaw_rom_admit_signals": (i32, [vp, u64, vp, vp, vp, vp, u64, vp, u64, vp]),
        "shaw_basis_info": (i32, [vp, P(u64), P(u64), P(u64), P(u64), P(u64), P(u64)]),
        "shaw_basis_sigma": (i32, [vp, i32, vp]),
        "shaw_basis_rows": (i32, [vp, i32, vp, u64, u64, vp]),
        "shaw_basis_download": (i32, [vp, i32, u64, vp]),
        "shaw_basis_project": (i32, [vp, vp, u64, u64, vp, vp, vp]),
        "shaw_basis_destroy": (None, [vp]),
    }
    optional = {
        "shaw_rom_project": (i32, [vp, vp, u64, vp, u64, vp, vp]),
        "shaw_rom_project_energy": (i32, [vp, vp, u64, vp, u64, vp, vp, vp]),
        "shaw_rom_create": (i32, [P(_RomDesc), P(vp)]),
        "shaw_rom_admit": (i32, [vp, u64, vp, u64, vp, vp, u64, vp]),
        "shaw_rom_advance": (i32, [vp, u64]),
        "shaw_rom_sync": (i32, [vp]),
        "shaw_rom_retire": (i32, [vp, u64, vp, vp, vp, P(u64)]),
        "shaw_rom_retire_batch": (i32, [vp, u64, vp, u64, vp, vp]),
        "shaw_rom_free_slots": (i32, [vp, P(u64)]),
        "shaw_rom_set_macro_steps": (i32, [vp, C.c_uint32]),
        "shaw_rom_set_time_budget": (i32, [vp, C.c_uint32]),
        "shaw_rom_steps": (i32, [vp, u64, vp, vp]),
        "shaw_rom_state": (i32, [vp, u64, vp, vp, P(u64)]),
        "shaw_debug_mfma": (i32, [i32, vp, vp, vp]),
        "shaw_pod_basis": (i32, [vp, u64, u64, u64, vp, vp]),
        "shaw_pod_basis_upto": (i32, [vp, u64, u64, u64, vp, vp, P(u64)]),
        "shaw_reconstruct": (i32, [vp, u64, u64, vp, u64, vp]),
        "shaw_debug_copy": (i32, [u64, i32, P(dbl)]),
        "shaw_debug_symeig": (i32, [i32, vp, vp, vp]),
        "shaw_rom_time": (i32, [vp, u64, P(dbl)]),
        "shaw_rom_destroy": (None, [vp]),
    }
    for name, (res, args) in list(sig.items()) + list(optional.items()):
        if not hasattr(L, name):
            if name in sig:
                raise ShawError(f"{LIB_PATH} does not export {name}")
            continue
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise ShawError(f"[{rc}] {lib().shaw_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---------------------------------------------------------------- material
def material_params(mat: dict) -> tuple[int, np.ndarray]:
    """Flatten a ``material`` input section (parser_material_model.hpp) into
    the C ABI's (kind, params)."""
    kind = str(mat["kind"])
    pad = lambda v: [float(x) for x in v] + [0.0] * (3 - len(v))
    k = kind.lower()
    if k == "unilayer":
        lay = mat["layer"]
        return 0, np.array(pad(lay["density"]) + pad(lay["velocity"]))
    if k == "bilayer":
        l1, l2 = mat["layer1"], mat["layer2"]
        return 1, np.array([float(l2["depth"])] + pad(l1["density"]) + pad(l1["velocity"])
                           + pad(l2["density"]) + pad(l2["velocity"]))
    if k == "prem":
        return 2, np.zeros(0)
    if k == "layered":
        r = [float(x) for x in mat["radius"]]
        return 3, np.array([float(len(r))] + r + [float(x) for x in mat["density"]]
                           + [float(x) for x in mat["velocity"]])
    raise ShawError(f"Unknown material model kind '{kind}'")


def _material_struct(mat: dict):
    kind, params = material_params(mat)
    params = _f64(params)
    m = _Material(kind, len(params), params.ctypes.data_as(C.POINTER(C.c_double)))
    return m, params  # keep params alive


def mesh_write(directory: str, nr: int, nth: int, th_left=0.0, th_right=180.0, r_cmb=3480.0, r_surf=6371.0):
    """Write mesh_info.dat/graph_vp.dat/coeff_vp.dat/graph_sp.dat as meshing/ does."""
    os.makedirs(directory, exist_ok=True)
    _check(lib().shaw_mesh_write(directory.encode(), nr, nth, th_left, th_right, r_cmb, r_surf))


class _DeviceBuffer:
    """Device memory of the current GPU through the HIP runtime (for the
    entry points that write device buffers); `host()` copies it back."""
    _hip = None

    def __init__(self, count: int):
        if _DeviceBuffer._hip is None:
            _DeviceBuffer._hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
            for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
                getattr(_DeviceBuffer._hip, f).restype = C.c_int
            _DeviceBuffer._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            _DeviceBuffer._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            _DeviceBuffer._hip.hipFree.argtypes = [C.c_void_p]
        self.count = count
        self.ptr = C.c_void_p()
        if _DeviceBuffer._hip.hipMalloc(C.byref(self.ptr), max(8, count * 8)) != 0:
            raise ShawError("hipMalloc failed")

    def host(self, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.count, dtype=dtype)
        if _DeviceBuffer._hip.hipDeviceSynchronize() != 0 or _DeviceBuffer._hip.hipMemcpy(
                out.ctypes.data_as(C.c_void_p), self.ptr, self.count * out.itemsize, 2) != 0:
            raise ShawError("device to host copy failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _DeviceBuffer._hip is not None:
            _DeviceBuffer._hip.hipFree(self.ptr)
            self.ptr = None


def _signal_arrays(kinds, delays, periods):
    kinds = np.ascontiguousarray([SIGNAL_KINDS[k] if isinstance(k, str) else int(k) for k in kinds], dtype=np.int32)
    return kinds, _f64(delays).ravel(), _f64(periods).ravel()


def forcing_table_device(kinds, delays, periods, dt: float, nsteps: int) -> np.ndarray:
    """forcing_table evaluated on the GPU (shaw_forcing_table_device), copied back."""
    k, d, p = _signal_arrays(kinds, delays, periods)
    buf = _DeviceBuffer(len(k) * nsteps)
    _check(lib().shaw_forcing_table_device(len(k), _p(k), _p(d), _p(p), dt, nsteps, buf.ptr, None))
    return buf.host().reshape(nsteps, len(k))


def forcing_table(kinds, delays, periods, dt: float, nsteps: int) -> np.ndarray:
    """[nsteps][m] forcing values f_m((s-1) dt) for steps s = 1..nsteps."""
    kinds = np.ascontiguousarray([SIGNAL_KINDS[k] if isinstance(k, str) else int(k) for k in kinds], dtype=np.int32)
    delays, periods = _f64(delays), _f64(periods)
    m = len(kinds)
    out = np.zeros((nsteps, m))
    _check(lib().shaw_forcing_table(m, _p(kinds), _p(delays), _p(periods), dt, nsteps, _p(out)))
    return out


class App:
    """Operators of one mesh + material (the ShWavePP replacement)."""

    def __init__(self, handle):
        self._h = handle
        info = _AppInfo()
        _check(lib().shaw_app_info_get(self._h, C.byref(info)))
        self.nr, self.nth, self.nVp, self.nSp = int(info.nr), int(info.nth), int(info.nvp), int(info.nsp)
        self.bounds = np.array([info.th_left_deg, info.th_right_deg, info.r_cmb_m, info.r_surf_m])
        self.dth, self.drr = info.dth, info.drr
        self.minVs, self.maxVs = info.min_vs, info.max_vs

    @classmethod
    def from_mesh(cls, mesh_dir: str, material: dict) -> "App":
        m, keep = _material_struct(material)
        h = C.c_void_p()
        _check(lib().shaw_app_create_from_mesh(mesh_dir.encode(), C.byref(m), C.byref(h)))
        return cls(h)

    @classmethod
    def structured(cls, nr: int, nth: int, material: dict, th_left=0.0, th_right=180.0,
                   r_cmb=3480.0, r_surf=6371.0) -> "App":
        m, keep = _material_struct(material)
        h = C.c_void_p()
        _check(lib().shaw_app_create_structured(nr, nth, th_left, th_right, r_cmb, r_surf,
                                                C.byref(m), C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.shaw_app_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def operators(self):
        cv = np.zeros((self.nVp, 4))
        cs = np.zeros((self.nSp, 2))
        ri = np.zeros(self.nVp)
        g = np.zeros(self.nSp)
        _check(lib().shaw_app_operators(self._h, _p(cv), _p(cs), _p(ri), _p(g)))
        return cv, cs, ri, g

    def coords(self, dof: str) -> np.ndarray:
        n = self.nVp if dof == "vp" else self.nSp
        c = np.zeros((n, 2))
        _check(lib().shaw_app_coords(self._h, 0 if dof == "vp" else 1, _p(c)))
        return c

    def energy_weights(self):
        """(Mv [nVp], Ms [nSp]): the FOM energy weights (shaw_app_energy_weights)."""
        mv, ms = np.zeros(self.nVp), np.zeros(self.nSp)
        _check(lib().shaw_app_energy_weights(self.handle, _p(mv), _p(ms)))
        return mv, ms

    def map_source(self, depth_km: float, angle_deg: float) -> int:
        g = C.c_uint64()
        _check(lib().shaw_map_source(self._h, depth_km, angle_deg, C.byref(g)))
        return int(g.value)

    def map_receiver(self, angle_deg: float) -> int:
        g = C.c_uint64()
        _check(lib().shaw_map_receiver(self._h, angle_deg, C.byref(g)))
        return int(g.value)

    def map_points_device(self, angles_deg, depths_km=None) -> np.ndarray:
        """gids of many sources (depths_km given) or surface receivers, mapped
        on the GPU in one launch (shaw_map_points_device)."""
        a = _f64(angles_deg).ravel()
        d = None if depths_km is None else _f64(depths_km).ravel()
        if d is not None and d.size != a.size:
            raise ShawError("map_points_device: depths and angles differ in length")
        g = np.zeros(a.size, dtype=np.uint64)
        _check(lib().shaw_map_points_device(self._h, a.size, None if d is None else _p(d), _p(a),
                                            g.ctypes.data_as(C.c_void_p)))
        return g

    def operators_device(self, material: dict):
        """(coef_vp [nVp x 4], coef_sp [nSp x 2], rho_inv [nVp]) of `material`
        on this mesh, assembled on the GPU (shaw_app_assemble_device)."""
        m, keep = _material_struct(material)
        cv, cs, ri = _DeviceBuffer(4 * self.nVp), _DeviceBuffer(2 * self.nSp), _DeviceBuffer(self.nVp)
        _check(lib().shaw_app_assemble_device(self._h, C.byref(m), cv.ptr, cs.ptr, ri.ptr, None))
        return cv.host().reshape(self.nVp, 4), cs.host().reshape(self.nSp, 2), ri.host()


class Fom:
    """Full-order leapfrog for m forcing columns on the GPU."""

    def __init__(self, app: App, m: int, dt: float, precision=FP64, kernel=KERNEL_AUTO,
                 rank1_forcing=None, cols_per_thread=0, strip=0, segment=0, prefetch_rows=0, steps_per_launch=0,
                 graph_steps=0, col_block=0):
        if rank1_forcing is None:
            rank1_forcing = m == 1
        o = _FomOptions(precision, kernel, int(bool(rank1_forcing)), cols_per_thread, strip, segment, prefetch_rows,
                        steps_per_launch, graph_steps, col_block)
        h = C.c_void_p()
        _check(lib().shaw_fom_create(app.handle, m, dt, C.byref(o), C.byref(h)))
        self._h, self.app, self.m, self.dt = h, app, m, dt
        self.R = 0
        self._seis_cap = 0
        self._snap_cols = {0: 0, 1: 0}
        self._snap_cap = {0: 0, 1: 0}

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.shaw_fom_destroy(self._h)
            self._h = None

    def close(self):
        self.__del__()

    def set_sources(self, gids):
        g = _u64(gids)
        assert len(g) == self.m
        _check(lib().shaw_fom_set_sources(self._h, _p(g)))

    def set_forcing(self, table):
        t = _f64(table)
        assert t.ndim == 2 and t.shape[1] == self.m
        _check(lib().shaw_fom_set_forcing(self._h, t.shape[0], _p(t)))

    def set_signals(self, kinds, delays, periods, nsteps: int):
        """The forcing table generated on the GPU (shaw_fom_set_signals)."""
        k, d, p = _signal_arrays(kinds, delays, periods)
        assert len(k) == self.m
        _check(lib().shaw_fom_set_signals(self._h, _p(k), _p(d), _p(p), nsteps))

    def set_material(self, app: "App", material: dict):
        """Operators of another material on the same mesh, assembled on the
        GPU (shaw_fom_set_material); the state is kept."""
        m, keep = _material_struct(material)
        _check(lib().shaw_fom_set_material(self._h, app.handle, C.byref(m)))

    def set_receivers(self, gids, every: int, max_samples: int):
        g = _u64(gids)
        _check(lib().shaw_fom_set_receivers(self._h, len(g), _p(g), every, max_samples))
        self.R, self._seis_cap = len(g), max_samples

    def set_snapshots(self, dof: str, every: int, cols, max_samples: int):
        c = _u64(cols)
        d = 0 if dof == "vp" else 1
        _check(lib().shaw_fom_set_snapshots(self._h, d, every, len(c), _p(c), max_samples))
        self._snap_cols[d], self._snap_cap[d] = len(c), max_samples

    def reset(self):
        _check(lib().shaw_fom_reset(self._h))

    def advance(self, nsteps: int):
        _check(lib().shaw_fom_advance(self._h, nsteps))

    def sync(self):
        _check(lib().shaw_fom_sync(self._h))

    @property
    def step(self) -> int:
        s = C.c_uint64()
        _check(lib().shaw_fom_step(self._h, C.byref(s)))
        return int(s.value)

    def seismogram(self) -> np.ndarray:
        """[m][T][R] (= LayoutLeft [R x T x m])."""
        out = np.zeros(max(1, self.R * self._seis_cap * self.m))  # at most max_samples samples
        t = C.c_uint64()
        _check(lib().shaw_fom_get_seismogram(self._h, _p(out), C.byref(t)))
        T = int(t.value)
        return out[: self.R * T * self.m].reshape(self.m, T, self.R)

    def snapshots(self, dof: str) -> np.ndarray:
        """[ncols][T][n] (= LayoutLeft [n x T x ncols])."""
        d = 0 if dof == "vp" else 1
        n = self.app.nVp if d == 0 else self.app.nSp
        nc = self._snap_cols[d]
        out = np.zeros(max(1, n * self._snap_cap[d] * nc))  # at most max_samples samples
        t = C.c_uint64()
        _check(lib().shaw_fom_get_snapshots(self._h, d, _p(out), C.byref(t)))
        T = int(t.value)
        return out[: n * T * nc].reshape(nc, T, n)

    def state(self):
        vp = np.zeros((self.app.nVp, self.m))
        sp = np.zeros((self.app.nSp, self.m))
        _check(lib().shaw_fom_get_state(self._h, _p(vp), _p(sp)))
        return vp, sp

    def set_state(self, vp, sp):
        _check(lib().shaw_fom_set_state(self._h, _p(_f64(vp)), _p(_f64(sp))))

    def time(self, nsteps: int):
        tot, per = C.c_double(), C.c_double()
        _check(lib().shaw_fom_time(self._h, nsteps, C.byref(tot), C.byref(per)))
        return tot.value, per.value

    def time_chunks(self, nsteps: int, chunk: int) -> np.ndarray:
        """Durations [ms] of the run's chunks of `chunk` steps (events between
        chunks, one wait at the end)."""
        out = np.zeros((nsteps + chunk - 1) // chunk)
        _check(lib().shaw_fom_time_chunks(self._h, nsteps, chunk, _p(out)))
        return out

    def bytes_per_step(self):
        k, m = C.c_double(), C.c_double()
        _check(lib().shaw_fom_bytes_per_step(self._h, C.byref(k), C.byref(m)))
        return k.value, m.value

    def pod(self, dof: str, k: int, nsnap: int, weights=None):
        """(Phi [n x k], sigma) of this run's snapshots of `dof` (consumes them);
        with `weights` [n] the POD is in the inner product diag(weights)."""
        n = self.app.nVp if dof == "vp" else self.app.nSp
        phi = np.zeros((k, n))
        sig = np.zeros(nsnap)
        d = 0 if dof == "vp" else 1
        if weights is None:
            _check(lib().shaw_fom_pod(self._h, d, k, _p(phi), _p(sig)))
        else:
            w = _f64(weights).ravel()
            if w.size != n:
                raise ShawError(f"POD weights: expected {n} values, got {w.size}")
            _check(lib().shaw_fom_pod_weighted(self._h, d, k, _p(w), _p(phi), _p(sig)))
        return phi.T, sig

    def pod_basis(self, kv: int, ks: int, wv=None, ws=None) -> "Basis":
        """POD of both dofs' snapshots (consumed) kept on the device
        (shaw_fom_pod_basis): Euclidean, or in the inner products diag(wv),
        diag(ws)."""
        if (wv is None) != (ws is None):
            raise ShawError("pod_basis: give both weight vectors or neither")
        w1 = None if wv is None else _f64(wv).ravel()
        w2 = None if ws is None else _f64(ws).ravel()
        if w1 is not None and (w1.size != self.app.nVp or w2.size != self.app.nSp):
            raise ShawError("pod_basis: weights must have nVp and nSp values")
        h = C.c_void_p()
        _check(lib().shaw_fom_pod_basis(self._h, kv, ks, None if w1 is None else _p(w1),
                                        None if w2 is None else _p(w2), C.byref(h)))
        return Basis(h)

    def autotune(self, trial_steps: int = 3) -> float:
        """Pick the fastest fused-kernel tile shape (state unchanged); returns
        its ms per step."""
        ms = C.c_double()
        _check(lib().shaw_fom_autotune(self._h, trial_steps, C.byref(ms)))
        return ms.value

    def config(self):
        v = [C.c_int() for _ in range(7)]
        _check(lib().shaw_fom_config(self._h, *[C.byref(x) for x in v]))
        return {"kernel": v[0].value, "cols_per_thread": v[1].value, "strip": v[2].value, "segment": v[3].value,
                "prefetch_rows": v[4].value, "steps_per_launch": v[5].value, "col_block": v[6].value}


def fom_run(app: App, dt: float, nsteps: int, sources, rank2: bool, receivers=(), seis_freq=0,
            snap_freq_vp=0, snap_freq_sp=0, signal="ricker", **fom_kw):
    """One runFom call (fom_run.hpp:57-161) on the GPU for the realizations
    ``sources`` = [(depth_km, angle_deg, period_s, delay_s), ...]; same result
    layout as the oracle: seismogram [m][T][R], snaps [m][T][n], final states."""
    m = len(sources)
    if not rank2 and m != 1:
        raise ShawError("rank-1 run needs one source")
    rank1_forcing = fom_kw.pop("rank1_forcing", not rank2)
    tune = fom_kw.pop("autotune", 0)
    fom = Fom(app, m, dt, rank1_forcing=rank1_forcing, **fom_kw)
    fom.set_sources([app.map_source(d, a) for d, a, p, dl in sources])
    fom.set_forcing(forcing_table([signal] * m, [s[3] for s in sources], [s[2] for s in sources], dt, nsteps))
    if seis_freq and len(receivers):
        fom.set_receivers([app.map_receiver(a) for a in receivers], seis_freq, nsteps // seis_freq)
    cols = list(range(m))
    if snap_freq_vp:
        fom.set_snapshots("vp", snap_freq_vp, cols, nsteps // snap_freq_vp)
    if snap_freq_sp:
        fom.set_snapshots("sp", snap_freq_sp, cols, nsteps // snap_freq_sp)
    if tune:
        fom.autotune(tune)
    fom.advance(nsteps)
    fom.sync()
    vp, sp = fom.state()
    out = {"final_vp": vp, "final_sp": sp}
    if seis_freq and len(receivers):
        out["seismogram"] = fom.seismogram()
    if snap_freq_vp:
        out["snaps_vp"] = fom.snapshots("vp")
    if snap_freq_sp:
        out["snaps_sp"] = fom.snapshots("sp")
    fom.close()
    return out


# ---------------------------------------------------------------- ROM
def rom_project(app: App, phi_v: np.ndarray, phi_s: np.ndarray):
    """Galerkin operators Ar_v = Phi_v^T J_vp Phi_s, Ar_s = Phi_s^T J_sp Phi_v
    (FP64, on the GPU). phi_v [nVp x Kv], phi_s [nSp x Ks]."""
    kv, ks = phi_v.shape[1], phi_s.shape[1]
    pv = np.asfortranarray(phi_v, dtype=np.float64)
    ps = np.asfortranarray(phi_s, dtype=np.float64)
    arv = np.zeros((kv, ks))
    ars = np.zeros((ks, kv))
    _check(lib().shaw_rom_project(app.handle, _p(pv), kv, _p(ps), ks, _p(arv), _p(ars)))
    return arv, ars


def rom_project_energy(app: App, phi_v: np.ndarray, phi_s: np.ndarray, mv: np.ndarray):
    """Energy-consistent operators Ar_v = Phi_v^T diag(mv) J_vp Phi_s and
    Ar_s = -Ar_v^T (FP64, on the GPU), for bases orthonormal in the FOM energy
    weights (an addition to the reference's Euclidean projection)."""
    kv, ks = phi_v.shape[1], phi_s.shape[1]
    pv = np.asfortranarray(phi_v, dtype=np.float64)
    ps = np.asfortranarray(phi_s, dtype=np.float64)
    w = _f64(mv).ravel()
    if w.size != app.nVp:
        raise ShawError(f"energy weights: expected {app.nVp} values, got {w.size}")
    arv = np.zeros((kv, ks))
    ars = np.zeros((ks, kv))
    _check(lib().shaw_rom_project_energy(app.handle, _p(pv), kv, _p(ps), ks, _p(w), _p(arv), _p(ars)))
    return arv, ars


class Basis:
    """POD bases of both dofs resident on the device (shaw_basis_*). dof:
    "vp" or "sp"; the first k columns of a basis are its k-mode basis."""

    def __init__(self, handle):
        self._h = handle
        vals = [C.c_uint64() for _ in range(6)]
        _check(lib().shaw_basis_info(self._h, *[C.byref(v) for v in vals]))
        self.n = {"vp": vals[0].value, "sp": vals[1].value}
        self.k = {"vp": vals[2].value, "sp": vals[3].value}
        self.nsnap = {"vp": vals[4].value, "sp": vals[5].value}

    @staticmethod
    def _dof(dof):
        if dof not in ("vp", "sp"):
            raise ShawError(f"unknown dof {dof!r}")
        return 0 if dof == "vp" else 1

    def sigma(self, dof: str) -> np.ndarray:
        out = np.zeros(self.nsnap[dof])
        _check(lib().shaw_basis_sigma(self._h, self._dof(dof), _p(out)))
        return out

    def rows(self, dof: str, gids, k: int = 0) -> np.ndarray:
        """[m x k]: the first k modes (all by default) at the grid points gids."""
        g = _u64(gids)
        k = k or self.k[dof]
        out = np.zeros((g.size, k))
        _check(lib().shaw_basis_rows(self._h, self._dof(dof), _p(g), g.size, k, _p(out)))
        return out

    def download(self, dof: str, k: int = 0) -> np.ndarray:
        """Phi [n x k] (Fortran order), the first k modes."""
        k = k or self.k[dof]
        out = np.zeros((k, self.n[dof]))
        _check(lib().shaw_basis_download(self._h, self._dof(dof), k, _p(out)))
        return out.T

    def project(self, app: "App", kv: int, ks: int, mv=None):
        """(Ar_v [kv x ks], Ar_s [ks x kv]) of the first kv / ks modes: the
        reference's Galerkin pair, or with mv the energy-consistent pair."""
        arv, ars = np.zeros((kv, ks)), np.zeros((ks, kv))
        w = None if mv is None else _f64(mv).ravel()
        _check(lib().shaw_basis_project(app.handle, self._h, kv, ks, None if w is None else _p(w), _p(arv),
                                        _p(ars)))
        return arv, ars

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.shaw_basis_destroy(self._h)
            self._h = None

    def close(self):
        self.__del__()


class Rom:
    """Resident reduced-order models for many forcing columns (slots)."""

    def __init__(self, ar_v, ar_s, dt: float, precision=ROM_FP64, receivers=None, max_cols=16, max_samples=0):
        """receivers: [R x Kv] rows of Phi_v at the receiver gids (or None)."""
        self._arv, self._ars = _f64(ar_v), _f64(ar_s)
        self.kv, self.ks = self._arv.shape
        self._C = None if receivers is None else _f64(receivers)
        self.R = 0 if receivers is None else self._C.shape[0]
        d = _RomDesc(self.kv, self.ks, dt, precision, _p(self._arv), _p(self._ars), self.R, _p(self._C),
                     max_cols, max_samples)
        hh = C.c_void_p()
        _check(lib().shaw_rom_create(C.byref(d), C.byref(hh)))
        self._h = hh
        self.every = {}

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.shaw_rom_destroy(self._h)
            self._h = None

    def close(self):
        self.__del__()

    def admit(self, b, ftab, horizons=None, every=0):
        """b [Kv x n], ftab [nsteps x n]; returns the slot ids."""
        b, ftab = _f64(b), _f64(ftab)
        n = b.shape[1]
        hz = _u64(horizons if horizons is not None else [ftab.shape[0]] * n)
        slots = np.zeros(n, np.uint64)
        _check(lib().shaw_rom_admit(self._h, n, _p(b), ftab.shape[0], _p(ftab), _p(hz), every, _p(slots)))
        for s in slots:
            self.every[int(s)] = every
        return [int(s) for s in slots]

    def admit_signals(self, b, kinds, delays, periods, nsteps: int, horizons=None, every=0):
        """admit with each column's forcing series evaluated on the GPU from
        its signal (shaw_rom_admit_signals)."""
        b = _f64(b)
        n = b.shape[1]
        k, d, p = _signal_arrays(kinds, delays, periods)
        hz = _u64(horizons if horizons is not None else [nsteps] * n)
        slots = np.zeros(n, np.uint64)
        _check(lib().shaw_rom_admit_signals(self._h, n, _p(b), _p(k), _p(d), _p(p), nsteps, _p(hz), every,
                                            _p(slots)))
        for sl in slots:
            self.every[int(sl)] = every
        return [int(sl) for sl in slots]

    def set_macro_steps(self, m: int):
        """BF16 engines: quiet, aligned tiles advance m steps per product
        (shaw_rom_set_macro_steps); 0 turns it off."""
        _check(lib().shaw_rom_set_macro_steps(self._h, m))

    def set_time_budget(self, w: int):
        """Engines with macro steps: advance(n) gives every tile the time of
        n one-steps, a macro step counting as w (shaw_rom_set_time_budget);
        0 turns it off."""
        _check(lib().shaw_rom_set_time_budget(self._h, w))

    def steps(self, slots):
        """Step count of each slot once the enqueued work is done."""
        sl = np.ascontiguousarray(slots, dtype=np.uint64)
        out = np.zeros(max(1, sl.size), dtype=np.uint64)
        _check(lib().shaw_rom_steps(self._h, sl.size, sl.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
        return [int(x) for x in out[:sl.size]]

    def advance(self, nsteps: int):
        _check(lib().shaw_rom_advance(self._h, nsteps))

    def sync(self):
        _check(lib().shaw_rom_sync(self._h))

    def state(self, slot: int):
        qv, qs = np.zeros(self.kv), np.zeros(self.ks)
        st = C.c_uint64()
        _check(lib().shaw_rom_state(self._h, slot, _p(qv), _p(qs), C.byref(st)))
        return qv, qs, int(st.value)

    def retire(self, slot: int, max_samples: int = 0):
        """-> (seismogram [T][R], qv, qs)."""
        seis = np.zeros(max(1, self.R * max(1, max_samples)))
        qv, qs = np.zeros(self.kv), np.zeros(self.ks)
        t = C.c_uint64()
        _check(lib().shaw_rom_retire(self._h, slot, _p(seis), _p(qv), _p(qs), C.byref(t)))
        T = int(t.value)
        return seis[: self.R * T].reshape(T, self.R), qv, qs

    def retire_many(self, slots, max_samples: int):
        """Seismograms [T_j][R] of several slots, retired with one sync and one
        staged copy (states are not returned)."""
        sl = np.ascontiguousarray(slots, dtype=np.uint64)
        n = sl.size
        seis = np.empty(max(1, n * self.R * max(1, max_samples)))
        t = np.zeros(max(1, n), dtype=np.uint64)
        _check(lib().shaw_rom_retire_batch(self._h, n, sl.ctypes.data_as(C.c_void_p), max_samples, _p(seis),
                                           t.ctypes.data_as(C.c_void_p)))
        out = []
        for j in range(n):
            T = int(t[j])
            blk = seis[j * self.R * max_samples: j * self.R * max_samples + self.R * T]
            out.append(blk.reshape(T, self.R))
        return out

    def free_slots(self) -> int:
        n = C.c_uint64()
        _check(lib().shaw_rom_free_slots(self._h, C.byref(n)))
        return int(n.value)

    def time(self, nsteps: int) -> float:
        t = C.c_double()
        _check(lib().shaw_rom_time(self._h, nsteps, C.byref(t)))
        return t.value


def rom_run(ar_v, ar_s, B, ftab, dt, nsteps, C=None, seis_freq=0, precision=ROM_FP64):
    """Batched reduced leapfrog of all columns of B [Kv x M] for nsteps; same
    result layout as the oracle (seismogram [M][T][R], qv [Kv x M], qs [Ks x M])."""
    M = B.shape[1]
    T = nsteps // seis_freq if seis_freq else 0
    rom = Rom(ar_v, ar_s, dt, precision, C, max_cols=M, max_samples=T)  # C: receiver rows
    slots = rom.admit(B, np.asarray(ftab)[:nsteps], [nsteps] * M, seis_freq if C is not None else 0)
    rom.advance(nsteps)
    rom.sync()
    R = 0 if C is None else C.shape[0]
    seis = np.zeros((M, T, R))
    qv = np.zeros((ar_v.shape[0], M))
    qs = np.zeros((ar_s.shape[0], M))
    for j, s in enumerate(slots):
        y, qv[:, j], qs[:, j] = rom.retire(s, T)
        if R and T:
            seis[j] = y
    rom.close()
    return {"seismogram": seis if (R and T) else None, "qv": qv, "qs": qs}


def debug_mfma(kind: int, a, b) -> np.ndarray:
    """Run one MFMA with lane-wise operands (see include/shaw_amd.h)."""
    a, b = _f64(a).ravel(), _f64(b).ravel()
    d = np.zeros(256)
    _check(lib().shaw_debug_mfma(kind, _p(a), _p(b), _p(d)))
    return d.reshape(64, 4)


def pod_basis(S: np.ndarray, k: int):
    """(Phi [n x k], sigma [ns]) of the snapshot matrix S [n x ns] (GPU, FP64)."""
    S = np.asfortranarray(S, dtype=np.float64)
    n, ns = S.shape
    phi = np.zeros((k, n))  # column-major [n x k]
    sig = np.zeros(ns)
    _check(lib().shaw_pod_basis(_p(S), n, ns, k, _p(phi), _p(sig)))
    return phi.T, sig


def pod_basis_upto(S: np.ndarray, kmax: int):
    """(Phi [n x k], sigma [ns]) with k <= kmax stopping at the numerical rank."""
    S = np.asfortranarray(S, dtype=np.float64)
    n, ns = S.shape
    kmax = min(kmax, n, ns)
    phi = np.zeros((kmax, n))
    sig = np.zeros(ns)
    k = C.c_uint64()
    _check(lib().shaw_pod_basis_upto(_p(S), n, ns, kmax, _p(phi), _p(sig), C.byref(k)))
    return phi[: k.value].T, sig


def reconstruct(phi: np.ndarray, q: np.ndarray) -> np.ndarray:
    """phi [n x k] @ q [k x ncols] on the GPU (FP64, sums in mode order, no FMA)."""
    phi = np.asfortranarray(phi, dtype=np.float64)
    q = np.asfortranarray(q, dtype=np.float64)
    n, k = phi.shape
    assert q.shape[0] == k
    out = np.zeros((q.shape[1], n))  # column-major [n x ncols]
    _check(lib().shaw_reconstruct(_p(phi), n, k, _p(q), q.shape[1], _p(out)))
    return out.T


def fom_complexity(nvp: int, nsp: int, nnz_vp: int, nnz_sp: int, m: int, rank2: bool):
    """(memMB, flops) of one FOM step in the reference's printPerf model."""
    mb, fl = C.c_double(), C.c_double()
    _check(lib().shaw_fom_complexity(nvp, nsp, nnz_vp, nnz_sp, m, int(rank2), C.byref(mb), C.byref(fl)))
    return mb.value, fl.value


def basis_orthonormalize(phi: np.ndarray, w: np.ndarray):
    """(phi_out, R): phi [n x k] made orthonormal in diag(w), phi_out = phi R^-1
    (shaw_basis_orthonormalize)."""
    n, k = phi.shape
    p = np.asfortranarray(phi, dtype=np.float64)
    out = np.zeros((n, k), order="F")
    R = np.zeros((k, k))
    _check(lib().shaw_basis_orthonormalize(_p(p), n, k, _p(_f64(w).ravel()), _p(out), _p(R)))
    return out, R


def debug_copy(nbytes: int, reps: int = 10) -> float:
    """ms per launch of the plain copy kernel over nbytes (read + write)."""
    ms = C.c_double()
    _check(lib().shaw_debug_copy(nbytes, reps, C.byref(ms)))
    return ms.value


def debug_symeig(a: np.ndarray):
    """(w, V) of the host eigensolver the POD uses (CPU only)."""
    a = _f64(a)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    _check(lib().shaw_debug_symeig(n, _p(a), _p(w), _p(v)))
    return w, v
