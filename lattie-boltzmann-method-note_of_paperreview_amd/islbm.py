"""ctypes binding of the gfx950 ISLBM solver library (include/islbm.h).

The library is loaded from this package's lib/ folder; a missing library is an error,
never a silent fallback.  Arrays are numpy float64 in unique-node layout:
populations (19, nz, ny, nx), macroscopic fields (nz, ny, nx), with the reference's axes
(x spanwise, y streamwise, z wall-normal).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libislbm_amd.so")

FP64, FP32, BF16, FP16 = 0, 1, 2, 3
MESH_BISECT, MESH_STRETCH = 0, 1
KERNEL_AUTO, KERNEL_PERNODE, KERNEL_TILED, KERNEL_TWOPASS, KERNEL_TILED_WIDE = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16": BF16, "fp16": FP16}

HEADER_SYMBOLS = [
    "islbm_default_params", "islbm_required_bytes", "islbm_create", "islbm_destroy",
    "islbm_init_equilibrium", "islbm_upload_f", "islbm_download_f", "islbm_download_macros",
    "islbm_step", "islbm_sync", "islbm_modify_forcing", "islbm_get_force", "islbm_set_force",
    "islbm_set_rho_modify", "islbm_get_flux_sum", "islbm_scalars", "islbm_step_count", "islbm_set_step_count",
    "islbm_kernel_in_use", "islbm_tile_shape", "islbm_get_table", "islbm_time_steps", "islbm_last_error",
    "islbm_selftest_collision", "islbm_copy_bandwidth",
    "islbm_ref_init", "islbm_ref_shutdown", "islbm_ref_device_arrays",
    "Launch_CollisionStreaming", "Launch_ModifyForcingTerm",
    "islbm_step_graph", "islbm_stream", "islbm_device_bytes",
    "islbm_batch_create", "islbm_batch_destroy", "islbm_batch_submit", "islbm_batch_advance",
    "islbm_batch_status", "islbm_batch_result", "islbm_batch_summary", "islbm_batch_bytes_in_use",
    "islbm_batch_last_error", "islbm_batch_now", "islbm_batch_group_stats",
    "islbm_upload_macros", "islbm_get_params", "islbm_mass_correct", "islbm_mass_correct_noalloc", "islbm_monitor",
    "islbm_stats_enable", "islbm_stats_disable", "islbm_stats_accumulate", "islbm_stats_reset", "islbm_stats_count",
    "islbm_stats_set_count", "islbm_stats_names", "islbm_stats_download", "islbm_stats_upload",
    "islbm_checkpoint_write", "islbm_checkpoint_read", "islbm_stats_write", "islbm_stats_read",
    "islbm_monitor_append", "islbm_io_last_error",
]
STAT_NAMES = ["U", "V", "W", "P", "UU", "UV", "UW", "VV", "VW", "WW", "PU", "PV", "PW", "PP", "KT",
              "DUDX2", "DUDY2", "DUDZ2", "DVDX2", "DVDY2", "DVDZ2", "DWDX2", "DWDY2", "DWDZ2",
              "UUU", "UUV", "UUW", "VVU", "VVV", "VVW", "WWU", "WWV", "WWW"]
REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class Params(ctypes.Structure):
    """Run-time form of the reference's compile-time parameter set (variables.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("cfl", ctypes.c_double), ("lxi", ctypes.c_double), ("hill_alpha", ctypes.c_double),
        ("flat", ctypes.c_int32), ("mesh_mode", ctypes.c_int32), ("stretch", ctypes.c_double),
        ("tau", ctypes.c_double), ("re", ctypes.c_double), ("u_ref", ctypes.c_double),
        ("cs_smag", ctypes.c_double), ("ndtfrc", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("kernel", ctypes.c_int32), ("device", ctypes.c_int32), ("xi_near_wall", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ISLBMError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the solver library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ISLBMError(f"solver library missing: {LIB_PATH} (run build.py)")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, i32 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32
    pp = ctypes.POINTER(Params)
    L.islbm_default_params.argtypes = [pp]
    L.islbm_required_bytes.argtypes = [pp, ctypes.POINTER(ctypes.c_size_t)]
    L.islbm_create.argtypes = [pp, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.islbm_destroy.argtypes = [vp]
    L.islbm_init_equilibrium.argtypes = [vp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_uint64]
    L.islbm_upload_f.argtypes = [vp, dp]
    L.islbm_download_f.argtypes = [vp, dp]
    L.islbm_download_macros.argtypes = [vp, dp, dp, dp, dp]
    L.islbm_step.argtypes = [vp, i32]
    L.islbm_sync.argtypes = [vp]
    L.islbm_modify_forcing.argtypes = [vp]
    L.islbm_get_force.argtypes = [vp, dp]
    L.islbm_set_force.argtypes = [vp, ctypes.c_double]
    L.islbm_set_rho_modify.argtypes = [vp, ctypes.c_double]
    L.islbm_get_flux_sum.argtypes = [vp, dp]
    L.islbm_scalars.argtypes = [vp, dp]
    L.islbm_step_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.islbm_set_step_count.argtypes = [vp, ctypes.c_int64]
    L.islbm_kernel_in_use.argtypes = [vp, ctypes.POINTER(i32)]
    L.islbm_tile_shape.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.islbm_get_table.argtypes = [vp, ctypes.c_char_p, dp, i32]
    L.islbm_time_steps.argtypes = [vp, i32, dp]
    L.islbm_last_error.restype = ctypes.c_char_p
    L.islbm_selftest_collision.argtypes = [dp]
    L.islbm_copy_bandwidth.argtypes = [i32, ctypes.c_size_t, i32, dp]
    L.islbm_ref_init.argtypes = [pp]
    L.islbm_ref_device_arrays.argtypes = [ctypes.POINTER(vp)] * 6
    L.Launch_CollisionStreaming.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.Launch_CollisionStreaming.restype = None
    L.Launch_ModifyForcingTerm.argtypes = []
    L.Launch_ModifyForcingTerm.restype = None
    L.islbm_step_graph.argtypes = [vp, i32]
    L.islbm_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.islbm_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    i64, i64p = ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)
    L.islbm_batch_create.argtypes = [i32, ctypes.c_size_t, i32, i32, ctypes.POINTER(vp)]
    L.islbm_batch_destroy.argtypes = [vp]
    L.islbm_batch_submit.argtypes = [vp, pp, dp, ctypes.c_double, ctypes.c_uint64, i64, i32, i64p]
    L.islbm_batch_advance.argtypes = [vp, i32, i64p]
    L.islbm_batch_status.argtypes = [vp, i64, ctypes.POINTER(i32), dp, i64p]
    L.islbm_batch_result.argtypes = [vp, i64, dp, dp, dp, dp]
    L.islbm_batch_summary.argtypes = [vp, i64, dp]
    L.islbm_batch_bytes_in_use.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    L.islbm_batch_group_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.islbm_batch_last_error.restype = ctypes.c_char_p
    L.islbm_batch_now.restype = ctypes.c_double
    L.islbm_upload_macros.argtypes = [vp, dp, dp, dp, dp]
    L.islbm_get_params.argtypes = [vp, pp]
    L.islbm_mass_correct.argtypes = [vp, ctypes.c_double, dp]
    L.islbm_mass_correct_noalloc.argtypes = [vp, ctypes.c_double, dp]
    L.islbm_monitor.argtypes = [vp, dp]
    L.islbm_stats_enable.argtypes = [vp, i32]
    for nm in ("islbm_stats_disable", "islbm_stats_accumulate", "islbm_stats_reset"):
        getattr(L, nm).argtypes = [vp]
    L.islbm_stats_count.argtypes = [vp, i64p]
    L.islbm_stats_set_count.argtypes = [vp, i64]
    L.islbm_stats_names.argtypes = [ctypes.POINTER(ctypes.c_char_p), i32]
    L.islbm_stats_download.argtypes = [vp, ctypes.c_char_p, dp]
    L.islbm_stats_upload.argtypes = [vp, ctypes.c_char_p, dp]
    for nm in ("islbm_checkpoint_write", "islbm_checkpoint_read", "islbm_stats_write", "islbm_stats_read"):
        getattr(L, nm).argtypes = [vp, ctypes.c_char_p, i32]
    L.islbm_monitor_append.argtypes = [vp, ctypes.c_char_p]
    L.islbm_io_last_error.restype = ctypes.c_char_p
    for name in HEADER_SYMBOLS:
        fn = getattr(L, name)
        if fn.restype is ctypes.c_int and name not in ("islbm_last_error",):
            fn.restype = ctypes.c_int
    _lib = L
    return L


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().islbm_last_error().decode(errors="replace")
        raise ISLBMError(f"{what} failed ({rc}): {msg}")
    return rc


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(**kw) -> Params:
    """The reference's shipped parameter set (variables.h), updated by keyword overrides."""
    p = Params()
    _check(lib().islbm_default_params(ctypes.byref(p)), "islbm_default_params")
    names = {f for f, _ in Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise KeyError(f"unknown parameter {k}")
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(p, k, v)
    return p


def required_bytes(p: Params) -> int:
    b = ctypes.c_size_t()
    _check(lib().islbm_required_bytes(ctypes.byref(p), ctypes.byref(b)), "islbm_required_bytes")
    return b.value


class Solver:
    """One device-resident periodic-hill solve."""

    def __init__(self, params: Params, hbm_budget: int = 0):
        self.p = params
        h = ctypes.c_void_p()
        _check(lib().islbm_create(ctypes.byref(params), int(hbm_budget), ctypes.byref(h)), "islbm_create")
        self.h = h
        self.shape = (params.nz, params.ny, params.nx)
        self.n = params.nx * params.ny * params.nz

    def close(self):
        if getattr(self, "h", None):
            lib().islbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------------------
    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.0, 0.0), amp=0.0, seed=20261203):
        u = np.ascontiguousarray(u0, dtype=np.float64)
        _check(lib().islbm_init_equilibrium(self.h, float(rho0), _dp(u), float(amp), int(seed)), "init")

    def upload_f(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.size != 19 * self.n:
            raise ValueError("population array must hold 19 * nx * ny * nz values")
        _check(lib().islbm_upload_f(self.h, _dp(f)), "islbm_upload_f")

    def f(self) -> np.ndarray:
        out = np.empty((19,) + self.shape)
        _check(lib().islbm_download_f(self.h, _dp(out)), "islbm_download_f")
        return out

    def macros(self):
        r, u, v, w = (np.empty(self.shape) for _ in range(4))
        _check(lib().islbm_download_macros(self.h, _dp(r), _dp(u), _dp(v), _dp(w)), "islbm_download_macros")
        return r, u, v, w

    # -- stepping ------------------------------------------------------------------------
    def step(self, n: int = 1, sync: bool = True):
        _check(lib().islbm_step(self.h, int(n)), "islbm_step")
        if sync:
            self.sync()

    def step_graph(self, n: int, sync: bool = True):
        """n steps (even) replayed from a HIP graph recorded on the first call of this n."""
        _check(lib().islbm_step_graph(self.h, int(n)), "islbm_step_graph")
        if sync:
            self.sync()

    def device_bytes(self) -> int:
        v = ctypes.c_size_t()
        _check(lib().islbm_device_bytes(self.h, ctypes.byref(v)), "islbm_device_bytes")
        return v.value

    def sync(self):
        _check(lib().islbm_sync(self.h), "islbm_sync")

    def modify_forcing(self):
        _check(lib().islbm_modify_forcing(self.h), "islbm_modify_forcing")

    @property
    def force(self) -> float:
        v = ctypes.c_double()
        _check(lib().islbm_get_force(self.h, ctypes.byref(v)), "islbm_get_force")
        return v.value

    @force.setter
    def force(self, val: float):
        _check(lib().islbm_set_force(self.h, float(val)), "islbm_set_force")

    def set_rho_modify(self, v: float):
        _check(lib().islbm_set_rho_modify(self.h, float(v)), "islbm_set_rho_modify")

    def flux_sum(self) -> float:
        v = ctypes.c_double()
        _check(lib().islbm_get_flux_sum(self.h, ctypes.byref(v)), "islbm_get_flux_sum")
        return v.value

    def scalars(self) -> dict:
        o = np.zeros(8)
        _check(lib().islbm_scalars(self.h, _dp(o)), "islbm_scalars")
        return dict(zip(["dt", "tau", "nu", "uref", "a", "dx", "dy", "force0"], o.tolist()))

    def step_count(self) -> int:
        v = ctypes.c_int64()
        _check(lib().islbm_step_count(self.h, ctypes.byref(v)), "islbm_step_count")
        return v.value

    def set_step_count(self, steps: int):
        _check(lib().islbm_set_step_count(self.h, int(steps)), "islbm_set_step_count")

    def kernel_in_use(self) -> int:
        v = ctypes.c_int32()
        _check(lib().islbm_kernel_in_use(self.h, ctypes.byref(v)), "islbm_kernel_in_use")
        return v.value

    def tile_shape(self) -> tuple:
        """(columns, rows) of a tiled-kernel workgroup's tile; (0, 0) for the other kernels."""
        c, r = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().islbm_tile_shape(self.h, ctypes.byref(c), ctypes.byref(r)), "islbm_tile_shape")
        return c.value, r.value

    def table(self, name: str) -> np.ndarray:
        cap = 16
        while True:
            out = np.zeros(cap)
            n = lib().islbm_get_table(self.h, name.encode(), _dp(out), cap)
            if n >= 0:
                return out[:n].copy()
            if n == -1 or -n <= cap:
                _check(n, f"islbm_get_table({name})")
            cap = -n

    # -- restart, monitor, mass correction -------------------------------------------------
    def upload_macros(self, rho=None, u=None, v=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rho, u, v, w)]
        for a in arrs:
            if a is not None and a.size != self.n:
                raise ValueError("macroscopic arrays must hold nx * ny * nz values")
        ptrs = [None if a is None else _dp(a) for a in arrs]
        _check(lib().islbm_upload_macros(self.h, *ptrs), "islbm_upload_macros")

    def mass_correct(self, rho0: float = 1.0) -> float:
        """Sets rho_modify = rho0 - mean(rho); returns the mean density before the correction."""
        m = ctypes.c_double()
        _check(lib().islbm_mass_correct(self.h, float(rho0), ctypes.byref(m)), "islbm_mass_correct")
        return m.value

    def monitor(self):
        """(v, u, w, force) at the reference's monitor node."""
        o = np.zeros(4)
        _check(lib().islbm_monitor(self.h, _dp(o)), "islbm_monitor")
        return tuple(o.tolist())

    def _io(self, rc: int, what: str):
        if rc < 0:
            msg = lib().islbm_io_last_error().decode(errors="replace")
            inner = lib().islbm_last_error().decode(errors="replace")
            raise ISLBMError(f"{what} failed ({rc}): {msg} {inner}".strip())

    def checkpoint_write(self, directory: str, rank: int = 0):
        self._io(lib().islbm_checkpoint_write(self.h, directory.encode(), int(rank)), "islbm_checkpoint_write")

    def checkpoint_read(self, directory: str, rank: int = 0):
        self._io(lib().islbm_checkpoint_read(self.h, directory.encode(), int(rank)), "islbm_checkpoint_read")

    def monitor_append(self, path: str):
        self._io(lib().islbm_monitor_append(self.h, path.encode()), "islbm_monitor_append")

    # -- turbulence statistics -------------------------------------------------------------
    def stats_enable(self, every: int = 1):
        _check(lib().islbm_stats_enable(self.h, int(every)), "islbm_stats_enable")

    def stats_disable(self):
        _check(lib().islbm_stats_disable(self.h), "islbm_stats_disable")

    def stats_accumulate(self):
        _check(lib().islbm_stats_accumulate(self.h), "islbm_stats_accumulate")

    def stats_reset(self):
        _check(lib().islbm_stats_reset(self.h), "islbm_stats_reset")

    def stats_count(self) -> int:
        v = ctypes.c_int64()
        _check(lib().islbm_stats_count(self.h, ctypes.byref(v)), "islbm_stats_count")
        return v.value

    def stats(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        _check(lib().islbm_stats_download(self.h, name.encode(), _dp(out)), f"islbm_stats_download({name})")
        return out

    def stats_write(self, directory: str, rank: int = 0):
        self._io(lib().islbm_stats_write(self.h, directory.encode(), int(rank)), "islbm_stats_write")

    def stats_read(self, directory: str, rank: int = 0):
        self._io(lib().islbm_stats_read(self.h, directory.encode(), int(rank)), "islbm_stats_read")

    def time_steps(self, n: int):
        """(ms for n steps, summed ms of the stream-collide launches) from HIP events."""
        o = np.zeros(2)
        _check(lib().islbm_time_steps(self.h, int(n), _dp(o)), "islbm_time_steps")
        return float(o[0]), float(o[1])


def _bcheck(rc: int, what: str) -> int:
    if rc < 0:
        L = lib()
        msg = L.islbm_batch_last_error().decode(errors="replace")
        inner = L.islbm_last_error().decode(errors="replace")
        raise ISLBMError(f"{what} failed ({rc}): {msg} {inner}".strip())
    return rc


class Batch:
    """Continuous-batching serving loop over many device-resident solves (islbm_batch_*)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_active: int = 8, chunk: int = 50):
        h = ctypes.c_void_p()
        _bcheck(lib().islbm_batch_create(int(device), int(hbm_budget), int(max_active), int(chunk), ctypes.byref(h)),
                "islbm_batch_create")
        self.h = h
        self.params = {}

    def close(self):
        if getattr(self, "h", None):
            lib().islbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def now() -> float:
        return float(lib().islbm_batch_now())

    def submit(self, params: Params, nsteps: int, u0=(0.0, 0.0, 0.0), amp: float = 0.0, seed: int = 20261203,
               keep_fields: bool = False) -> int:
        u = np.ascontiguousarray(u0, dtype=np.float64)
        rid = ctypes.c_int64()
        _bcheck(lib().islbm_batch_submit(self.h, ctypes.byref(params), _dp(u), float(amp), int(seed), int(nsteps),
                                         int(bool(keep_fields)), ctypes.byref(rid)), "islbm_batch_submit")
        self.params[rid.value] = (params.nz, params.ny, params.nx)
        return rid.value

    def advance(self, wait: bool = True):
        """One pass of the loop; returns (queued, active, done, failed)."""
        c = (ctypes.c_int64 * 4)()
        _bcheck(lib().islbm_batch_advance(self.h, int(bool(wait)), c), "islbm_batch_advance")
        return tuple(int(x) for x in c)

    def run(self):
        """Advance until nothing is queued or active."""
        while True:
            q, a, d, f = self.advance(True)
            if q == 0 and a == 0:
                return d, f

    def status(self, rid: int):
        st = ctypes.c_int32()
        t = np.zeros(3)
        steps = ctypes.c_int64()
        _bcheck(lib().islbm_batch_status(self.h, int(rid), ctypes.byref(st), _dp(t), ctypes.byref(steps)),
                "islbm_batch_status")
        return {"state": st.value, "t_submit": t[0], "t_admit": t[1], "t_done": t[2], "steps": steps.value}

    def result(self, rid: int):
        shape = self.params[rid]
        r, u, v, w = (np.empty(shape) for _ in range(4))
        _bcheck(lib().islbm_batch_result(self.h, int(rid), _dp(r), _dp(u), _dp(v), _dp(w)), "islbm_batch_result")
        return r, u, v, w

    def summary(self, rid: int) -> dict:
        o = np.zeros(4)
        _bcheck(lib().islbm_batch_summary(self.h, int(rid), _dp(o)), "islbm_batch_summary")
        return dict(zip(["nodes", "mean_rho", "max_u", "mean_v"], o.tolist()))

    def bytes_in_use(self) -> int:
        v = ctypes.c_size_t()
        _bcheck(lib().islbm_batch_bytes_in_use(self.h, ctypes.byref(v)), "islbm_batch_bytes_in_use")
        return v.value

    def group_stats(self) -> dict:
        """Grouped chunks launched and times their graph was recorded."""
        o = (ctypes.c_int64 * 2)()
        _bcheck(lib().islbm_batch_group_stats(self.h, o), "islbm_batch_group_stats")
        return {"chunks": int(o[0]), "records": int(o[1])}


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20, device: int = 0) -> float:
    """GB/s of a plain device copy kernel (bytes read + written)."""
    v = ctypes.c_double()
    _check(lib().islbm_copy_bandwidth(int(device), int(nbytes), int(iters), ctypes.byref(v)), "islbm_copy_bandwidth")
    return v.value


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.linalg.norm(b.ravel()))
    return float(np.linalg.norm((a - b).ravel())) / (den if den > 0 else 1.0)


def rel_l2_vec(us, vs) -> float:
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(us, vs))
    den = sum(float(np.sum(b ** 2)) for b in vs)
    return (num ** 0.5) / (den ** 0.5 if den > 0 else 1.0)
