"""ctypes binding of the FP64 CPU oracle (oracle/islbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.  The
product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ISLBM_ORACLE_LIB names another build of the oracle (the mutation tests load planted-mistake
# builds through it)
LIB_PATH = os.environ.get("ISLBM_ORACLE_LIB") or os.path.join(HERE, "_build", "libislbm_oracle.so")
REF_GEOM = os.path.join(HERE, "_ref", "ref_geom")


class OrParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("cfl", ctypes.c_double), ("lxi", ctypes.c_double), ("hill_alpha", ctypes.c_double),
        ("flat", ctypes.c_int), ("mesh_mode", ctypes.c_int), ("stretch", ctypes.c_double),
        ("tau", ctypes.c_double), ("re", ctypes.c_double), ("u_ref", ctypes.c_double),
        ("cs_smag", ctypes.c_double), ("ndtfrc", ctypes.c_int), ("xi_near_wall", ctypes.c_int),
    ]


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference driver when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.POINTER(OrParams)]
        L.or_destroy.argtypes = [vp]
        L.or_threads.argtypes = [ip]
        L.or_threads.restype = ip
        L.or_scalars.argtypes = [vp, dp]
        L.or_init_equilibrium.argtypes = [vp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_uint64]
        L.or_set_f.argtypes = [vp, dp]
        L.or_get_f.argtypes = [vp, dp]
        L.or_get_macros.argtypes = [vp, dp, dp, dp, dp]
        L.or_step.argtypes = [vp, ip]
        L.or_modify_forcing.argtypes = [vp]
        L.or_get_force.argtypes = [vp]
        L.or_get_force.restype = ctypes.c_double
        L.or_set_force.argtypes = [vp, ctypes.c_double]
        L.or_set_rho_modify.argtypes = [vp, ctypes.c_double]
        L.or_flux_sum.argtypes = [vp]
        L.or_flux_sum.restype = ctypes.c_double
        L.or_step_count.argtypes = [vp]
        L.or_step_count.restype = ctypes.c_long
        L.or_get_table.argtypes = [vp, ctypes.c_char_p, dp, ip]
        L.or_get_table.restype = ip
        L.or_hill.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ip]
        L.or_hill.restype = ctypes.c_double
        L.or_lagrange6.argtypes = [ctypes.c_double, dp, ip]
        L.or_lagrange6.restype = ctypes.c_double
        L.or_bisect_a.argtypes = [ctypes.c_double, ctypes.c_double, ip, ctypes.c_double]
        L.or_bisect_a.restype = ctypes.c_double
        L.or_delta_horizontal.argtypes = [ctypes.c_double] * 7
        L.or_delta_horizontal.restype = ctypes.c_double
        L.or_delta_45.argtypes = [ctypes.c_double] * 7
        L.or_delta_45.restype = ctypes.c_double
        L.or_tanh_map.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ip, ip]
        L.or_tanh_map.restype = ctypes.c_double
        L.or_mrt_matrices.argtypes = [dp, dp]
        L.or_stats_enable.argtypes = [vp]
        L.or_stats_accumulate.argtypes = [vp]
        L.or_stats_count.argtypes = [vp]
        L.or_stats_count.restype = ctypes.c_long
        L.or_stats_get.argtypes = [vp, ctypes.c_char_p, dp]
        L.or_stats_get.restype = ip
        L.or_set_macros.argtypes = [vp, dp, dp, dp, dp]
        L.or_stream_only.argtypes = [vp]
        L.or_collide_one.argtypes = [vp, dp, ip, ip, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(**kw) -> OrParams:
    """Laminar periodic-hill workload, 32 x 64 x 64 (spanwise x streamwise x wall-normal)."""
    p = dict(nx=32, ny=64, nz=64, lx=4.5, ly=9.0, lz=3.036, cfl=0.6, lxi=10.0, hill_alpha=1.0,
             flat=0, mesh_mode=1, stretch=2.0, tau=0.0, re=200.0, u_ref=0.05, cs_smag=0.0,
             ndtfrc=10000, xi_near_wall=0)
    p.update(kw)
    return OrParams(**p)


class Oracle:
    """FP64 CPU restatement of the reference step on one single-rank padded domain."""

    def __init__(self, params: OrParams):
        self.p = params
        self.h = lib().or_create(ctypes.byref(params))
        self.n = params.nx * params.ny * params.nz

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalars(self) -> dict:
        o = np.zeros(8)
        lib().or_scalars(self.h, _dp(o))
        return dict(zip(["dt", "tau", "nu", "uref", "a", "dx", "dy", "force"], o.tolist()))

    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.05, 0.0), amp=1e-3, seed=20261203):
        u = np.ascontiguousarray(u0, dtype=np.float64)
        lib().or_init_equilibrium(self.h, rho0, _dp(u), amp, seed)

    def set_f(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == 19 * self.n
        lib().or_set_f(self.h, _dp(f))

    def get_f(self) -> np.ndarray:
        f = np.zeros((19, self.p.nz, self.p.ny, self.p.nx))
        lib().or_get_f(self.h, _dp(f))
        return f

    # statistics (statistics.h:122-224)
    def stats_enable(self):
        lib().or_stats_enable(self.h)

    def stats_accumulate(self):
        lib().or_stats_accumulate(self.h)

    def stats_count(self) -> int:
        return int(lib().or_stats_count(self.h))

    def stats(self, name: str) -> np.ndarray:
        p = self.p
        out = np.empty((p.nz, p.ny, p.nx))
        if lib().or_stats_get(self.h, name.encode(), _dp(out)) != 0:
            raise KeyError(name)
        return out

    def set_macros(self, rho, u, v, w):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rho, u, v, w)]
        lib().or_set_macros(self.h, *[_dp(a) for a in arrs])

    def macros(self):
        shape = (self.p.nz, self.p.ny, self.p.nx)
        r, u, v, w = (np.zeros(shape) for _ in range(4))
        lib().or_get_macros(self.h, _dp(r), _dp(u), _dp(v), _dp(w))
        return r, u, v, w

    def step(self, n: int):
        lib().or_step(self.h, int(n))

    def stream_only(self):
        """Interpolated streaming + bounce-back + BFL of one step, no collision (test entry)."""
        lib().or_stream_only(self.h)

    def collide_one(self, f_in, j: int, k: int):
        """Collision of one node (j, k) on f_in[19]: (f_out[19], [rho, u, v, w, tau_eff])."""
        fi = np.ascontiguousarray(f_in, dtype=np.float64)
        fo = np.zeros(19)
        mac = np.zeros(5)
        lib().or_collide_one(self.h, _dp(fi), int(j), int(k), _dp(fo), _dp(mac))
        return fo, mac

    def modify_forcing(self):
        lib().or_modify_forcing(self.h)

    @property
    def force(self) -> float:
        return lib().or_get_force(self.h)

    @force.setter
    def force(self, v: float):
        lib().or_set_force(self.h, float(v))

    def set_rho_modify(self, v: float):
        lib().or_set_rho_modify(self.h, float(v))

    def flux_sum(self) -> float:
        return lib().or_flux_sum(self.h)

    def table(self, name: str) -> np.ndarray:
        cap = 7 * 2 * (self.p.ny + 8) * (self.p.nz + 8) + 64
        out = np.zeros(cap)
        n = lib().or_get_table(self.h, name.encode(), _dp(out), cap)
        if n <= 0:
            raise KeyError(name)
        return out[:n].copy()


def threads(n: int) -> int:
    return lib().or_threads(int(n))


def hill(y: float, ly: float = 9.0, alpha: float = 1.0, flat: int = 0) -> float:
    return lib().or_hill(float(y), float(ly), float(alpha), int(flat))


def mrt_matrices():
    M = np.zeros((19, 19))
    Mi = np.zeros((19, 19))
    lib().or_mrt_matrices(_dp(M), _dp(Mi))
    return M, Mi
