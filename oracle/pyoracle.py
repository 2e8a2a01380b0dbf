"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
oracle restates the reference update (Propagator_without_cutoff.cu:1-477,
Lattice.cu:2-214, Observables_without_cutoff.cu:35-101, 456-488, 1155-1634)
in plain C++; see clv_oracle.cpp for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from cl_var_dt_amd._abi import Observables, Params, Shape, StepStats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        vp, dbl, i32, i64, u32, u64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
        sig = {
            "orc_philox4x32_10": (None, [P(u32), P(u32), P(u32)]),
            "orc_normals": (None, [u64, u32, u64, u32, u32, P(dbl)]),
            "orc_fft1d": (None, [P(dbl), i32, i32]),
            "orc_fft3d": (None, [P(dbl), i32, i32, i32, i32]),
            "orc_new_dt": (dbl, [dbl, dbl, P(Params), P(i32)]),
            "orc_create": (vp, [P(Shape), P(Params), i32]),
            "orc_destroy": (None, [vp]),
            "orc_set_params": (None, [vp, P(Params)]),
            "orc_set_fields": (None, [vp, P(dbl), P(dbl)]),
            "orc_get_fields": (None, [vp, P(dbl), P(dbl)]),
            "orc_init_fields": (None, [vp, dbl, dbl]),
            "orc_kinetic": (None, [vp, P(dbl), P(dbl)]),
            "orc_drift": (None, [vp, P(dbl), P(dbl)]),
            "orc_max_mean": (None, [vp, P(dbl), P(dbl)]),
            "orc_propagate": (None, [vp, i64, i32]),
            "orc_apply_update": (None, [vp, dbl]),
            "orc_get_step_stats": (None, [vp, P(StepStats)]),
            "orc_evaluate": (None, [vp]),
            "orc_read_observables": (None, [vp, P(Observables), P(dbl), P(dbl), P(dbl)]),
            "orc_reset_observables": (None, [vp]),
            "orc_observable_sums": (None, [vp, P(dbl)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def normals(seed: int, stream: int, step: int, site: int, pair: int) -> np.ndarray:
    o = np.zeros(4)
    lib().orc_normals(seed, stream, step, site, pair, _dp(o))
    return o


def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(_dp(a), a.size, sign)
    return a


def fft3d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    nx, ny, nz = a.shape
    lib().orc_fft3d(_dp(a), nx, ny, nz, sign)
    return a


def new_dt(mean: float, mx: float, p: Params):
    cl = C.c_int32(0)
    dt = lib().orc_new_dt(mean, mx, C.byref(p), C.byref(cl))
    return dt, cl.value


class Oracle:
    """One chain on the CPU; fields as complex128 arrays of shape (4, Nt, X, Y, Z)."""

    def __init__(self, shape: Shape, params: Params, nthreads: int = 0):
        self.shape = Shape(shape.nx, shape.ny, shape.nz, shape.nt)
        self.params = params
        self._h = lib().orc_create(C.byref(self.shape), C.byref(params), nthreads)
        self.dims = (4, shape.nt, shape.nx, shape.ny, shape.nz)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: Params):
        self.params = p
        lib().orc_set_params(self._h, C.byref(p))

    def set_fields(self, psi: np.ndarray, psib: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.complex128).reshape(self.dims)
        b = np.ascontiguousarray(psib, dtype=np.complex128).reshape(self.dims)
        lib().orc_set_fields(self._h, _dp(a), _dp(b))

    def get_fields(self):
        a = np.zeros(self.dims, np.complex128)
        b = np.zeros(self.dims, np.complex128)
        lib().orc_get_fields(self._h, _dp(a), _dp(b))
        return a, b

    def init_fields(self, n0: float = 1.0, var: float = 0.1):
        lib().orc_init_fields(self._h, n0, var)

    def kinetic(self):
        a = np.zeros(self.dims, np.complex128)
        b = np.zeros(self.dims, np.complex128)
        lib().orc_kinetic(self._h, _dp(a), _dp(b))
        return a, b

    def drift(self):
        a = np.zeros(self.dims, np.complex128)
        b = np.zeros(self.dims, np.complex128)
        lib().orc_drift(self._h, _dp(a), _dp(b))
        return a, b

    def max_mean(self):
        m, s = C.c_double(), C.c_double()
        lib().orc_max_mean(self._h, C.byref(m), C.byref(s))
        return m.value, s.value

    def propagate(self, nsteps: int, eval_every: int = 0):
        lib().orc_propagate(self._h, nsteps, eval_every)

    def apply_update(self, dt: float):
        lib().orc_apply_update(self._h, dt)

    def stats(self) -> StepStats:
        s = StepStats()
        lib().orc_get_step_stats(self._h, C.byref(s))
        return s

    def evaluate(self):
        lib().orc_evaluate(self._h)

    def observables(self):
        o = Observables()
        vs = self.shape.nx * self.shape.ny * self.shape.nz
        F = [np.zeros(vs) for _ in range(3)]
        lib().orc_read_observables(self._h, C.byref(o), _dp(F[0]), _dp(F[1]), _dp(F[2]))
        return o, F

    def reset_observables(self):
        lib().orc_reset_observables(self._h)

    def observable_sums(self) -> np.ndarray:
        o = np.zeros(9)
        lib().orc_observable_sums(self._h, _dp(o))
        return o
