"""Python mirror of the engine's C ABI (include/clv.h) over ctypes.

`Engine` owns one clv context: a set of chains (possibly of different
shapes), each with its own parameters, fields, controller state and
observable accumulators.  Fields cross the boundary as complex128 arrays of
shape (4, Nt, X, Y, Z) -- the reference's device layout (Lattice.cu:2-5).
Every call goes to libclv.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import Observables, Params, Shape, StepStats, check, lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Engine:
    """A batch of complex-Langevin chains on one GPU."""

    def __init__(self, shapes: Sequence[Shape] | Shape, precision: int = _abi.CLV_FP64,
                 device: int = 0):
        if isinstance(shapes, Shape):
            shapes = [shapes]
        self.shapes = [Shape(s.nx, s.ny, s.nz, s.nt) for s in shapes]
        arr = (Shape * len(self.shapes))(*self.shapes)
        h = C.c_void_p()
        st = lib().clv_create(arr, len(self.shapes), precision, device, C.byref(h))
        if st != _abi.CLV_OK:
            raise _abi.ClvError(st, lib().clv_last_error(None).decode())
        self._h = h
        self.precision = precision
        self.params = [_abi.default_params(seed=i) for i in range(len(self.shapes))]

    # ---- lifetime ----
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().clv_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, status: int) -> None:
        check(status, self._h)

    @property
    def nchains(self) -> int:
        return len(self.shapes)

    def device_bytes(self) -> int:
        return int(lib().clv_device_bytes(self._h))

    def dims(self, chain: int):
        s = self.shapes[chain]
        return (4, s.nt, s.nx, s.ny, s.nz)

    # ---- parameters and fields ----
    def set_params(self, chain: int, p: Params) -> None:
        self._chk(lib().clv_set_params(self._h, chain, C.byref(p)))
        self.params[chain] = p

    def set_fields(self, chain: int, psi: np.ndarray, psib: np.ndarray) -> None:
        a = np.ascontiguousarray(psi, dtype=np.complex128).reshape(self.dims(chain))
        b = np.ascontiguousarray(psib, dtype=np.complex128).reshape(self.dims(chain))
        self._chk(lib().clv_set_fields(self._h, chain, _dp(a), _dp(b)))

    def get_fields(self, chain: int):
        a = np.zeros(self.dims(chain), np.complex128)
        b = np.zeros(self.dims(chain), np.complex128)
        self._chk(lib().clv_get_fields(self._h, chain, _dp(a), _dp(b)))
        return a, b

    def init_fields(self, chain: int, n0: float = 1.0, var: float = 0.1) -> None:
        self._chk(lib().clv_init_fields(self._h, chain, n0, var))

    # ---- stepping ----
    def propagate(self, nsteps: int, eval_every: int = 0) -> None:
        self._chk(lib().clv_propagate(self._h, nsteps, eval_every))

    def propagate_plain(self, chain: int, nsteps: int, eval_every: int = 0) -> None:
        self._chk(lib().clv_propagate_plain(self._h, chain, nsteps, eval_every))

    def drift_stats(self, chain: int):
        m, s = C.c_double(), C.c_double()
        self._chk(lib().clv_drift_stats(self._h, chain, C.byref(m), C.byref(s)))
        return m.value, s.value

    def apply_update(self, chain: int, dt: float) -> None:
        self._chk(lib().clv_apply_update(self._h, chain, dt))

    def stats(self, chain: int) -> StepStats:
        s = StepStats()
        self._chk(lib().clv_get_step_stats(self._h, chain, C.byref(s)))
        return s

    def set_stats(self, chain: int, st: StepStats) -> None:
        """Restore the controller state (exact resume together with set_fields)."""
        self._chk(lib().clv_set_step_stats(self._h, chain, C.byref(st)))

    def synchronize(self) -> None:
        self._chk(lib().clv_synchronize(self._h))

    # ---- observables ----
    def evaluate(self, chain: int = -1) -> None:
        self._chk(lib().clv_evaluate(self._h, chain))

    def observables(self, chain: int):
        o = Observables()
        s = self.shapes[chain]
        vs = s.nx * s.ny * s.nz
        F = [np.zeros(vs) for _ in range(3)]
        self._chk(lib().clv_read_observables(self._h, chain, C.byref(o), _dp(F[0]), _dp(F[1]),
                                             _dp(F[2])))
        return o, F

    def reset_observables(self, chain: int = -1) -> None:
        self._chk(lib().clv_reset_observables(self._h, chain))

    # ---- kernel schedule ----
    def autotune(self, cache_path: str | None = None) -> None:
        self._chk(lib().clv_autotune(self._h, cache_path.encode() if cache_path else None))

    def set_schedule(self, schedule: int) -> None:
        self._chk(lib().clv_set_schedule(self._h, schedule))

    def schedule(self, chain: int = 0) -> int:
        return int(lib().clv_get_schedule(self._h, chain))

    def set_row_split(self, chain: int, n1: int) -> None:
        """Row-transform split N1 (X = N1 * N2) of the chain's shape; 0 = default."""
        self._chk(lib().clv_set_row_split(self._h, chain, n1))

    def row_split(self, chain: int = 0) -> int:
        return int(lib().clv_get_row_split(self._h, chain))

    def set_fused_options(self, line_sets: int = 0, batch: int = 0) -> None:
        """Fused schedule: LDS line sets (1, 2) and K_0 batch (1, 2, 4); 0 keeps."""
        self._chk(lib().clv_set_fused_options(self._h, line_sets, batch))

    def set_fused_blocks(self, blocks: int = 0) -> None:
        """Fused schedule: tau blocks per (row tile, chain); 0 = untimed choice."""
        self._chk(lib().clv_set_fused_blocks(self._h, blocks))

    def set_graphs(self, on: bool) -> None:
        self._chk(lib().clv_set_graphs(self._h, 1 if on else 0))

    def prepare_graph(self, eval_every: int = 0) -> None:
        """Capture the step graph the next propagate(n, eval_every) replays,
        without running it (keeps the one-time capture out of a timed region)."""
        self._chk(lib().clv_prepare_graph(self._h, eval_every))

    # ---- kernel timing (HIP events on the context's stream) ----
    def set_profiling(self, on: bool) -> None:
        self._chk(lib().clv_set_profiling(self._h, 1 if on else 0))

    def kernel_time(self, kind: str):
        ms, n = C.c_double(), C.c_int64()
        self._chk(lib().clv_kernel_time(self._h, _abi.KERNEL_KINDS[kind], C.byref(ms), C.byref(n)))
        return ms.value, n.value


def compute_new_dt(mean: float, mx: float, p: Params):
    """clv_compute_new_dt: the reference's controller on the host."""
    cl = C.c_int32(0)
    dt = lib().clv_compute_new_dt(mean, mx, C.byref(p), C.byref(cl))
    return dt, cl.value
