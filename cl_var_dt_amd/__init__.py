"""cl_var_dt_amd -- MI355X-native complex-Langevin engine (host side).

The compute path is libclv.so (hand-written gfx950 HIP kernels behind the C
ABI of include/clv.h); this package binds it with ctypes and mirrors the
reference's Lattice / Propagator / Observables roles for Python callers.
"""
from ._abi import (CLV_FP32, CLV_FP64, CLV_CONJ_DRIFT_ACTION, CLV_CONJ_DRIFT_REFERENCE, ClvError,  # noqa: F401
                   Observables, Params, Shape, StepStats, default_params)

__all__ = ["CLV_FP32", "CLV_FP64", "CLV_CONJ_DRIFT_ACTION", "CLV_CONJ_DRIFT_REFERENCE", "ClvError",
           "Observables", "Params", "Shape", "StepStats", "default_params"]
