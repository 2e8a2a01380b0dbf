"""ctypes mirror of include/clv.h and the loader of libclv.so.

The structures here are byte-for-byte the C structs of clv.h.  The loader looks
for the in-tree build (cl_var_dt_amd/libclv.so) and raises if it is missing:
there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

CLV_OK = 0
CLV_ERR_ARG = -1
CLV_ERR_HIP = -2
CLV_ERR_NOMEM = -3
CLV_ERR_BUDGET = -4
CLV_ERR_STATE = -5
CLV_ERR_NOTFOUND = -6
CLV_ERR_CANCELLED = -7

CLV_FP64 = 0
CLV_FP32 = 1
CLV_CONJ_DRIFT_REFERENCE = 0
CLV_CONJ_DRIFT_ACTION = 1

SCHED_AUTO = 0
SCHED_TWO_PASS = 1
SCHED_FUSED = 2

KERNEL_KINDS = {"drift_stats": 0, "controller": 1, "update": 2, "obs": 3, "obs_finish": 4,
                "fused": 5, "transverse": 6}


class Shape(C.Structure):
    """clv_shape: one chain's X x Y x Z x Nt lattice (Parameters.h:8-11)."""

    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nt", C.c_int32)]

    @property
    def vs(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def sites(self) -> int:
        return self.vs * self.nt

    def __repr__(self) -> str:  # pragma: no cover - debugging aid
        return f"Shape({self.nx}x{self.ny}x{self.nz}x{self.nt})"


class Params(C.Structure):
    """clv_params (see include/clv.h for the meaning of each field)."""

    _fields_ = [
        ("mu", C.c_double), ("P", C.c_double), ("G", C.c_double), ("c1", C.c_double),
        ("Q", C.c_double), ("eps", C.c_double), ("a_x", C.c_double), ("DT", C.c_double),
        ("dt_min", C.c_double), ("dt_max", C.c_double), ("alpha", C.c_double),
        ("factor", C.c_double), ("seed", C.c_uint64), ("kin_dims", C.c_int32),
        ("conj_drift_mode", C.c_int32), ("noise_on", C.c_int32), ("reserved", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("max_drift", C.c_double), ("mean_drift", C.c_double),
        ("t_langevin", C.c_double), ("steps", C.c_int64), ("clamp_low", C.c_int64),
        ("clamp_high", C.c_int64), ("nonfinite", C.c_int32), ("reserved", C.c_int32),
    ]


class Observables(C.Structure):
    _fields_ = [
        ("count", C.c_int64), ("n_tot", C.c_double), ("n_comp", C.c_double * 4),
        ("energy", C.c_double), ("Fx_tot", C.c_double), ("Fy_tot", C.c_double),
        ("Fz_tot", C.c_double), ("density", C.c_double), ("Sz", C.c_double),
    ]


class Request(C.Structure):
    """clv_request: one Langevin run submitted to a server."""

    _fields_ = [("shape", Shape), ("params", Params), ("nsteps", C.c_int64),
                ("eval_every", C.c_int32), ("reserved", C.c_int32), ("eval_from", C.c_int64),
                ("n0", C.c_double), ("var", C.c_double)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("reserved", C.c_int32),
                ("obs", Observables), ("stats", StepStats), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


class ServerInfo(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("running", C.c_int64),
                ("queued", C.c_int64), ("failed", C.c_int64), ("bytes_in_use", C.c_int64),
                ("budget_bytes", C.c_int64), ("pools", C.c_int64), ("site_updates", C.c_double),
                ("busy_seconds", C.c_double), ("cancelled", C.c_int64),
                ("pool_setup_seconds", C.c_double)]


CLV_RUN_QUEUED = 0
CLV_RUN_RUNNING = 1
CLV_RUN_DONE = 2


def default_params(**kw) -> Params:
    """Reference defaults (Parameters.h:13-31, Propagator_without_cutoff.cu:391-402)."""
    p = Params(mu=0.0, P=0.0, G=0.1, c1=0.0, Q=0.0, eps=0.01, a_x=1.0, DT=0.002,
               dt_min=1e-8, dt_max=1e-2, alpha=0.9, factor=1.0, seed=0,
               kin_dims=1, conj_drift_mode=CLV_CONJ_DRIFT_REFERENCE, noise_on=1, reserved=0)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"clv_params has no field {k!r}")
        setattr(p, k, v)
    return p


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libclv.so")

_lib = None


class ClvError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"clv error {status}: {msg}")
        self.status = status


def _declare(lib):
    P = C.POINTER
    vp = C.c_void_p
    i32, i64, dbl = C.c_int32, C.c_int64, C.c_double
    sig = {
        "clv_version": (C.c_char_p, []),
        "clv_last_error": (C.c_char_p, [vp]),
        "clv_default_params": (None, [P(Params)]),
        "clv_compute_new_dt": (dbl, [dbl, dbl, P(Params), P(i32)]),
        "clv_create": (C.c_int, [P(Shape), i32, i32, i32, P(vp)]),
        "clv_destroy": (C.c_int, [vp]),
        "clv_num_chains": (C.c_int, [vp]),
        "clv_device_bytes": (i64, [vp]),
        "clv_set_params": (C.c_int, [vp, i32, P(Params)]),
        "clv_set_fields": (C.c_int, [vp, i32, P(dbl), P(dbl)]),
        "clv_get_fields": (C.c_int, [vp, i32, P(dbl), P(dbl)]),
        "clv_init_fields": (C.c_int, [vp, i32, dbl, dbl]),
        "clv_propagate": (C.c_int, [vp, i64, i32]),
        "clv_propagate_plain": (C.c_int, [vp, i32, i64, i32]),
        "clv_drift_stats": (C.c_int, [vp, i32, P(dbl), P(dbl)]),
        "clv_apply_update": (C.c_int, [vp, i32, dbl]),
        "clv_get_step_stats": (C.c_int, [vp, i32, P(StepStats)]),
        "clv_set_step_stats": (C.c_int, [vp, i32, P(StepStats)]),
        "clv_synchronize": (C.c_int, [vp]),
        "clv_evaluate": (C.c_int, [vp, i32]),
        "clv_read_observables": (C.c_int, [vp, i32, P(Observables), P(dbl), P(dbl), P(dbl)]),
        "clv_reset_observables": (C.c_int, [vp, i32]),
        "clv_autotune": (C.c_int, [vp, C.c_char_p]),
        "clv_set_schedule": (C.c_int, [vp, i32]),
        "clv_get_schedule": (C.c_int, [vp, i32]),
        "clv_set_row_split": (C.c_int, [vp, i32, i32]),
        "clv_get_row_split": (C.c_int, [vp, i32]),
        "clv_set_fused_options": (C.c_int, [vp, i32, i32]),
        "clv_set_fused_blocks": (C.c_int, [vp, i32]),
        "clv_set_graphs": (C.c_int, [vp, i32]),
        "clv_prepare_graph": (C.c_int, [vp, i32]),
        "clv_set_profiling": (C.c_int, [vp, i32]),
        "clv_kernel_time": (C.c_int, [vp, i32, P(dbl), P(i64)]),
        "clv_step_noalloc": (C.c_int, [P(Shape), i32, P(Params), vp, vp, vp, vp, P(StepStats), i32, vp]),
        "clv_server_create": (C.c_int, [i32, i32, i64, P(vp)]),
        "clv_server_destroy": (C.c_int, [vp]),
        "clv_server_submit": (C.c_int, [vp, P(Request), P(i64)]),
        "clv_server_poll": (C.c_int, [vp, i64, P(Result)]),
        "clv_server_wait": (C.c_int, [vp, i64, dbl, P(Result)]),
        "clv_server_release": (C.c_int, [vp, i64]),
        "clv_server_cancel": (C.c_int, [vp, i64]),
        "clv_server_info_get": (C.c_int, [vp, P(ServerInfo)]),
        "clv_server_clock": (dbl, []),
        "clv_server_error": (C.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def exported_symbols() -> list[str]:
    """Every entry point that include/clv.h declares (checked by the tests)."""
    return [
        "clv_version", "clv_last_error", "clv_default_params", "clv_compute_new_dt",
        "clv_create", "clv_destroy", "clv_num_chains", "clv_device_bytes",
        "clv_set_params", "clv_set_fields", "clv_get_fields", "clv_init_fields",
        "clv_propagate", "clv_propagate_plain", "clv_drift_stats", "clv_apply_update",
        "clv_get_step_stats", "clv_set_step_stats", "clv_synchronize", "clv_evaluate", "clv_read_observables",
        "clv_reset_observables", "clv_autotune", "clv_set_schedule", "clv_get_schedule",
        "clv_set_row_split", "clv_get_row_split",
        "clv_set_fused_options", "clv_set_fused_blocks", "clv_set_profiling", "clv_kernel_time", "clv_set_graphs", "clv_prepare_graph", "clv_step_noalloc",
        "clv_server_create", "clv_server_destroy", "clv_server_submit", "clv_server_poll",
        "clv_server_wait", "clv_server_release", "clv_server_cancel", "clv_server_info_get", "clv_server_clock", "clv_server_error",
    ]


def lib():
    """Load libclv.so (in-tree build).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python __graft_entry__.py` "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        # PyTorch's ROCm wheels bring their own HIP and HSA runtimes.  A process
        # may hold only one working HSA runtime: when torch is loaded first,
        # libclv's HIP runtime binds to torch's already-loaded HSA runtime (the
        # same soname), and both see the GPU; when libclv is loaded first, the
        # two HSA runtimes differ and whichever initialises second finds no
        # device.  So torch, where installed, is loaded before libclv.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def check(status: int, ctx=None) -> None:
    if status != CLV_OK:
        msg = lib().clv_last_error(ctx).decode()
        raise ClvError(status, msg)
