"""Device memory for GPU tests through the HIP runtime libclv.so links
(/opt/rocm's libamdhip64), so test buffers live in the same runtime as the
library.  torch ships its own HIP runtime: in a process where libclv has
initialised the device first, torch sees no GPU, so the tests do not use it
for buffers."""
from __future__ import annotations

import ctypes as C

import numpy as np

_rt = None
H2D, D2H = 1, 2


def rt():
    global _rt
    if _rt is None:
        from cl_var_dt_amd import _abi
        _abi.lib()  # loads libamdhip64 as a dependency of libclv
        _rt = C.CDLL("libamdhip64.so.7")
        _rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _rt.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _rt.hipSetDevice.argtypes = [C.c_int]
    return _rt


def _ok(st, what):
    if st != 0:
        raise RuntimeError(f"{what} failed with hipError {st}")


def mem_get_info():
    free, total = C.c_size_t(), C.c_size_t()
    _ok(rt().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """A device allocation of `nbytes`, freed by close()."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _ok(rt().hipMalloc(C.byref(p), self.nbytes), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        _ok(rt().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        _ok(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def close(self) -> None:
        if self.ptr:
            _ok(rt().hipFree(self.ptr), "hipFree")
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
