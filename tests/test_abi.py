"""Checks of the C ABI: libclv.so loads, exports every entry point that
include/clv.h declares, and its host-only functions behave (no GPU call); on
a GPU, torch and libclv both see the device whichever is used first."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

from cl_var_dt_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "clv.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clv_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_match_binding():
    assert declared_functions() == sorted(_abi.exported_symbols())


@pytest.fixture(scope="module")
def libclv():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libclv.so not built (run __graft_entry__.build())")
    return _abi.lib()


def test_library_exports_every_declared_symbol(libclv):
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (clv_[a-z_0-9]+)", out))
    missing = set(declared_functions()) - exported
    assert not missing, missing


def test_struct_sizes_match_header(libclv):
    # clv_params: 12 doubles + u64 + 4 int32 = 96 + 8 + 16
    assert C.sizeof(_abi.Params) == 120
    assert C.sizeof(_abi.StepStats) == 4 * 8 + 3 * 8 + 8
    assert C.sizeof(_abi.Shape) == 16


def test_default_params_and_host_dt(libclv):
    p = _abi.Params()
    libclv.clv_default_params(C.byref(p))
    assert (p.G, p.eps, p.a_x, p.DT, p.dt_min, p.dt_max, p.alpha) == (0.1, 0.01, 1.0, 0.002, 1e-8, 1e-2, 0.9)
    assert p.kin_dims == 1 and p.noise_on == 1 and p.factor == 1.0
    cl = C.c_int32(7)
    assert libclv.clv_compute_new_dt(0.0, 0.0, C.byref(p), C.byref(cl)) == 0.002 and cl.value == 0
    dt = libclv.clv_compute_new_dt(0.05, 0.5, C.byref(p), C.byref(cl))
    assert cl.value == 0 and dt == (1 - 0.9) * 0.002 + 0.9 * (0.05 / 0.5 * 0.002)
    libclv.clv_compute_new_dt(1e-12, 1.0, C.byref(p), C.byref(cl))
    assert cl.value == -1
    libclv.clv_compute_new_dt(10.0, 1.0, C.byref(p), C.byref(cl))
    assert cl.value == 1


def test_host_dt_matches_oracle(libclv, oracle_lib):
    import numpy as np
    rng = np.random.default_rng(1)
    p = _abi.default_params(DT=1e-3)
    for _ in range(200):
        mx = float(rng.uniform(0, 3))
        mean = float(rng.uniform(0, 1)) * mx / 4
        cl = C.c_int32()
        a = libclv.clv_compute_new_dt(mean, mx, C.byref(p), C.byref(cl))
        b, cb = oracle_lib.new_dt(mean, mx, p)
        assert a == b and cl.value == cb


def test_create_without_gpu_fails_cleanly(libclv):
    import torch  # noqa: F401  (only to see whether a GPU is present)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    shp = (_abi.Shape * 1)(_abi.Shape(8, 1, 1, 4))
    h = C.c_void_p()
    st = libclv.clv_create(shp, 1, _abi.CLV_FP64, 0, C.byref(h))
    assert st == _abi.CLV_ERR_HIP
    assert b"no HIP device" in libclv.clv_last_error(None)


def test_create_rejects_bad_shapes(libclv):
    for bad in [_abi.Shape(0, 1, 1, 4), _abi.Shape(10, 1, 1, 4), _abi.Shape(8, 1, 1, 1)]:
        shp = (_abi.Shape * 1)(bad)
        h = C.c_void_p()
        st = libclv.clv_create(shp, 1, _abi.CLV_FP64, 0, C.byref(h))
        assert st == _abi.CLV_ERR_ARG, bad
    # 2^31 sites per chain: the 32-bit site counter of the noise would wrap
    shp = (_abi.Shape * 1)(_abi.Shape(256, 1024, 1024, 8))
    h = C.c_void_p()
    assert libclv.clv_create(shp, 1, _abi.CLV_FP32, 0, C.byref(h)) == _abi.CLV_ERR_ARG
    assert b"too large" in libclv.clv_last_error(None)
    shp = (_abi.Shape * 1)(_abi.Shape(8, 1, 1, 4))
    assert libclv.clv_create(shp, 1, 7, 0, C.byref(h)) == _abi.CLV_ERR_ARG
    assert b"precision" in libclv.clv_last_error(None)


_ORDER = r'''
import ctypes as C, sys
sys.path.insert(0, {root!r})
from cl_var_dt_amd import _abi
def create():
    shp = (_abi.Shape * 1)(_abi.Shape(8, 1, 1, 4)); h = C.c_void_p()
    st = _abi.lib().clv_create(shp, 1, _abi.CLV_FP64, 0, C.byref(h))
    if st == 0:
        _abi.lib().clv_destroy(h)
    return st
if {case!r} == "lib-first":
    _abi.lib()
    import torch
    ok = torch.cuda.is_available() and float(torch.ones(4, device="cuda").sum()) == 4.0
    print(int(ok), create())
else:
    st = create()
    import torch
    ok = torch.cuda.is_available() and float(torch.ones(4, device="cuda").sum()) == 4.0
    print(int(ok), st)
'''


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["lib-first", "create-first"])
def test_torch_and_libclv_share_the_gpu_in_any_order(case):
    """A Python process that uses both torch (ROCm wheel, its own HIP/HSA
    runtimes) and libclv sees the GPU from both, whichever it touches first:
    loading libclv first used to leave the second runtime without a device."""
    print("case: runtime order", case, flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", _ORDER.format(root=root, case=case)],
                         capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    torch_ok, st = out.stdout.split()[-2:]
    assert torch_ok == "1" and st == "0", out.stdout
