"""clv_step_noalloc, the allocation-free step (GPU): one step of one chain in
caller-owned device buffers, with the parameters, controller state, row
tables and partials in the library's static device scratch array.  Checked
bit for bit against clv_propagate on a context (same kernels, same tables),
on two calls in a row -- the second on a smaller lattice, so the scratch keeps
the first call's larger contents -- in FP32 and FP64; shapes whose partials do
not fit the scratch are rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, Shape, default_params
from cl_var_dt_amd import _abi

pytestmark = pytest.mark.gpu


def _to_internal(f):
    # reference layout (c, t, x, y, z) -> internal (t, y, z, x, c)
    return np.ascontiguousarray(f.transpose(1, 3, 4, 2, 0))


def _to_reference(g):
    return np.ascontiguousarray(g.transpose(4, 0, 3, 1, 2))


def _params(seed, nt):
    return default_params(mu=0.9, G=0.1, c1=-0.05, Q=0.1, P=0.02, eps=1.0 / nt, DT=1e-3, seed=seed)


def _context_run(shape, prec, p, nsteps):
    from cl_var_dt_amd.engine import Engine
    e = Engine([shape], precision=prec)
    e.set_params(0, p)
    e.init_fields(0, 1.0, 0.1)
    start = e.get_fields(0)
    e.propagate(nsteps, 0)
    out = e.get_fields(0), e.stats(0)
    e.close()
    return start, out


def _noalloc_run(shape, prec, p, start, nsteps):
    from tests import hip_rt
    dtype = np.complex64 if prec == CLV_FP32 else np.complex128
    dims = (shape.nt, shape.ny, shape.nz, shape.nx, 4)
    nbytes = int(np.prod(dims)) * np.dtype(dtype).itemsize
    bufs = [hip_rt.DeviceBuffer(nbytes) for _ in range(4)]
    try:
        bufs[0].upload(_to_internal(start[0]).astype(dtype))
        bufs[1].upload(_to_internal(start[1]).astype(dtype))
        st = _abi.StepStats()
        cur, nxt = (0, 1), (2, 3)
        for _ in range(nsteps):
            rc = _abi.lib().clv_step_noalloc(C.byref(shape), prec, C.byref(p), bufs[cur[0]].ptr,
                                             bufs[cur[1]].ptr, bufs[nxt[0]].ptr, bufs[nxt[1]].ptr,
                                             C.byref(st), 0, None)
            _abi.check(rc)
            cur, nxt = nxt, cur
        a = _to_reference(bufs[cur[0]].download(dtype, dims)).astype(np.complex128)
        b = _to_reference(bufs[cur[1]].download(dtype, dims)).astype(np.complex128)
    finally:
        for buf in bufs:
            buf.close()
    return (a, b), st


@pytest.mark.parametrize("prec", [CLV_FP32, CLV_FP64])
def test_noalloc_step_equals_propagate(prec):
    print("case: clv_step_noalloc vs clv_propagate", prec, flush=True)
    for shape, nsteps, seed in [(Shape(64, 16, 1, 32), 12, 3), (Shape(16, 4, 1, 8), 5, 4)]:
        p = _params(seed, shape.nt)
        start, ((a, b), sc) = _context_run(shape, prec, p, nsteps)
        (c, d), sn = _noalloc_run(shape, prec, p, start, nsteps)
        print(f"  {shape}: max |delta| = {max(np.max(abs(a - c)), np.max(abs(b - d))):.3e}", flush=True)
        assert np.array_equal(a, c) and np.array_equal(b, d), shape
        assert (sn.steps, sn.dt, sn.max_drift, sn.mean_drift, sn.t_langevin, sn.nonfinite) == \
               (sc.steps, sc.dt, sc.max_drift, sc.mean_drift, sc.t_langevin, sc.nonfinite), shape


def test_noalloc_rejects_what_the_scratch_cannot_hold():
    print("case: clv_step_noalloc limits", flush=True)
    lib = _abi.lib()
    st = _abi.StepStats()
    p = _params(1, 256)
    fake = [C.c_void_p(k << 40) for k in range(1, 5)]   # never dereferenced: rejected first
    rc = lib.clv_step_noalloc(C.byref(Shape(256, 512, 1, 256)), CLV_FP32, C.byref(p), *fake,
                              C.byref(st), 0, None)
    assert rc == _abi.CLV_ERR_ARG and b"scratch" in lib.clv_last_error(None)
    same = [fake[0], fake[1], fake[0], fake[3]]
    rc = lib.clv_step_noalloc(C.byref(Shape(16, 4, 1, 8)), CLV_FP32, C.byref(p), *same,
                              C.byref(st), 0, None)
    assert rc == _abi.CLV_ERR_ARG and b"overlap" in lib.clv_last_error(None)
