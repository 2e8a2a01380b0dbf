"""Failure handling and exact resume (GPU; every kernel exits normally).

- A non-finite value injected into one chain through clv_set_fields raises
  that chain's `nonfinite` flag and leaves the other chains of the batch bit
  for bit untouched (no cross-chain reductions or atomics).
- Bad arguments come back as error codes with a message, never as faults.
- A server wait that times out returns an error and the run still completes.
- Saving fields + controller state and restoring them in a fresh context
  resumes a run exactly (counter-based noise keyed by seed and step).
"""
import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, ClvError, Shape, default_params

pytestmark = pytest.mark.gpu


def _engine(shapes, precision=CLV_FP32):
    from cl_var_dt_amd.engine import Engine
    return Engine([Shape(*s) for s in shapes], precision=precision)


def _p(seed, **kw):
    base = dict(mu=1.0, G=0.1, c1=-0.05, Q=0.1, eps=1 / 8, DT=1e-3, seed=seed)
    base.update(kw)
    return default_params(**base)


@pytest.mark.parametrize("sched", [1, 2])
def test_nan_in_one_chain_is_contained(sched):
    print("case: nan containment", sched, flush=True)
    shapes = [(32, 4, 1, 8)] * 3
    runs = []
    for poison in (False, True):
        e = _engine(shapes)
        for i in range(3):
            e.set_params(i, _p(200 + i))
            e.init_fields(i, 1.0, 0.1)
        e.set_schedule(sched)
        if poison:
            f, g = e.get_fields(1)
            f[2, 3, 5, 1, 0] = np.nan
            e.set_fields(1, f, g)
        e.propagate(10, 5)
        runs.append(e)
    clean, bad = runs
    assert bad.stats(1).nonfinite == 1 and clean.stats(1).nonfinite == 0
    for i in (0, 2):
        assert bad.stats(i).nonfinite == 0
        a, _ = clean.get_fields(i)
        b, _ = bad.get_fields(i)
        assert np.array_equal(a, b)
        assert clean.observables(i)[0].n_tot == bad.observables(i)[0].n_tot
    for e in runs:
        e.close()


def test_errors_are_codes_not_faults():
    print("case: error codes", flush=True)
    e = _engine([(16, 2, 1, 4)])
    with pytest.raises(ClvError, match="out of range"):
        e.set_params(5, _p(1))
    with pytest.raises(ClvError, match="kin_dims"):
        e.set_params(0, _p(1, kin_dims=2))
    with pytest.raises(ClvError):
        e.apply_update(0, -1.0)
    with pytest.raises(ClvError):
        e.init_fields(0, -1.0, 0.1)
    e.propagate(2)          # still usable
    assert e.stats(0).steps == 2
    e.close()


def test_server_wait_timeout_then_completion():
    print("case: server timeout", flush=True)
    from cl_var_dt_amd import serving
    with serving.Server(precision=CLV_FP32) as srv:
        rid = srv.submit(Shape(128, 128, 1, 32), _p(9, eps=1 / 32), nsteps=400)
        with pytest.raises(ClvError):
            srv.wait(rid, timeout=1e-4)
        r = srv.wait(rid, timeout=300)
        assert r.steps == 400 and r.count >= 1


@pytest.mark.parametrize("prec", [CLV_FP32, CLV_FP64])
def test_checkpoint_resume_is_exact(prec):
    print("case: resume", prec, flush=True)
    shape = (32, 4, 1, 8)
    straight = _engine([shape], prec)
    straight.set_params(0, _p(33))
    straight.init_fields(0, 1.0, 0.1)
    straight.propagate(30, 0)
    first = _engine([shape], prec)
    first.set_params(0, _p(33))
    first.init_fields(0, 1.0, 0.1)
    first.propagate(15, 0)
    f, g = first.get_fields(0)
    st = first.stats(0)
    first.close()
    resumed = _engine([shape], prec)
    resumed.set_params(0, _p(33))
    resumed.set_fields(0, f, g)
    resumed.set_stats(0, st)
    resumed.propagate(15, 0)
    a, b = straight.get_fields(0)
    c, d = resumed.get_fields(0)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    s1, s2 = straight.stats(0), resumed.stats(0)
    assert s1.steps == s2.steps == 30 and s1.t_langevin == s2.t_langevin
    straight.close()
    resumed.close()


def test_out_of_device_memory_is_an_error():
    """A context larger than the device's HBM fails with an error code and
    message (the first field buffer, 5 x 68.7 GB in FP64, cannot be had), frees
    what it took, and leaves nothing behind for later launches to report."""
    print("case: out of device memory", flush=True)
    from cl_var_dt_amd.engine import Engine
    with pytest.raises(ClvError, match="hipMalloc"):
        Engine([Shape(128, 128, 128, 512)] * 5, precision=CLV_FP64)
    e = _engine([(16, 2, 1, 4)])
    e.set_params(0, _p(5))
    e.init_fields(0, 1.0, 0.1)
    e.propagate(6, 2)
    assert e.stats(0).steps == 6 and e.stats(0).nonfinite == 0
    e.close()
