"""Serving: continuous batching through clv_server_* (GPU).

Checks that a run's result is the same alone and in a mix of other runs that
arrive while it is on the device (bit for bit: noise keyed by seed and step,
per-chain fixed-order reductions), that it equals the same chain run through a
plain context, that the HBM budget is enforced, and that the secant mu search
drives follow-up runs toward a target density.  CPU-only checks cover the
secant update (mylib.py:561-584).
"""
import time

import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, Shape, default_params
from cl_var_dt_amd import serving


def _p(seed, mu=1.0, nt=8, **kw):
    base = dict(mu=mu, G=0.1, c1=-0.05, Q=0.1, eps=1.0 / nt, DT=1e-3, seed=seed)
    base.update(kw)
    return default_params(**base)


# ------------------------------------------------------------------ CPU ----
def test_new_mu_is_the_secant_root_of_a_line():
    # residual r(mu) = 2 (mu - 0.7): one secant step lands on the root
    mus = [0.2, 1.5]
    r = [2 * (m - 0.7) for m in mus]
    assert serving.new_mu(r, mus) == pytest.approx(0.7, abs=1e-15)
    assert serving.check_to_stop(0.0005, 1e-3) and not serving.check_to_stop(0.002, 1e-3)


def test_secant_search_on_a_smooth_curve():
    f = lambda mu: 0.5 + np.tanh(mu)          # noqa: E731  density(mu), monotone
    mu, hist = serving.secant_mu_search(f, 0.0, 1.0, target=1.0, tol=1e-9)
    assert f(mu) == pytest.approx(1.0, abs=1e-9)
    assert np.arctanh(0.5) == pytest.approx(mu, abs=1e-8)
    assert len(hist) < 10


# ------------------------------------------------------------------ GPU ----
@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix():
    print("case: serving alone vs mix", flush=True)
    shape_a = Shape(64, 64, 1, 32)
    pa = _p(11, mu=0.9, nt=32)
    with serving.Server(precision=CLV_FP32) as srv:
        alone = srv.wait(srv.submit(shape_a, pa, nsteps=120, eval_every=10))
    with serving.Server(precision=CLV_FP32) as srv:
        others = [(Shape(128, 128, 1, 32), _p(21, mu=1.2, nt=32), 200),
                  (Shape(64, 64, 1, 64), _p(22, mu=0.6, nt=64), 90),
                  (Shape(64, 64, 1, 32), _p(23, mu=1.4, nt=32), 150)]
        ids = [srv.submit(s, p, nsteps=n, eval_every=10) for s, p, n in others[:2]]
        time.sleep(0.05)
        rid = srv.submit(shape_a, pa, nsteps=120, eval_every=10)     # arrives mid-flight
        time.sleep(0.02)
        ids.append(srv.submit(*others[2][:2], nsteps=others[2][2], eval_every=10))
        mixed = srv.wait(rid)
        rest = [srv.wait(i) for i in ids]
        info = srv.info()
    assert info.completed == 4 and info.failed == 0
    assert mixed.count == alone.count == 12
    for f in ("n_tot", "density", "Sz", "energy", "t_langevin", "steps"):
        assert getattr(mixed, f) == getattr(alone, f), f
    assert all(r.steps == n for r, (_, _, n) in zip(rest, [others[0], others[1], others[2]]))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [CLV_FP32, CLV_FP64])
def test_server_run_equals_plain_context(prec):
    print("case: serving vs context", prec, flush=True)
    from cl_var_dt_amd.engine import Engine
    shape = Shape(32, 32, 1, 16)
    p = _p(5, mu=0.8, nt=16)
    with serving.Server(precision=prec) as srv:
        r = srv.wait(srv.submit(shape, p, nsteps=75, eval_every=5, eval_from=20))
    eng = Engine([shape], precision=prec)
    eng.set_params(0, p)
    eng.init_fields(0, 1.0, 0.1)
    for i in range(75):
        eng.propagate(1, 0)
        if i >= 20 and i % 5 == 0:
            eng.evaluate(0)
    o, _ = eng.observables(0)
    st = eng.stats(0)
    eng.close()
    assert r.count == o.count == 11
    assert r.n_tot == o.n_tot and r.energy == o.energy and r.t_langevin == st.t_langevin


@pytest.mark.gpu
def test_budget_is_enforced():
    print("case: serving budget", flush=True)
    from cl_var_dt_amd import _abi
    small = 64 << 20
    with serving.Server(precision=CLV_FP32, hbm_budget_bytes=small) as srv:
        big = srv.submit(Shape(256, 256, 1, 128), _p(1, nt=128), nsteps=5)
        ok = srv.submit(Shape(16, 16, 1, 8), _p(2), nsteps=5)
        with pytest.raises(_abi.ClvError):
            srv.wait(big, timeout=60)
        assert srv.wait(ok, timeout=60).steps == 5
        info = srv.info()
        assert info.bytes_in_use <= small and info.failed == 1


@pytest.mark.gpu
def test_secant_mu_search_through_the_server():
    print("case: serving secant", flush=True)
    shape = Shape(16, 16, 1, 8)
    with serving.Server(precision=CLV_FP32) as srv:
        def density(mu):
            return srv.wait(srv.submit(shape, _p(77, mu=mu), nsteps=400, eval_every=10,
                                       eval_from=100)).density
        target = 4.0
        mu, hist = serving.secant_mu_search(density, 0.2, 0.6, target=target, tol=0.05,
                                            max_iter=8)
    assert abs(hist[-1][1]) < 0.05 or len(hist) == 10
    # the search moved toward the target
    assert abs(hist[-1][1]) < abs(hist[0][1])


@pytest.mark.gpu
def test_release_forgets_finished_runs_only():
    print("case: serving release", flush=True)
    from cl_var_dt_amd import _abi
    shape = Shape(16, 16, 1, 8)
    with serving.Server(precision=CLV_FP32) as srv:
        rid = srv.submit(shape, _p(5), nsteps=2000)
        with pytest.raises(_abi.ClvError) as e:       # queued or running: kept
            srv.release(rid)
        assert e.value.status == _abi.CLV_ERR_STATE
        assert srv.wait(rid, timeout=120).steps == 2000
        srv.release(rid)
        with pytest.raises(_abi.ClvError) as e:
            srv.poll(rid)
        assert e.value.status == _abi.CLV_ERR_NOTFOUND
        with pytest.raises(_abi.ClvError):
            srv.release(rid)
        rid2 = srv.submit(shape, _p(6), nsteps=20)
        assert srv.wait(rid2, timeout=60, release=True).steps == 20
        with pytest.raises(_abi.ClvError):
            srv.poll(rid2)


@pytest.mark.gpu
def test_largest_request_same_alone_and_in_the_mix():
    """A run of the largest serving shape (256x256x1x128, 8.4 M sites) gives the
    same result alone and admitted while other shapes are in flight."""
    print("case: serving largest request", flush=True)
    big = Shape(256, 256, 1, 128)
    pb = _p(31, mu=1.1, nt=128)
    with serving.Server(precision=CLV_FP32) as srv:
        alone = srv.wait(srv.submit(big, pb, nsteps=30, eval_every=10), release=True)
    with serving.Server(precision=CLV_FP32) as srv:
        ids = [srv.submit(Shape(128, 128, 1, 64), _p(32, mu=0.7, nt=64), nsteps=60, eval_every=10),
               srv.submit(Shape(64, 64, 1, 32), _p(33, mu=1.3, nt=32), nsteps=80, eval_every=10)]
        time.sleep(0.02)
        rid = srv.submit(big, pb, nsteps=30, eval_every=10)
        ids.append(srv.submit(Shape(256, 256, 1, 32), _p(34, mu=0.9, nt=32), nsteps=40, eval_every=10))
        mixed = srv.wait(rid, release=True)
        for i in ids:
            srv.wait(i, release=True)
        assert srv.info().failed == 0
    assert mixed.count == alone.count == 3
    for f in ("n_tot", "density", "Sz", "energy", "t_langevin", "steps"):
        assert getattr(mixed, f) == getattr(alone, f), f


def test_windowed_means_closed_forms():
    # no cut: the window means are the cumulative means themselves
    m = np.array([[1.0, 2.0, 3.0, 4.0], [3.0, 2.0, 5.0, 0.0]])
    evol, err = serving.windowed_means(m, 0.0)
    assert np.allclose(evol, m[:, 1:].mean(axis=0)) and np.allclose(err, m[:, 1:].std(axis=0) / 2 ** 0.5)
    # constant series: the window mean is the constant, with zero spread
    evol, err = serving.windowed_means(np.full((3, 10), 0.7), 0.4)
    assert np.allclose(evol, 0.7) and np.allclose(err, 0.0) and len(evol) == 10 - 4 - 1
    # by hand, rows = 4, cut 0.5: i0 = 2, i = 1: v = (4/2) (m3 - (2/4) m2) = 2 m3 - m2
    m = np.array([[0.0, 0.0, 1.0, 2.0], [0.0, 0.0, 3.0, 2.0]])
    evol, err = serving.windowed_means(m, 0.5)
    assert np.allclose(evol, [2.0]) and np.allclose(err, [1.0 / 2 ** 0.5])
    # excluded runs are 1-based, as in the reference
    evol, err = serving.windowed_means(m, 0.5, excluded=[2])
    assert np.allclose(evol, [3.0]) and np.allclose(err, [0.0])
    with pytest.raises(ValueError):
        serving.windowed_means(np.array([[1.0, 2.0, np.nan]]), 0.0)
