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


def test_new_mu_guard_falls_back_to_older_points():
    """new_mu (mylib.py:561-580): a step into [10 max(mu), 0] -- reachable only
    when no mu is positive -- moves j to an older point; none left raises
    GuessError on the first step and JError after it."""
    # i = 1: points (mu1, r1) = (-1.0, 1.0), (mu2, r2) = (-2.0, 0.5): secant root -3.0
    # lies in [10 max(mu), 0] = [-10, 0], so j = 0 is tried: (mu0, r0) = (-4.0, -1.0)
    # with (mu2, r2) gives (0.5 * -4 + 1.0 * -2) / 1.5 = -8/3, inside too: JError
    with pytest.raises(serving.JError):
        serving.new_mu([-1.0, 1.0, 0.5], [-4.0, -1.0, -2.0], 1)
    # with mu0 positive the guard is inert: the plain secant of the last pair
    r, mu = [-1.0, 1.0, 0.5], [4.0, -1.0, -2.0]
    assert serving.new_mu(r, mu, 1) == pytest.approx((0.5 * -1.0 - 1.0 * -2.0) / (0.5 - 1.0))
    # first step into the refused range: GuessError
    with pytest.raises(serving.GuessError):
        serving.new_mu([1.0, 0.5], [-1.0, -2.0], 0)
    # fallback that succeeds: the last pair's root (-0.4) is refused, the older pair's
    # root (-1.3) lies below 10 max(mu) = -1 and is taken
    r, mu = [0.3, 0.5, 0.25], [-0.1, -0.2, -0.3]
    last = (r[2] * mu[1] - r[1] * mu[2]) / (r[2] - r[1])
    older = (r[2] * mu[0] - r[0] * mu[2]) / (r[2] - r[0])
    assert -1.0 <= last <= 0.0 and older < -1.0
    assert serving.new_mu(r, mu, 1) == pytest.approx(older)


def test_new_mu_degenerate_step_is_refused():
    # equal residuals: no secant root; the older point is used, or an error raised
    with pytest.raises(serving.GuessError):
        serving.new_mu([0.3, 0.3], [0.5, 0.7], 0)
    r, mu = [-0.2, 0.3, 0.3], [0.4, 0.5, 0.7]
    assert serving.new_mu(r, mu, 1) == pytest.approx((0.3 * 0.4 + 0.2 * 0.7) / 0.5)


def test_change_guess_and_secant_method_stop_on_the_error():
    assert serving.change_guess([0.5, 0.2], 0.8) == 0.4
    with pytest.raises(serving.GuessError):
        serving.change_guess([0.5, -0.2], 0.8)
    f = lambda mu: 0.5 + np.tanh(mu)          # noqa: E731  density(mu), monotone
    calls = []

    def run_many(mus):
        calls.append(list(mus))
        return [(f(m), 0.02) for m in mus]
    mu, hist, ok = serving.secant_method(run_many, 0.8, 0.25)
    assert ok and abs(f(mu) - 1.0) < 0.02
    assert calls[0] == [1.0, 0.6] and all(len(c) == 1 for c in calls[1:])
    # the residual error, not a fixed tolerance, ends the search
    mu2, hist2, ok2 = serving.secant_method(lambda mus: [(f(m), 1e-9) for m in mus], 0.8, 0.25)
    assert ok2 and len(hist2) > len(hist) and abs(f(mu2) - 1.0) < 1e-9
    mu3, hist3 = serving.secant_mu_search(lambda m: (f(m), 0.02), 1.0, 0.6)
    assert abs(hist3[-1][1]) < 0.02 and len(hist3) == len(hist)


# ------------------------------------------------------------------ GPU ----
def _wait_running(srv, rid, timeout=30.0):
    t0 = time.perf_counter()
    while srv.poll(rid).status != 1:          # CLV_RUN_RUNNING
        assert time.perf_counter() - t0 < timeout, "run never started"
        time.sleep(0.001)


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix():
    """A run admitted while other runs are on the device (shown from the
    server's own admit/done stamps) returns the same result, bit for bit, as
    the run alone."""
    print("case: serving alone vs mix", flush=True)
    shape_a = Shape(64, 64, 1, 32)
    pa = _p(11, mu=0.9, nt=32)
    with serving.Server(precision=CLV_FP32) as srv:
        alone = srv.wait(srv.submit(shape_a, pa, nsteps=120, eval_every=10))
    with serving.Server(precision=CLV_FP32) as srv:
        others = [(Shape(128, 128, 1, 32), _p(21, mu=1.2, nt=32), 6000),
                  (Shape(64, 64, 1, 64), _p(22, mu=0.6, nt=64), 3000),
                  (Shape(64, 64, 1, 32), _p(23, mu=1.4, nt=32), 150)]
        ids = [srv.submit(s, p, nsteps=n, eval_every=10) for s, p, n in others[:2]]
        for i in ids:
            _wait_running(srv, i)
        rid = srv.submit(shape_a, pa, nsteps=120, eval_every=10)     # arrives mid-flight
        ids.append(srv.submit(*others[2][:2], nsteps=others[2][2], eval_every=10))
        mixed = srv.wait(rid)
        rest = [srv.wait(i) for i in ids]
        info = srv.info()
    assert info.completed == 4 and info.failed == 0
    m = mixed.raw
    # admitted while both earlier runs were in flight, and ended before they did
    for r in rest[:2]:
        assert r.raw.t_admit < m.t_admit < r.raw.t_done, (r.raw.t_admit, m.t_admit, r.raw.t_done)
    assert m.t_done < min(r.raw.t_done for r in rest[:2])
    assert mixed.count == alone.count == 12
    for f in ("n_tot", "density", "Sz", "energy", "t_langevin", "steps"):
        assert getattr(mixed, f) == getattr(alone, f), f
    assert all(r.steps == n for r, (_, _, n) in zip(rest, [others[0], others[1], others[2]]))


@pytest.mark.gpu
def test_submit_does_not_wait_for_pool_creation():
    """clv_server_submit never queues behind the worker's device work: while a
    4 GB pool for a 256x256x1x128 run is being created and stepped, every
    submit returns in under 1 ms, and t_submit is stamped on entry."""
    print("case: serving submit latency", flush=True)
    with serving.Server(precision=CLV_FP32) as srv:
        t0 = srv.clock()
        big = srv.submit(Shape(256, 256, 1, 128), _p(41, nt=128), nsteps=20, eval_every=10)
        durations, stamps, ids = [], [], []
        while srv.clock() - t0 < 0.3:
            ts = srv.clock()
            ids.append(srv.submit(Shape(16, 16, 1, 8), _p(42 + len(ids)), nsteps=2, eval_every=0))
            te = srv.clock()
            durations.append(te - ts)
            stamps.append(ts - t0)
            time.sleep(5e-4)
        raw = srv.poll(ids[0])
        assert ts - 1e-3 <= srv.poll(ids[-1]).t_submit <= te
        assert stamps[0] <= raw.t_submit - t0 <= stamps[0] + durations[0]
        srv.wait(big, timeout=120)
        for i in ids:
            srv.wait(i, timeout=120, release=True)
        info = srv.info()
    setup = info.pool_setup_seconds
    during = sum(1 for t in stamps if t < setup)
    print(f"{len(durations)} submits, max {1e3 * max(durations):.3f} ms, pool setup {1e3 * setup:.1f} ms, "
          f"{during} submits during it", flush=True)
    assert info.completed == len(ids) + 1 and info.failed == 0
    assert max(durations) < 1e-3


@pytest.mark.gpu
def test_cancel_frees_the_slot_and_leaves_other_runs_unchanged():
    """clv_server_cancel: a queued run never starts; a running run's slot is
    freed at its pool's next tick boundary while the pool's other run goes on
    to the same result as alone; the memory then held by no run is used by a
    run of another shape that did not fit the budget before."""
    print("case: serving cancel", flush=True)
    from cl_var_dt_amd import _abi
    a, b = Shape(64, 64, 1, 32), Shape(64, 64, 1, 64)
    pa = _p(51, mu=0.9, nt=32)
    n_keep = 20000
    with serving.Server(precision=CLV_FP32) as srv:
        alone = srv.wait(srv.submit(a, pa, nsteps=n_keep, eval_every=100), timeout=300, release=True)
    # ~20 MB per a-slot and ~40 MB per b-slot (fields, accumulators): one pool
    # of two a-slots (35 MB), no room for a b-pool beside it
    budget = 50 << 20
    with serving.Server(precision=CLV_FP32, hbm_budget_bytes=budget) as srv:
        long_run = srv.submit(a, _p(52, nt=32), nsteps=10 ** 7)
        keep = srv.submit(a, pa, nsteps=n_keep, eval_every=100)
        _wait_running(srv, long_run)
        _wait_running(srv, keep)
        queued = srv.submit(a, _p(53, nt=32), nsteps=10)   # no free slot, no room for a pool
        other = srv.submit(b, _p(54, nt=64), nsteps=30, eval_every=10)
        time.sleep(0.05)
        assert srv.poll(queued).status == 0 and srv.poll(other).status == 0
        srv.cancel(queued)
        with pytest.raises(_abi.ClvError) as e:
            srv.wait(queued, timeout=30)
        assert e.value.status == _abi.CLV_ERR_CANCELLED and srv.poll(queued).t_admit == 0.0
        srv.cancel(long_run)
        with pytest.raises(_abi.ClvError) as e:
            srv.wait(long_run, timeout=30)
        assert e.value.status == _abi.CLV_ERR_CANCELLED
        lr = srv.poll(long_run)
        assert 0 < lr.stats.steps < 10 ** 7
        # the pool still serves `keep`, so the other shape still has no room
        assert srv.poll(keep).status == 1 and srv.poll(other).status == 0
        kept = srv.wait(keep, timeout=300)
        done = srv.wait(other, timeout=60)
        assert done.steps == 30 and done.raw.t_admit >= max(lr.t_done, kept.raw.t_done)
        with pytest.raises(_abi.ClvError) as e:
            srv.cancel(keep)                           # finished: nothing to cancel
        assert e.value.status == _abi.CLV_ERR_STATE
        info = srv.info()
        assert info.cancelled == 2 and info.completed == 2 and info.failed == 0
        assert info.bytes_in_use <= budget
    for f in ("n_tot", "density", "Sz", "energy", "t_langevin", "steps"):
        assert getattr(kept, f) == getattr(alone, f), f


@pytest.mark.gpu
def test_device_allocation_failure_fails_the_run():
    """A run that fits the budget but not the device memory left (taken by
    another allocation after the server started), with nothing else in
    flight, fails with CLV_ERR_NOMEM and a message instead of waiting
    forever; once the memory is back the same request runs."""
    print("case: serving allocation failure", flush=True)
    from cl_var_dt_amd import _abi
    from tests import hip_rt
    big = Shape(256, 256, 1, 128)                   # a pool of 3 slots, 3.3 GB
    with serving.Server(precision=CLV_FP32, hbm_budget_bytes=16 << 30) as srv:
        free, _ = hip_rt.mem_get_info()
        hog = hip_rt.DeviceBuffer(free - (1 << 30))
        rid = srv.submit(big, _p(61, nt=128), nsteps=1)
        with pytest.raises(_abi.ClvError) as e:
            srv.wait(rid, timeout=120)
        assert e.value.status == _abi.CLV_ERR_NOMEM and "hipMalloc" in str(e.value)
        hog.close()
        assert srv.wait(srv.submit(big, _p(61, nt=128), nsteps=1), timeout=120).steps == 1
        info = srv.info()
        assert info.failed == 1 and info.completed == 1


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
