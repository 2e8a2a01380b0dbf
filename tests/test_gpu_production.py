"""Parity with the FP64 oracle at the production shapes and tilings (GPU).

Every other engine-vs-oracle test runs at one or two row tiles per tau-slice.
These run each workload's exact chain shape through the path that workload
uses in production -- the full batch, the autotuned schedule, graph-replayed
steps, the server -- so the multi-tile row indexing (r = tile * rows + j), the
unrolled partial reductions (hundreds to thousands of partials per chain) and
the fused march over long tau ranges are compared with the oracle directly:

  workload                             tiles x Nt per chain   steps   bar
  128 chains of 128x128x1x32, FP32     64 x 32                50      1e-4 rms-relative per site
  one 128x128x1x32 chain, FP64         64 x 32                3       1e-11 per site, both schedules
  16 chains of 64x64x1x64, FP64        32 x 64, 2 tau blocks  50      1e-10 per site, fused in tau blocks
  2-D 64x64x1x64, FP64 (device dt)     32 x 64                1000    1e-10 per site (the 1000-step trajectory gate)
  uneven 3-D batch, 24^3 / 32^3, FP32  144 / 128 x Nt          20      1e-4 rms-relative per site
  256x256x1x128, FP32 (context+server) 256 x 128              10      1e-4 rms-relative; observables 1e-4

The oracle runs on the host's cores (at most 16 threads).
"""
import os
import random

import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, Shape, default_params

pytestmark = pytest.mark.gpu


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _engine(shapes, precision):
    from cl_var_dt_amd.engine import Engine
    return Engine([Shape(*s) for s in shapes], precision=precision)


def _params(seed, nt, mu=1.0, beta=1.0, **kw):
    base = dict(mu=mu, G=0.1, c1=-0.05, Q=0.1, P=0.0, eps=beta / nt, DT=1e-3, seed=seed)
    base.update(kw)
    return default_params(**base)


def _oracle(oracle_lib, shape, p, steps, eval_every):
    orc = oracle_lib.Oracle(Shape(*shape), p, nthreads=_threads())
    orc.init_fields(1.0, 0.1)
    orc.propagate(steps, eval_every)
    out = orc.get_fields(), orc.stats(), orc.observables()[0]
    orc.close()
    return out


def _rel(a, b, c, d):
    scale = np.sqrt(np.mean(abs(c) ** 2))
    return max(np.max(abs(a - c)), np.max(abs(b - d))) / scale


def test_batched_128x128x32_fp32_chains_match_oracle(oracle_lib):
    """The batched workload as the bench runs it: 128 chains, mu on a grid in
    [0.5, 1.5], autotuned schedule, graph-replayed steps, observables every 10;
    chains 0 and 127 against the oracle after 50 steps."""
    print("case: batched 128 x 128x128x1x32 fp32 vs oracle", flush=True)
    nch, shape = 128, (128, 128, 1, 32)
    mus = [0.5 + i / (nch - 1) for i in range(nch)]
    e = _engine([shape] * nch, CLV_FP32)
    for i in range(nch):
        e.set_params(i, _params(1000 + i, 32, mu=mus[i]))
        e.init_fields(i, 1.0, 0.1)
    e.autotune(None)
    print("schedule", e.schedule(0), flush=True)
    e.propagate(50, 10)
    e.synchronize()
    for i in (0, 127):
        (c, d), so, oo = _oracle(oracle_lib, shape, _params(1000 + i, 32, mu=mus[i]), 50, 10)
        a, b = e.get_fields(i)
        sg, og = e.stats(i), e.observables(i)[0]
        rel = _rel(a, b, c, d)
        print(f"chain {i}: max per-site |delta| / rms = {rel:.3e}, dt {sg.dt:.9e} vs {so.dt:.9e}",
              flush=True)
        assert rel <= 1e-4, i
        assert sg.steps == so.steps == 50 and sg.nonfinite == 0
        assert sg.t_langevin == pytest.approx(so.t_langevin, rel=1e-5)
        assert og.count == oo.count == 5
        for f in ("density", "Sz", "energy"):
            assert getattr(og, f) == pytest.approx(getattr(oo, f), rel=1e-4, abs=1e-6), f
    e.close()


@pytest.mark.parametrize("sched", [1, 2])
def test_128x128x32_fp64_steps_match_oracle(oracle_lib, sched):
    """FP64 at the batched workload's tiling (64 row tiles x 32 slices = 2048
    partials per chain, the unrolled reduction), both schedules, with noise:
    3 steps within 1e-11 per site, dt and max |D_0| to 1e-12."""
    print("case: 128x128x1x32 fp64 vs oracle, schedule", sched, flush=True)
    shape = (128, 128, 1, 32)
    p = _params(7, 32, mu=0.8)
    e = _engine([shape, shape], CLV_FP64)
    for i in range(2):
        e.set_params(i, p if i == 0 else _params(8, 32))
        e.init_fields(i, 1.0, 0.1)
    e.set_schedule(sched)
    e.propagate(3, 0)
    (c, d), so, _ = _oracle(oracle_lib, shape, p, 3, 0)
    a, b = e.get_fields(0)
    err = max(np.max(abs(a - c)), np.max(abs(b - d)))
    sg = e.stats(0)
    print(f"max per-site |delta| = {err:.3e}", flush=True)
    assert err <= 1e-11
    assert sg.dt == pytest.approx(so.dt, rel=1e-12)
    assert sg.max_drift == pytest.approx(so.max_drift, rel=1e-12)
    assert sg.mean_drift == pytest.approx(so.mean_drift, rel=1e-12)
    e.close()


def test_fp64_batch_64x64x64_tau_blocks_match_oracle(oracle_lib):
    """The FP64 batched workload (16 chains of 64x64x1x64, the bench's
    fp64_batched leg) on its untimed production plan -- the fused march in two
    tau blocks, the blocks' first slices completed by the strided
    drift-statistics pass, graph-replayed steps with observables every 10 --
    against the oracle for chains 0 and 15 after 50 steps: 1e-10 per site,
    dt and the drift statistics to 1e-12, observables to 1e-10."""
    print("case: 16 x 64x64x1x64 fp64 tau blocks vs oracle", flush=True)
    from cl_var_dt_amd._abi import SCHED_FUSED
    shape = (64, 64, 1, 64)
    nch = 16
    ps = [_params(300 + i, 64, mu=0.5 + i / (nch - 1)) for i in range(nch)]
    e = _engine([shape] * nch, CLV_FP64)
    for i in range(nch):
        e.set_params(i, ps[i])
        e.init_fields(i, 1.0, 0.1)
    e.set_schedule(0)
    assert e.schedule(0) == SCHED_FUSED  # 32 two-wave tiles x 16 chains x 2 blocks = 2048 waves
    e.propagate(50, 10)
    for i in (0, nch - 1):
        (c, d), so, oo = _oracle(oracle_lib, shape, ps[i], 50, 10)
        a, b = e.get_fields(i)
        err = max(np.max(abs(a - c)), np.max(abs(b - d)))
        sg = e.stats(i)
        o, _ = e.observables(i)
        print(f"chain {i}: max per-site |delta| = {err:.3e}", flush=True)
        assert err <= 1e-10
        assert sg.dt == pytest.approx(so.dt, rel=1e-12)
        assert sg.max_drift == pytest.approx(so.max_drift, rel=1e-12)
        assert sg.mean_drift == pytest.approx(so.mean_drift, rel=1e-12)
        assert o.count == oo.count
        assert o.n_tot == pytest.approx(oo.n_tot, rel=1e-10)
        assert o.energy == pytest.approx(oo.energy, rel=1e-10)
    e.close()


def test_plain_config_64x64x64_fp64_1000_steps(oracle_lib):
    """The plain-method FP64 workload's shape (2-D 64x64 x Nt=64, beta = 1,
    mu = 1, c0 = 0.1, c1 = -0.05, q = 0.1, seed 7): 1000 adaptive steps on the
    device-resident controller with graph replay and observables every 10,
    per site within 1e-10 of the oracle; then 100 host-driven (plain) steps
    from there, also within 1e-10."""
    print("case: 64x64x1x64 fp64 1000 steps vs oracle", flush=True)
    shape = (64, 64, 1, 64)
    p = _params(7, 64)
    e = _engine([shape], CLV_FP64)
    e.set_params(0, p)
    e.init_fields(0, 1.0, 0.1)
    e.propagate(1000, 10)
    orc = oracle_lib.Oracle(Shape(*shape), p, nthreads=_threads())
    orc.init_fields(1.0, 0.1)
    orc.propagate(1000, 10)
    a, b = e.get_fields(0)
    c, d = orc.get_fields()
    err = max(np.max(abs(a - c)), np.max(abs(b - d)))
    print(f"1000 steps: max per-site |delta| = {err:.3e}", flush=True)
    assert err <= 1e-10
    sg, so = e.stats(0), orc.stats()
    assert sg.steps == so.steps == 1000
    assert sg.t_langevin == pytest.approx(so.t_langevin, rel=1e-12)
    og, oo = e.observables(0)[0], orc.observables()[0]
    assert og.count == oo.count == 100
    for f in ("n_tot", "density", "Sz", "energy"):
        assert getattr(og, f) == pytest.approx(getattr(oo, f), rel=1e-10, abs=1e-12), f
    e.propagate_plain(0, 100, 10)
    orc.propagate(100, 10)
    a, b = e.get_fields(0)
    c, d = orc.get_fields()
    err = max(np.max(abs(a - c)), np.max(abs(b - d)))
    print(f"+100 plain steps: max per-site |delta| = {err:.3e}", flush=True)
    assert err <= 1e-10
    orc.close()
    e.close()


def test_uneven_3d_batch_chains_match_oracle(oracle_lib):
    """The uneven batch as the bench runs it (256 chains of L^3 x Nt, L in
    {16, 24, 32}, Nt in {16, 32, 64}, beta in [0.5, 2], drawn from seed 3,
    autotuned per-bucket schedules, graph-replayed steps): the first 24^3 x 32
    and 32^3 x 64 chains of the batch against the oracle after 20 steps."""
    print("case: uneven 3-D batch vs oracle", flush=True)
    rng = random.Random(3)
    shapes, betas = [], []
    for _ in range(256):
        L = rng.choice([16, 24, 32])
        nt = rng.choice([16, 32, 64])
        shapes.append((L, L, L, nt))
        betas.append(rng.uniform(0.5, 2.0))
    probe = [shapes.index(s) for s in [(24, 24, 24, 32), (32, 32, 32, 64)]]

    def params(i):
        return _params(5000 + i, shapes[i][3], mu=1.0, beta=betas[i])

    e = _engine(shapes, CLV_FP32)
    for i in range(len(shapes)):
        e.set_params(i, params(i))
        e.init_fields(i, 1.0, 0.1)
    e.autotune(None)
    e.propagate(20, 10)
    e.synchronize()
    for i in probe:
        (c, d), so, oo = _oracle(oracle_lib, shapes[i], params(i), 20, 10)
        a, b = e.get_fields(i)
        sg, og = e.stats(i), e.observables(i)[0]
        rel = _rel(a, b, c, d)
        print(f"chain {i} {shapes[i]} schedule {e.schedule(i)}: max per-site |delta| / rms = {rel:.3e}",
              flush=True)
        assert rel <= 1e-4, i
        assert sg.steps == so.steps == 20
        assert sg.t_langevin == pytest.approx(so.t_langevin, rel=1e-5)
        assert og.count == oo.count == 2
        assert og.density == pytest.approx(oo.density, rel=1e-4)
    e.close()


def test_largest_serving_shape_matches_oracle(oracle_lib):
    """The largest serving request, 256x256x1x128 FP32 (one row per tile: 256
    tiles x 128 slices = 32768 partials): 10 steps in a context (fields) and
    through the server (observables, Langevin time) against the oracle."""
    print("case: 256x256x1x128 fp32 vs oracle", flush=True)
    from cl_var_dt_amd import serving
    shape = (256, 256, 1, 128)
    p = _params(31, 128, mu=1.1)
    (c, d), so, oo = _oracle(oracle_lib, shape, p, 10, 5)
    e = _engine([shape], CLV_FP32)
    e.set_params(0, p)
    e.init_fields(0, 1.0, 0.1)
    e.propagate(10, 5)
    a, b = e.get_fields(0)
    rel = _rel(a, b, c, d)
    print(f"context: max per-site |delta| / rms = {rel:.3e}", flush=True)
    assert rel <= 1e-4
    e.close()
    with serving.Server(precision=CLV_FP32) as srv:
        r = srv.wait(srv.submit(Shape(*shape), p, nsteps=10, eval_every=5), release=True)
    print(f"server: density {r.density:.9f} vs {oo.density:.9f}", flush=True)
    assert r.steps == 10 and r.count == oo.count == 2
    assert r.t_langevin == pytest.approx(so.t_langevin, rel=1e-5)
    for f in ("density", "Sz", "energy"):
        assert getattr(r, f) == pytest.approx(getattr(oo, f), rel=1e-4, abs=1e-6), f
