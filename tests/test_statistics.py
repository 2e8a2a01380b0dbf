"""Statistical checks of the whole complex-Langevin machinery on the GPU.

1. Free theory (G = c1 = P = Q = 0, psibar drift as the action derivative):
   the action is Gaussian and diagonal in momentum, S = sum_k sum_t
   phibar_k(t) [phi_k(t) - a_k phi_k(t-1)] with a_k = 1 + eps (mu - p_k^2), so
   <phibar_k(t+1) phi_k(t)> = a_k^(Nt-1) / (1 - a_k^Nt) exactly and
   n_tot = 4 sum_k a_k^(Nt-1) / (1 - a_k^Nt).  Independent chains give the
   standard error; the engine must agree within 2 sigma (plus the O(dt) bias
   of Euler-Maruyama, bounded here by a 1% allowance).  This pins the kinetic
   sign and normalisation, the drift, the noise variance and the observable.
2. FP32 batched chains against the FP64 CPU oracle, and FP32 engine chains
   against FP64 engine chains, on the same Philox streams (chain i has seed i
   in both).  Two checks per observable, each printed:
   - statistical: the two means agree within two standard errors of the
     independent-chain spread (the acceptance criterion for FP32 observables);
   - paired: the per-chain differences d_i = fp32_i - fp64_i have a mean
     within two standard errors of the d_i themselves plus the FP32 rounding
     allowance 1e-5 (|mean| + chain spread) -- a bias ten times the observed
     one fails this,
     where the statistical check alone would pass biases a thousand times
     larger.
"""
import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, CLV_CONJ_DRIFT_ACTION, Shape, default_params

pytestmark = pytest.mark.gpu


def _engine(shapes, precision):
    from cl_var_dt_amd.engine import Engine
    return Engine(shapes, precision=precision)


NAMES = ["density", "Sz", "energy"]


def _compare(a, b, label, rel=1e-5):
    """a: FP32 runs, b: FP64 runs of the same seeds, shape (chains, observables)."""
    n = a.shape[0]
    for k, name in enumerate(NAMES):
        x, y = a[:, k], b[:, k]
        se = np.sqrt(x.var(ddof=1) / n + y.var(ddof=1) / n)
        d = x - y
        se_d = d.std(ddof=1) / np.sqrt(n)
        allow = rel * (abs(y.mean()) + y.std(ddof=1))
        print(f"{label} {name}: fp32 {x.mean():.8f} fp64 {y.mean():.8f} diff {d.mean():+.2e} | "
              f"statistical 2 sigma {2 * se:.2e} | paired mean {d.mean():+.2e} 2 sigma_d {2 * se_d:.2e} "
              f"+ allowance {allow:.2e} | max |d_i| {np.max(abs(d)):.2e}")
        assert abs(x.mean() - y.mean()) <= 2 * se + 1e-12, name
        assert abs(d.mean()) <= 2 * se_d + allow, name


def test_free_theory_occupation():
    print("case: free theory", flush=True)
    X, Nt, mu, eps = 8, 8, -0.5, 0.1
    nch = 64
    shape = Shape(X, 1, 1, Nt)
    eng = _engine([shape] * nch, CLV_FP32)
    for i in range(nch):
        eng.set_params(i, default_params(mu=mu, G=0.0, c1=0.0, P=0.0, Q=0.0, eps=eps, DT=1e-2,
                                         conj_drift_mode=CLV_CONJ_DRIFT_ACTION, seed=3000 + i))
        eng.init_fields(i, 0.0, 0.5)
    eng.propagate(20000, 0)          # thermalise (slowest mode relaxes in ~20 time units)
    eng.reset_observables()
    eng.propagate(60000, 10)
    n = np.array([eng.observables(i)[0].n_tot for i in range(nch)])
    eng.close()
    p = 2 * np.pi * np.fft.fftfreq(X)
    a = 1 + eps * (mu - p ** 2)
    exact = 4 * np.sum(a ** (Nt - 1) / (1 - a ** Nt))
    se = n.std(ddof=1) / np.sqrt(nch)
    print(f"free theory n_tot: engine {n.mean():.5f} +- {se:.5f}, exact {exact:.5f}")
    assert abs(n.mean() - exact) < 2 * se + 0.01 * abs(exact)


def test_fp32_observables_match_fp64_oracle(oracle_lib):
    print("case: fp32 vs oracle observables", flush=True)
    shape = Shape(8, 1, 1, 8)
    nch, therm, steps = 12, 2000, 12000
    kw = dict(mu=1.0, G=0.1, c1=-0.05, Q=0.1, eps=1 / 8, DT=1e-2)
    eng = _engine([shape] * nch, CLV_FP32)
    for i in range(nch):
        eng.set_params(i, default_params(seed=4000 + i, **kw))
        eng.init_fields(i, 1.0, 0.1)
    eng.propagate(therm, 0)
    eng.reset_observables()
    eng.propagate(steps, 10)
    g = np.array([[eng.observables(i)[0].density, eng.observables(i)[0].Sz,
                   eng.observables(i)[0].energy] for i in range(nch)])
    eng.close()
    o = []
    for i in range(nch):
        orc = oracle_lib.Oracle(shape, default_params(seed=4000 + i, **kw), nthreads=1)
        orc.init_fields(1.0, 0.1)
        orc.propagate(therm, 0)
        orc.reset_observables()
        orc.propagate(steps, 10)
        ob, _ = orc.observables()
        o.append([ob.density, ob.Sz, ob.energy])
        orc.close()
    _compare(g, np.array(o), "8x8 vs oracle")


def test_fp32_observables_match_fp64_oracle_2d(oracle_lib):
    """A 2-D lattice, 32^2 x 16, 16 independent chains (seeds 4100 + i) of the
    FP32 engine against the FP64 oracle on the same seeds: 300 steps to relax,
    then 1200 with observables every 10."""
    print("case: fp32 vs oracle observables 32x32x16", flush=True)
    import os
    shape = Shape(32, 32, 1, 16)
    nch, therm, steps = 16, 300, 1200
    kw = dict(mu=1.0, G=0.1, c1=-0.05, Q=0.1, eps=1 / 16, DT=5e-3)
    eng = _engine([shape] * nch, CLV_FP32)
    for i in range(nch):
        eng.set_params(i, default_params(seed=4100 + i, **kw))
        eng.init_fields(i, 1.0, 0.1)
    eng.propagate(therm, 0)
    eng.reset_observables()
    eng.propagate(steps, 10)
    g = np.array([[getattr(eng.observables(i)[0], f) for f in NAMES] for i in range(nch)])
    eng.close()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    o = []
    for i in range(nch):
        orc = oracle_lib.Oracle(shape, default_params(seed=4100 + i, **kw), nthreads=threads)
        orc.init_fields(1.0, 0.1)
        orc.propagate(therm, 0)
        orc.reset_observables()
        orc.propagate(steps, 10)
        ob, _ = orc.observables()
        o.append([getattr(ob, f) for f in NAMES])
        orc.close()
    _compare(g, np.array(o), "32x32x16 vs oracle")


def test_fp32_and_fp64_engine_chains_agree():
    """The 2-D configuration's couplings (mu = 1, c0 = 0.1, c1 = -0.05, q = 0.1)
    on a reduced 16^2 x 16 lattice: FP32 and FP64 chains of the same seeds,
    density / Sz / energy by the statistical and the paired check."""
    print("case: fp32 vs fp64 engine", flush=True)
    shape = Shape(16, 16, 1, 16)
    nch = 24
    res = {}
    for prec in (CLV_FP32, CLV_FP64):
        eng = _engine([shape] * nch, prec)
        for i in range(nch):
            eng.set_params(i, default_params(mu=1.0, G=0.1, c1=0.05, Q=0.1, eps=1 / 16, DT=5e-3,
                                             seed=7 + 100 * i))
            eng.init_fields(i, 1.0, 0.1)
        eng.propagate(1000, 0)
        eng.reset_observables()
        eng.propagate(4000, 10)
        res[prec] = np.array([[eng.observables(i)[0].density, eng.observables(i)[0].Sz,
                               eng.observables(i)[0].energy] for i in range(nch)])
        eng.close()
    _compare(res[CLV_FP32], res[CLV_FP64], "16x16x16")


@pytest.mark.parametrize("c1", [-0.05, 0.05])
def test_config1_fp32_vs_fp64_observables(c1):
    """The plain-method configuration at full length: 2-D 64x64 x Nt=64, beta = 1,
    mu = 1, c0 = 0.1, q = 0.1, c1 in {-0.05, +0.05}, 5x10^4 adaptive steps,
    Gaussian start around sqrt(n0 = 1), observables every 10 steps accumulated
    in FP64 from step 0 (as the reference does).  16 independent chains per
    precision (seeds 7, 8, ...): density, Sz and energy by the statistical and
    the paired check."""
    print("case: config 64x64x64 c1 =", c1, flush=True)
    shape = Shape(64, 64, 1, 64)
    nch, steps = 16, 50000
    res = {}
    for prec in (CLV_FP64, CLV_FP32):
        eng = _engine([shape] * nch, prec)
        for i in range(nch):
            eng.set_params(i, default_params(mu=1.0, G=0.1, c1=c1, Q=0.1, P=0.0, eps=1 / 64,
                                             DT=1e-3, seed=7 + i))
            eng.init_fields(i, 1.0, 0.1)
        eng.propagate(steps, 10)
        res[prec] = np.array([[eng.observables(i)[0].density, eng.observables(i)[0].Sz,
                               eng.observables(i)[0].energy] for i in range(nch)])
        assert all(eng.stats(i).steps == steps and eng.stats(i).nonfinite == 0 for i in range(nch))
        eng.close()
    _compare(res[CLV_FP32], res[CLV_FP64], f"64x64x64 c1={c1:+.2f}")
