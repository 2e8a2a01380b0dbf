"""Statistical checks of the whole complex-Langevin machinery on the GPU.

1. Free theory (G = c1 = P = Q = 0, psibar drift as the action derivative):
   the action is Gaussian and diagonal in momentum, S = sum_k sum_t
   phibar_k(t) [phi_k(t) - a_k phi_k(t-1)] with a_k = 1 + eps (mu - p_k^2), so
   <phibar_k(t+1) phi_k(t)> = a_k^(Nt-1) / (1 - a_k^Nt) exactly and
   n_tot = 4 sum_k a_k^(Nt-1) / (1 - a_k^Nt).  Independent chains give the
   standard error; the engine must agree within 2 sigma (plus the O(dt) bias
   of Euler-Maruyama, bounded here by a 1% allowance).  This pins the kinetic
   sign and normalisation, the drift, the noise variance and the observable.
2. FP32 batched chains against the FP64 CPU oracle on the same Philox
   streams: density, Sz and energy means over independent chains agree within
   two standard errors; the measured differences are printed.
"""
import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, CLV_CONJ_DRIFT_ACTION, Shape, default_params

pytestmark = pytest.mark.gpu


def _engine(shapes, precision):
    from cl_var_dt_amd.engine import Engine
    return Engine(shapes, precision=precision)


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
    o = np.array(o)
    for k, name in enumerate(["density", "Sz", "energy"]):
        se = np.sqrt(g[:, k].var(ddof=1) / nch + o[:, k].var(ddof=1) / nch)
        diff = g[:, k].mean() - o[:, k].mean()
        print(f"{name}: fp32 {g[:, k].mean():.6f}  fp64 oracle {o[:, k].mean():.6f}  "
              f"diff {diff:+.2e}  2 sigma {2 * se:.2e}")
        assert abs(diff) <= 2 * se + 1e-12


def test_fp32_and_fp64_engine_chains_agree():
    """The 2-D configuration's couplings (mu = 1, c0 = 0.1, c1 = -0.05, q = 0.1)
    on a reduced 16^2 x 16 lattice: FP32 and FP64 chains, density / Sz /
    energy within two standard errors."""
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
    for k, name in enumerate(["density", "Sz", "energy"]):
        a, b = res[CLV_FP32][:, k], res[CLV_FP64][:, k]
        se = np.sqrt(a.var(ddof=1) / nch + b.var(ddof=1) / nch)
        print(f"{name}: fp32 {a.mean():.6f} fp64 {b.mean():.6f} diff {a.mean() - b.mean():+.2e} "
              f"2 sigma {2 * se:.2e} max per-chain |diff| {np.max(abs(a - b)):.2e}")
        assert abs(a.mean() - b.mean()) <= 2 * se + 1e-12


@pytest.mark.parametrize("c1", [-0.05, 0.05])
def test_config1_fp32_vs_fp64_observables(c1):
    """The plain-method configuration at full length: 2-D 64x64 x Nt=64, beta = 1,
    mu = 1, c0 = 0.1, q = 0.1, c1 in {-0.05, +0.05}, 5x10^4 adaptive steps,
    Gaussian start around sqrt(n0 = 1), observables every 10 steps accumulated
    in FP64 from step 0 (as the reference does).  16 independent chains per
    precision (seeds 7, 8, ...): the FP32 and FP64 means of density, Sz and
    energy agree within two standard errors of their difference."""
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
    for k, name in enumerate(["density", "Sz", "energy"]):
        a, b = res[CLV_FP32][:, k], res[CLV_FP64][:, k]
        se = np.sqrt(a.var(ddof=1) / nch + b.var(ddof=1) / nch)
        print(f"c1={c1:+.2f} {name}: fp32 {a.mean():.6f} fp64 {b.mean():.6f} "
              f"diff {a.mean() - b.mean():+.2e} 2 sigma {2 * se:.2e} "
              f"max per-chain |diff| {np.max(abs(a - b)):.2e}")
        assert abs(a.mean() - b.mean()) <= 2 * se + 1e-12
