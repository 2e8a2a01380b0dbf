"""kin_dims = 3 (kinetic term p_x^2 + p_y^2 + p_z^2) on the GPU against the
FP64 oracle, which implements the full Laplacian as a rank-d transform
(the reference keeps p_x^2 only, Propagator_without_cutoff.cu:196).

The engine adds the y and z second derivatives from a strided line transform
(k_transverse) to the x part of its row transform; the oracle multiplies by
p^2 in the 3-D spectrum.  Tolerances as in test_gpu_parity.py: FP64 1e-12 for
one step and single evaluations, 1e-10 per site after long runs; FP32 1e-4.
"""
import numpy as np
import pytest

from cl_var_dt_amd import CLV_FP32, CLV_FP64, Shape, default_params

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 1, 8), (8, 4, 4, 4), (24, 8, 1, 4), (4, 16, 8, 4), (32, 4, 1, 6)]


def _engine(shapes, precision=CLV_FP64):
    from cl_var_dt_amd.engine import Engine
    return Engine([Shape(*s) for s in shapes], precision=precision)


def _params(seed=1, **kw):
    base = dict(mu=1.0, G=0.1, c1=-0.05, Q=0.1, P=0.02, eps=1 / 16, DT=1e-3, seed=seed, kin_dims=3,
                a_x=0.9)
    base.update(kw)
    return default_params(**base)


def _pair(oracle_lib, shape, precision=CLV_FP64, **kw):
    p = _params(**kw)
    eng = _engine([shape], precision)
    eng.set_params(0, p)
    orc = oracle_lib.Oracle(Shape(*shape), p, nthreads=4)
    return eng, orc, p


@pytest.mark.parametrize("shape", SHAPES)
def test_one_step_full_laplacian(oracle_lib, shape):
    print("case: kin_dims=3 one step", shape, flush=True)
    eng, orc, p = _pair(oracle_lib, shape)
    eng.init_fields(0, 1.0, 0.1)
    orc.init_fields(1.0, 0.1)
    mx_g, mean_g = eng.drift_stats(0)
    mx_o, mean_o = orc.max_mean()
    assert mx_g == pytest.approx(mx_o, rel=1e-12) and mean_g == pytest.approx(mean_o, rel=1e-12)
    eng.propagate(1)
    orc.propagate(1)
    a, b = eng.get_fields(0)
    c, d = orc.get_fields()
    assert max(np.max(abs(a - c)), np.max(abs(b - d))) < 1e-12
    # the full Laplacian changes the step against kin_dims = 1
    one = _engine([shape])
    one.set_params(0, _params(kin_dims=1))
    one.init_fields(0, 1.0, 0.1)
    one.propagate(1)
    e, _ = one.get_fields(0)
    assert np.max(abs(e - a)) > 1e-9
    one.close()
    eng.close()


@pytest.mark.parametrize("plain", [False, True])
def test_trajectory_fp64(oracle_lib, plain):
    print("case: kin_dims=3 fp64 trajectory plain=%s" % plain, flush=True)
    shape = (16, 8, 4, 8)
    eng, orc, p = _pair(oracle_lib, shape, seed=3)
    eng.init_fields(0, 1.0, 0.1)
    orc.init_fields(1.0, 0.1)
    if plain:
        eng.propagate_plain(0, 300, 10)
    else:
        eng.propagate(300, 10)
    orc.propagate(300, 10)
    a, b = eng.get_fields(0)
    c, d = orc.get_fields()
    err = max(np.max(abs(a - c)), np.max(abs(b - d)))
    print("max per-site |delta| =", err)
    assert err <= 1e-10
    assert eng.stats(0).t_langevin == pytest.approx(orc.stats().t_langevin, rel=1e-12)
    og, Fg = eng.observables(0)
    oo, Fo = orc.observables()
    assert og.count == oo.count == 30
    assert og.energy == pytest.approx(oo.energy, rel=1e-10)
    assert og.n_tot == pytest.approx(oo.n_tot, rel=1e-11)
    eng.close()


def test_trajectory_fp32(oracle_lib):
    print("case: kin_dims=3 fp32 trajectory", flush=True)
    shape = (8, 8, 8, 8)
    eng, orc, p = _pair(oracle_lib, shape, precision=CLV_FP32, seed=4)
    eng.init_fields(0, 1.0, 0.1)
    orc.init_fields(1.0, 0.1)
    eng.propagate(500, 10)
    orc.propagate(500, 10)
    a, b = eng.get_fields(0)
    c, d = orc.get_fields()
    rel = max(np.max(abs(a - c)), np.max(abs(b - d))) / np.sqrt(np.mean(abs(c) ** 2))
    print("max per-site |delta| / rms =", rel)
    assert rel <= 1e-4
    og, _ = eng.observables(0)
    oo, _ = orc.observables()
    assert og.energy == pytest.approx(oo.energy, rel=1e-4)
    eng.close()


def test_single_evaluation(oracle_lib):
    print("case: kin_dims=3 observables", flush=True)
    for shape in [(16, 16, 1, 8), (8, 4, 4, 4)]:
        eng, orc, p = _pair(oracle_lib, shape, c1=0.07, P=0.1)
        rng = np.random.default_rng(9)
        dims = eng.dims(0)
        f = 1 + 0.3 * (rng.normal(size=dims) + 1j * rng.normal(size=dims))
        g = 1 + 0.3 * (rng.normal(size=dims) + 1j * rng.normal(size=dims))
        eng.set_fields(0, f, g)
        orc.set_fields(f, g)
        eng.evaluate(0)
        orc.evaluate()
        og, _ = eng.observables(0)
        oo, _ = orc.observables()
        assert og.energy == pytest.approx(oo.energy, rel=1e-12)
        eng.close()


def test_mixed_bucket_and_graphs():
    """kin_dims 1 and 3 chains in one bucket: each equals the chain alone;
    graph replay equals eager steps."""
    print("case: kin_dims mixed bucket", flush=True)
    shape = (16, 8, 2 * 2, 8)
    kd = [3, 1, 3]
    runs = {}
    for graphs in (True, False):
        mix = _engine([shape] * 3, CLV_FP32)
        mix.set_graphs(graphs)
        for i in range(3):
            mix.set_params(i, _params(seed=50 + i, kin_dims=kd[i]))
            mix.init_fields(i, 1.0, 0.1)
        mix.propagate(12, 4)
        runs[graphs] = [mix.get_fields(i) for i in range(3)] + [mix.observables(i)[0].energy for i in range(3)]
        mix.close()
    for i in range(3):
        assert np.array_equal(runs[True][i][0], runs[False][i][0])
        assert runs[True][3 + i] == runs[False][3 + i]
        one = _engine([shape], CLV_FP32)
        one.set_params(0, _params(seed=50 + i, kin_dims=kd[i]))
        one.init_fields(0, 1.0, 0.1)
        one.propagate(12, 4)
        a, b = one.get_fields(0)
        assert np.array_equal(a, runs[True][i][0]) and np.array_equal(b, runs[True][i][1]), i
        assert one.schedule(0) == 1 or kd[i] == 1
        one.close()
