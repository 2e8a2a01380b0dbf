"""The C++ host driver (clv_main): run-time Parameters.h / params.txt ingestion,
the reference's stdout strings and output files (main.cu:53-162,
SaveResults.cu:52-607), and -- on a GPU -- its Skalars.txt numbers against the
oracle run through the same main loop."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cl_var_dt_amd", "driver", "clv_main")

PARAMS_H = """#ifndef PARAMETERS_H_
#define PARAMETERS_H_
#define COMPONENTS 4
#define XSIZE 32
#define YSIZE 1
#define ZSIZE 1
#define TAUSIZE 8
#define EPS     0.125        //beta=EPS*TAUSIZE
#define a_x     1.0
#define G       0.10000
#define M       0.00000
#define SIGMA   0.00000
#define D       0.27735
#define OMEGAX 0.
#define OMEGAY 0.
#define OMEGAZ 0.
#define Q 0.0
#define STEPS 120
#define DT 0.002
#define SNAPSHOT_TRIGGER 50
#define EVALUATE_TRIGGER 10
#define BLOCKSIZE 64
#endif
"""


def _run(tmp_path, *args, params="mu=-0.5\nP=0\nrun=1\nseed=42\nseed2=7\n", params_h=PARAMS_H):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    (tmp_path / "Parameters.h").write_text(params_h)
    (tmp_path / "params.txt").write_text(params)
    return subprocess.run([DRIVER, *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)


def test_dry_run_outputs(tmp_path):
    r = _run(tmp_path, "--dry-run")
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert out[0] == 'Directory created successfully: "./DATA/BoseCL_Sim_mu_-0.5_P_0/Run1"'
    assert out[1] == "Saved parameters"
    assert out[2] == "Directory for raw data was successfully created"
    assert out[3] == "Started simulation with parameters"
    assert "T 1" in out and "lattice 32x1x1x8" in out and "seed 42" in out and "steps 120" in out
    run1 = tmp_path / "DATA" / "BoseCL_Sim_mu_-0.5_P_0" / "Run1"
    assert (run1 / "rawdata").is_dir()
    par = (run1 / "Parameters.txt").read_text().split("\n")
    assert par[0] == "T\t1" and par[1] == "mu\t-0.5" and par[2] == "g\t0.1"
    assert "lattice\t32x1x1x8" in par and par[-1] == "seed\t42"
    # a second run goes to Run2 (SaveResults.cu:59-67)
    r = _run(tmp_path, "--dry-run")
    assert '"./DATA/BoseCL_Sim_mu_-0.5_P_0/Run2"' in r.stdout


def test_missing_mu_is_an_error(tmp_path):
    r = _run(tmp_path, "--dry-run", params="P=0\nseed=1\n")
    assert r.returncode == 1
    assert "params.txt" in r.stderr


@pytest.mark.gpu
def test_driver_run_matches_oracle(tmp_path, oracle_lib):
    print("case: driver run", flush=True)
    r = _run(tmp_path, "0")
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[-1].startswith("Finished simulation after computation time of ")
    saved = [ln for ln in lines if ln.startswith("Saved scalars at time ")]
    # snapshots at i = 1, 50, 100 and after the loop (main.cu:146-158)
    assert [int(s.split()[-1]) for s in saved] == [1, 5, 10, 12]
    base = tmp_path / "DATA" / "BoseCL_Sim_mu_-0.5_P_0"
    sk = [ln.split("\t") for ln in (base / "Skalars.txt").read_text().splitlines()]
    assert len(sk) == 4 and all(len(row) == 8 for row in sk)
    fz = (base / "Run1" / "F_z.txt").read_text().splitlines()
    assert len(fz) == 4 and len(fz[0].rstrip("\t").split("\t")) == 32

    # the same main loop on the oracle, from the reference's zero start
    from cl_var_dt_amd import Shape, default_params
    p = default_params(mu=-0.5, P=0.0, G=0.1, Q=0.0, eps=0.125, a_x=1.0, DT=0.002, seed=42)
    o = oracle_lib.Oracle(Shape(32, 1, 1, 8), p, nthreads=2)
    z = np.zeros((4, 8, 32, 1, 1), complex)
    o.set_fields(z, z)
    rows = []
    for i in range(120):
        if (i % 50 == 0 and i not in (0, 119)) or i == 1:
            obs, _ = o.observables()
            rows.append([obs.n_tot, *obs.n_comp])
        o.propagate(1, 0)
        if i % 10 == 0:
            o.evaluate()
    obs, _ = o.observables()
    rows.append([obs.n_tot, *obs.n_comp])
    got = np.array([[float(v) for v in row[:5]] for row in sk])
    assert np.allclose(got, np.array(rows), atol=2e-6)


def _ext_params_h(dt_device):
    """A 3-D lattice with the engine's Parameters.h extensions."""
    h = PARAMS_H.replace("#define XSIZE 32\n#define YSIZE 1\n#define ZSIZE 1\n",
                         "#define XSIZE 16\n#define YSIZE 4\n#define ZSIZE 4\n")
    ext = ("#define C1 0.05\n#define KIN_DIMS 3\n#define CONJ_DRIFT_MODE 1\n#define INIT gaussian\n"
           "#define INIT_N0 1.0\n#define INIT_VAR 0.05\n#define DT_DEVICE %d\n#endif\n" % dt_device)
    return h.replace("#endif\n", ext)


@pytest.mark.gpu
def test_driver_extensions_match_oracle(tmp_path, oracle_lib):
    """The Parameters.h extensions (C1, KIN_DIMS = 3, CONJ_DRIFT_MODE = 1, a
    Gaussian start) on a 16x4x4x8 lattice: host-driven and device-resident dt
    write identical files, and Skalars.txt equals the oracle run through the
    same main loop (main.cu:144-158)."""
    print("case: driver extensions", flush=True)
    outs = []
    for dtd in (0, 1):
        d = tmp_path / ("dt%d" % dtd)
        d.mkdir()
        r = _run(d, "0", params_h=_ext_params_h(dtd))
        assert r.returncode == 0, r.stdout + r.stderr
        base = d / "DATA" / "BoseCL_Sim_mu_-0.5_P_0"
        outs.append([(base / "Skalars.txt").read_text()] +
                    [(base / "Run1" / ("F_%s.txt" % a)).read_text() for a in "xyz"])
    assert outs[0] == outs[1]
    from cl_var_dt_amd import CLV_CONJ_DRIFT_ACTION, Shape, default_params
    p = default_params(mu=-0.5, P=0.0, G=0.1, Q=0.0, c1=0.05, eps=0.125, a_x=1.0, DT=0.002, seed=42,
                       kin_dims=3, conj_drift_mode=CLV_CONJ_DRIFT_ACTION)
    o = oracle_lib.Oracle(Shape(16, 4, 4, 8), p, nthreads=2)
    o.init_fields(1.0, 0.05)
    rows = []
    for i in range(120):
        if (i % 50 == 0 and i not in (0, 119)) or i == 1:
            obs, _ = o.observables()
            rows.append([obs.n_tot, *obs.n_comp])
        o.propagate(1, 0)
        if i % 10 == 0:
            o.evaluate()
    obs, _ = o.observables()
    rows.append([obs.n_tot, *obs.n_comp])
    sk = [ln.split("\t") for ln in outs[0][0].splitlines()]
    got = np.array([[float(v) for v in row[:5]] for row in sk])
    assert got.shape == (4, 5) and np.all(np.isfinite(got))
    assert np.allclose(got, np.array(rows), atol=2e-6)


def _oracle_main_loop(o, steps=120):
    """Skalars rows (n_tot, n_0..n_3) of main.cu's loop (:144-158) on the oracle."""
    rows = []
    for i in range(steps):
        if (i % 50 == 0 and i not in (0, steps - 1)) or i == 1:
            obs, _ = o.observables()
            rows.append([obs.n_tot, *obs.n_comp])
        o.propagate(1, 0)
        if i % 10 == 0:
            o.evaluate()
    obs, _ = o.observables()
    rows.append([obs.n_tot, *obs.n_comp])
    return np.array(rows)


@pytest.mark.gpu
def test_driver_mean_field_start_and_fp32(tmp_path, oracle_lib):
    """INIT mean_field (set_value, Lattice.cu:50-81: sqrt(|mu|/G/4) in components
    0 and 3, 2/3 of it in 1 and 2) against the oracle from the same fields, and
    the same run with FP32 fields (PRECISION 1) within FP32 rounding."""
    print("case: driver mean field + fp32", flush=True)
    got = []
    for prec in (0, 1):
        d = tmp_path / ("p%d" % prec)
        d.mkdir()
        h = PARAMS_H.replace("#endif\n", "#define INIT mean_field\n#define PRECISION %d\n#endif\n" % prec)
        r = _run(d, "0", params_h=h)
        assert r.returncode == 0, r.stdout + r.stderr
        sk = (d / "DATA" / "BoseCL_Sim_mu_-0.5_P_0" / "Skalars.txt").read_text().splitlines()
        got.append(np.array([[float(v) for v in ln.split("\t")[:5]] for ln in sk]))
    from cl_var_dt_amd import Shape, default_params
    p = default_params(mu=-0.5, P=0.0, G=0.1, Q=0.0, eps=0.125, a_x=1.0, DT=0.002, seed=42)
    o = oracle_lib.Oracle(Shape(32, 1, 1, 8), p, nthreads=2)
    amp = np.sqrt(0.5 / 0.1 / 4)
    f = np.zeros((4, 8, 32, 1, 1), complex)
    for c, w in enumerate((1.0, 2.0 / 3.0, 2.0 / 3.0, 1.0)):
        f[c] = w * amp
    o.set_fields(f, np.conj(f))
    want = _oracle_main_loop(o)
    assert got[0].shape == want.shape == (4, 5)
    assert np.allclose(got[0], want, atol=2e-6)
    assert np.allclose(got[1], got[0], rtol=1e-4)


def test_integration_binding_compiles(tmp_path):
    """The documented reference-side C++ binding (the first ```cpp block of the
    integration guide: the reference's Propagator members over clv.h) compiles
    and links against libclv.so as written."""
    lib = os.path.join(ROOT, "cl_var_dt_amd", "libclv.so")
    if not os.path.exists(lib):
        pytest.skip("libclv.so not built")
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = doc.index("```cpp\n") + len("```cpp\n")
    code = doc[start:doc.index("```", start)]
    src = tmp_path / "binding.cpp"
    src.write_text("#include <iostream>\n"
                   "#define XSIZE 64\n#define YSIZE 1\n#define ZSIZE 1\n#define TAUSIZE 32\n"
                   "static const double G = 0.1, Q = 0.1, EPS = 1.0 / 32, a_x = 1.0, DT = 1e-3;\n"
                   + code +
                   "\nint main() {\n  Propagator p(1234ULL, 0.0, 1.0);\n  p.propagate(0.0, 1.0);\n"
                   "  p.propagate_without_noise(0.5, 0.0, 1.0);\n  return 0;\n}\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", str(tmp_path / "binding"), str(src), "-L", os.path.dirname(lib), "-lclv"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
