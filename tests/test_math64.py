"""Accuracy of the FP64 noise's logarithm and sine/cosine (clv_math64.h).

normals4<double> (clv_device.h) takes log(u) and sin/cos(2 pi v) of its
Philox uniforms u = (k + 1/2) 2^-32 from clv_math64.h instead of the general
library routines.  The header is plain C++ with every multiply-add an explicit
fma, so the host build below runs the device's operation sequence; it is
checked on 2^21 pseudo-random k and the edges (both ends, the quarter turns,
the sqrt(1/2) split of the logarithm, powers of two) against long-double
references: the logarithm and the sine/cosine within 2 ulp, and sin/cos
within 2^-52 absolute of the long-double angle 2 pi u (a reference that
shares no step with the reduction).  The Box-Muller formula itself follows
the oracle's (Propagator_without_cutoff.cu:432 draws the normals; the
oracle's normals, pinned by moments and a KS test, use libm).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "cl_var_dt_amd", "csrc")


def test_log_and_sincos_within_two_ulp(tmp_path):
    exe = str(tmp_path / "math64_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-I", CSRC,
                    "-o", exe, os.path.join(HERE, "math64_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vals = dict(zip(out[0::2], out[1::2]))
    print(" ".join(out), flush=True)
    assert int(vals["n"]) > 2_000_000
    assert float(vals["log_ulp"]) <= 2.0
    assert float(vals["sin_ulp"]) <= 2.0
    assert float(vals["cos_ulp"]) <= 2.0
    assert float(vals["sincos_abs"]) <= 2.0 ** -52


@pytest.mark.parametrize("planted", ["log_coef", "quarter_turn"])
def test_check_catches_planted_mistakes(tmp_path, planted):
    """The same check on a copy of the header with one plausible mistake:
    a wrong series coefficient (1/9 for 1/7), or the cosine's quarter-turn
    sign taken from j instead of j + 1."""
    src = open(os.path.join(CSRC, "clv_math64.h")).read()
    if planted == "log_coef":
        bad = src.replace("q = fma(q, z, 1.0 / 7.0);", "q = fma(q, z, 1.0 / 9.0);")
    else:
        bad = src.replace("*cp = ((j + 1) & 2) ? -b : b;", "*cp = (j & 2) ? -b : b;")
    assert bad != src
    d = tmp_path / "inc"
    d.mkdir()
    (d / "clv_math64.h").write_text(bad)
    exe = str(tmp_path / "math64_bad")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(d),
                    "-o", exe, os.path.join(HERE, "math64_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vals = dict(zip(out[0::2], out[1::2]))
    worst = max(float(vals["log_ulp"]), float(vals["sin_ulp"]), float(vals["cos_ulp"]))
    assert worst > 2.0 or float(vals["sincos_abs"]) > 2.0 ** -52
