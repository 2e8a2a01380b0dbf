"""Build libclv.so (gfx950) in-tree with hipcc, and the host driver clv_main.

    python -m cl_var_dt_amd.build          # library + driver
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libclv.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 in the generated code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(extra: list[str] | None = None) -> str:
    srcs = [os.path.join(CSRC, "clv_engine.hip"), os.path.join(CSRC, "clv_server.hip")]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *(extra or []), "-I", os.path.join(ROOT, "include"),
          "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str | None:
    src = os.path.join(PKG, "driver", "clv_main.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(PKG, "driver", "clv_main")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          "-o", out, src, "-L", PKG, "-lclv", "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_all() -> None:
    build_lib()
    build_driver()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
