"""Builds the native libraries in-tree.

    python -m kutta_amd.build          # libkutta_hip.so + libkutta_amd.so + oracle

libkutta_hip.so  hand-written gfx950 HIP kernels and the device runtime (hipcc)
libkutta_amd.so  C++ host code (AMG setup, Krylov driver, mesh) over that ABI (g++)
oracle/          the CPU oracle used by tests and the CPU baseline (make)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "kutta_amd"
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["kh_runtime.hip", "kh_linalg.hip", "kh_fv.hip", "kh_gmres.hip", "kh_dense.hip",
               "kh_learned.hip"]
HOST_SOURCES = ["amg_setup.cpp", "krylov.cpp", "mesh.cpp", "pressure.cpp", "simplec.cpp", "capi.cpp",
                "capi_mesh.cpp", "capi_simplec.cpp", "learned.cpp", "capi_learned.cpp"]

# Device arithmetic contract: no contraction of a*b+c, IEEE division and sqrt,
# denormals kept, packed FP32 instructions disabled.
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++20", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wextra"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def _hip_include() -> list[str]:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return [f"-I{rocm}/include", "-D__HIP_PLATFORM_AMD__"]


def build_hip() -> Path:
    out = PKG / "libkutta_hip.so"
    srcs = [str(CSRC / s) for s in HIP_SOURCES if (CSRC / s).exists()]
    _run(["hipcc", *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", str(out), *srcs])
    return out


def build_host() -> Path:
    out = PKG / "libkutta_amd.so"
    host = CSRC / "host"
    srcs = [str(host / s) for s in HOST_SOURCES if (host / s).exists()]
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", f"-I{host}", "-o", str(out), *srcs,
          f"-L{PKG}", "-lkutta_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{rocm}/lib"])
    return out


def build_oracle() -> None:
    jobs = os.environ.get("MAX_JOBS", "8")
    _run(["make", "-C", str(ROOT / "oracle"), f"-j{min(int(jobs), 16)}"])


def build_all() -> None:
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        for w in what:
            {"hip": build_hip, "host": build_host, "oracle": build_oracle}[w]()
