"""In-tree build of the package's native code (python -m xfdtd_cuda_amd.build).

libxfdtd_hip.so  hipcc --offload-arch=gfx950: the time-stepping engine.
libxfdtd_host.so g++: scene set-up, SimulationHD drop-in, C ABI for Python.
tests/cpp/simulation_hd_driver  g++: the C++ drop-in driver test.

The FP64 kernels must round like the CPU oracle, so device code is built
with -ffp-contract=off (no fused multiply-add).  Packed FP32 instructions
are disabled with the target feature -packed-fp32-ops.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra",
              "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _run(cmd, cwd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line
                    and "(ignoring feature)" not in line)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_hip():
    src = os.path.join(HERE, "csrc", "xf_engine.hip")
    out = os.path.join(HERE, "libxfdtd_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-o", out + ".tmp", src], HERE)
    os.replace(out + ".tmp", out)


def build_host():
    srcs = [os.path.join(HERE, "csrc", "host", f)
            for f in ("xfdtd_core.cpp", "capi.cpp", "simulation_hd.cpp")]
    out = os.path.join(HERE, "libxfdtd_host.so")
    _run([CXX, *HOST_FLAGS, "-shared", "-o", out + ".tmp", *srcs, "-L" + HERE, "-lxfdtd_hip",
          "-Wl,-rpath,$ORIGIN"], HERE)
    os.replace(out + ".tmp", out)


def build_driver():
    """The C++ driver test of the SimulationHD drop-in (tests/cpp)."""
    src = os.path.join(ROOT, "tests", "cpp", "simulation_hd_driver.cpp")
    if not os.path.exists(src):
        return
    out = os.path.join(ROOT, "tests", "cpp", "simulation_hd_driver")
    _run([CXX, *HOST_FLAGS, "-o", out, src, "-L" + HERE, "-lxfdtd_host", "-lxfdtd_hip",
          "-Wl,-rpath,$ORIGIN/../../xfdtd_cuda_amd"], ROOT)


def build_all():
    build_hip()
    build_host()
    build_driver()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
