"""Build recipe for the native libraries (run on any host with ROCm; no GPU needed).

    python -m afquantumsim_amd.build        # or __graft_entry__.build()

Produces, in-tree:
  afquantumsim_amd/libaqs_hip.so   HIP kernels for gfx950 + the C ABI (include/aqs_hip.h)
  afquantumsim_amd/libaqs.so       the aqs:: C++ API over that C ABI (include/aqs/quantum.h)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CPPSRC = os.path.join(PKG, "cpp")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
HIP_LIB = os.path.join(PKG, "libaqs_hip.so")
AQS_LIB = os.path.join(PKG, "libaqs.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions stay off in every device pass.
# -fcx-limited-range: host complex products by the textbook formula (every
# operand is a gate or unitary entry, |x| <= 1) instead of a library call
# that rescues overflow and NaN.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-fcx-limited-range",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]

HIP_SOURCES = ["state.hip", "gates.hip", "mfma.hip", "diag.hip", "perm.hip", "tile.hip", "measure.hip", "batch.hip",
               "fusion.cpp", "tiling.cpp", "serve_core.cpp", "serve.cpp"]
CPP_SOURCES = ["quantum.cpp", "quantum_algo.cpp", "quantum_gates.cpp"]  # algo_bench.cpp is a program


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", src, "-o", _obj(src)])

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, srcs))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[_obj(s) for s in srcs]])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_cpp() -> str:
    srcs = [os.path.join(CPPSRC, s) for s in CPP_SOURCES if os.path.exists(os.path.join(CPPSRC, s))]
    if not srcs:
        return ""
    tmp = AQS_LIB + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-fcx-limited-range", "-shared", "-I", INCLUDE, "-o", tmp, *srcs,
          "-L", PKG, "-laqs_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, AQS_LIB)
    return AQS_LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "api_test.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "api_test")


def build_cpp_tests() -> str:
    """The fixture-driven test program of the aqs:: C++ API (runs on a GPU)."""
    if not os.path.exists(CPP_TEST_SRC):
        return ""
    tmp = CPP_TEST_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-o", tmp, CPP_TEST_SRC, "-L", PKG, "-laqs",
          "-laqs_hip", "-Wl,-rpath,$ORIGIN/../../afquantumsim_amd"])
    os.replace(tmp, CPP_TEST_BIN)
    return CPP_TEST_BIN


ALGO_BENCH_SRC = os.path.join(CPPSRC, "algo_bench.cpp")
ALGO_BENCH_BIN = os.path.join(PKG, "algo_bench")


def build_algo_bench() -> str:
    """The reference's published algorithm benchmarks over the aqs:: API (run by bench.py)."""
    if not os.path.exists(ALGO_BENCH_SRC):
        return ""
    tmp = ALGO_BENCH_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-o", tmp, ALGO_BENCH_SRC, "-L", PKG, "-laqs",
          "-laqs_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, ALGO_BENCH_BIN)
    return ALGO_BENCH_BIN


TILE_EMU_SRC = os.path.join(ROOT, "tests", "cpp", "tile_emu.cpp")
TILE_EMU_LIB = os.path.join(ROOT, "tests", "cpp", "libtile_emu.so")


def build_test_helpers() -> str:
    """Host emulation of compiled schedules for the CPU tests of the tile scheduler."""
    if not os.path.exists(TILE_EMU_SRC):
        return ""
    tmp = TILE_EMU_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-shared", "-o", tmp, TILE_EMU_SRC, "-L", PKG, "-laqs_hip",
          "-Wl,-rpath,$ORIGIN/../../afquantumsim_amd"])
    os.replace(tmp, TILE_EMU_LIB)
    return TILE_EMU_LIB


SERVE_CORE_SRC = os.path.join(CSRC, "serve_core.cpp")
SERVE_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "serve_core_test.cpp")
SERVE_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "serve_core_test")
SERVE_TEST_TSAN_BIN = os.path.join(ROOT, "tests", "cpp", "serve_core_test_tsan")


def build_serve_core_tests() -> str:
    """Host tests of the server's scheduling core (a stand-in backend, no GPU),
    plain and under ThreadSanitizer (host code only: g++, no device code)."""
    if not os.path.exists(SERVE_TEST_SRC):
        return ""
    common = ["-std=c++17", "-Wall", "-I", INCLUDE, "-I", CSRC, SERVE_CORE_SRC, SERVE_TEST_SRC, "-lpthread"]
    for out, flags in ((SERVE_TEST_BIN, ["-O2"]), (SERVE_TEST_TSAN_BIN, ["-O1", "-g", "-fsanitize=thread"])):
        tmp = out + ".tmp"
        _run(["g++", *flags, "-o", tmp, *common])
        os.replace(tmp, out)
    return SERVE_TEST_BIN


def build_all() -> None:
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        raise RuntimeError("hipcc not found")
    build_hip()
    build_cpp()
    build_cpp_tests()
    build_algo_bench()
    build_test_helpers()
    build_serve_core_tests()


if __name__ == "__main__":
    build_all()
    print("built", HIP_LIB, AQS_LIB)
