"""Build the gfx950 library (libgcctb.so), its command-line tools and the CPU oracle.

Every call recompiles all sources (hipcc cross-compiles for gfx950 without a
GPU).  Objects go to gpu_cc_tb_amd/_build/, the library and tools next to this
file, the oracle library to oracle/.
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
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgcctb.so")
BIN = os.path.join(PKG, "bin")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]
DEVICE_SRCS = [
    "device/prep.hip",
    "device/exec_tpl.hip",
    "device/exec_to.hip",
    "device/exec_mvcc.hip",
    "device/exec_silo.hip",
    "device/exec_tictoc.hip",
    "device/exec_ordered.hip",
    "device/exec_row.hip",
    "device/exec_evt.hip",
    "device/exec_row_evt.hip",
    "device/exec_latch.hip",
    "device/exec_tpcc.hip",
    "device/microbench.hip",
    "device/noalloc.hip",
]
HOST_SRCS = ["host/engine.cpp", "host/workload.cpp", "host/check.cpp", "host/output.cpp",
             "host/cpu_path.cpp", "host/tpcc.cpp"]
TOOLS = {"test_ycsb": "tools/test_ycsb.cpp", "gen_ycsb": "tools/gen_ycsb.cpp", "test_tpcc": "tools/test_tpcc.cpp"}
ORACLE_SRCS = ["oracle_core.c", "cpu_cc.c"]


def _run(cmd: list[str], quiet: bool = True) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if out and not quiet:
        print(out)


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".", "_") + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = DEVICE_SRCS + HOST_SRCS
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for name, src in TOOLS.items():
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, src),
              "-o", os.path.join(BIN, name), f"-L{PKG}", "-lgcctb", "-Wl,-rpath,$ORIGIN/..", "-lpthread"])
    return LIB


def build_oracle() -> str:
    out = os.path.join(ORACLE, "liboracle.so")
    srcs = [os.path.join(ORACLE, s) for s in ORACLE_SRCS if os.path.exists(os.path.join(ORACLE, s))]
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-o", out + ".tmp",
          *srcs, "-lm", "-lpthread"])
    os.replace(out + ".tmp", out)
    return out


def build(jobs: int = 8) -> None:
    build_oracle()
    build_library(jobs)


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", LIB)
