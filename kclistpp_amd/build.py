"""Build libkclpp.so (HIP for gfx950 + host C++) and the CLI programs in-tree.

    python -m kclistpp_amd.build            # library + CLIs
    python -m kclistpp_amd.build --lib-only

Outputs: kclistpp_amd/libkclpp.so, kclistpp_amd/bin/<program>.
Objects go to kclistpp_amd/_build/.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libkclpp.so")
BIN = os.path.join(PKG, "bin")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: fp64 rho arithmetic must not fuse into FMAs.
# -packed-fp32-ops off: no v_pk_*_f32 generated.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC,
]

SOURCES = ["host_io.cpp", "host_order.cpp", "graph.hip", "pass.hip", "plan.hip", "sim.hip", "batch.hip",
           "costmodel.hip", "trim.cpp", "exact.cpp", "noalloc.hip", "sampling.hip"]
CLIS = {
    "kCDensestCmpSync": "cli_cmpsync.c",
    "kClistCoreTrim": "cli_trim.c",
    "kCDensestMem": "cli_exact.c",
    "kCDensestSamplingKclistpp": "cli_sampling.c",
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


# One object per word class of the enumeration kernels (compiled in parallel).
PASS_CLASSES = 9


def _compile(job) -> str:
    src, defines, stem = job
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, stem + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *HIP_FLAGS, *defines, *lang, "-c", path, "-o", obj])
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    work = [(src, [], os.path.splitext(src)[0]) for src in SOURCES]
    work += [("pass_inst.hip", [f"-DKCL_PASS_CLASS={c}"], f"pass_inst_{c}")
             for c in range(PASS_CLASSES)]
    work.sort(key=lambda w: 0 if w[0] == "pass_inst.hip" else 1)  # longest first
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, work))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_clis() -> list[str]:
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name, src in CLIS.items():
        path = os.path.join(CSRC, "cli", src)
        if not os.path.exists(path):
            continue
        exe = os.path.join(BIN, name)
        _run(["gcc", "-O2", "-std=c11", "-Wall", "-I", INCLUDE, path, "-o", exe,
              f"-L{PKG}", "-lkclpp", f"-Wl,-rpath,$ORIGIN/..", "-lm"])
        out.append(exe)
    return out


def build(lib_only: bool = False) -> None:
    build_lib()
    if not lib_only:
        build_clis()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    a = ap.parse_args()
    build(a.lib_only)
    print(f"built {LIB}", flush=True)
    sys.exit(0)
