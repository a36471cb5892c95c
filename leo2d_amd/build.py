"""Build libleo2d.so (HIP kernels + C-ABI + host C++) and the `leo2d` CLI for gfx950.

Outputs stay in-tree: leo2d_amd/lib/libleo2d.so and leo2d_amd/bin/leo2d
(object files in leo2d_amd/_build/).  Compiles every time it is called.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libleo2d.so")
BIN = os.path.join(HERE, "bin", "leo2d")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LEO2D_ARCH", "gfx950")

SOURCES = ["kpm.hip", "sell.hip", "scratch.hip", "dgemm.hip", "batch.hip", "mlp.hip", "dos.hip", "cond.hip", "assemble.hip", "capi.cpp", "surrogate.cpp", "deck.cpp", "geom.cpp"]
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",                               # reference operation order, no contraction
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(SRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cli_obj = _compile("cli.cpp")
    r = subprocess.run([HIPCC, "-o", BIN, cli_obj, "-L" + os.path.dirname(LIB), "-lleo2d",
                        "-Wl,-rpath,$ORIGIN/../lib"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli link failed\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", LIB, BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
