"""Build the CPU oracle (test infrastructure only) into oracle/liboracle.so.

g++ -O3 -march=x86-64-v3 (AVX2 + FMA: every x86 host of an MI355X system has
them) -ffp-contract=off: the only fused multiply-adds are the explicit
std::fma calls of owtk_oracle.cpp.
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
SRC = HERE / "owtk_oracle.cpp"
OUT = HERE / "liboracle.so"


def build(verbose: bool = False) -> pathlib.Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-std=c++20", "-fPIC", "-shared",
           "-fopenmp", "-o", str(tmp), str(SRC)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
