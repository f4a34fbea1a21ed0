"""Build the CPU oracle (test infrastructure only): oracle/liboracle.so and
oracle/liboracle_v4.so.

g++ -O3 -ffp-contract=off -fopenmp: the only fused multiply-adds are the
explicit std::fma calls of owtk_oracle.cpp.  liboracle.so targets x86-64-v3
(AVX2 + FMA: every x86 host of an MI355X system has them) and is the one the
tests load; liboracle_v4.so targets x86-64-v4 (AVX-512) and is what the CPU
baseline uses on hosts with AVX-512 (the reference's benchmark asks for
-march=native; a native build here would target this machine's CPU, not the
GPU host's).
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
SRC = HERE / "owtk_oracle.cpp"
OUT = HERE / "liboracle.so"


OUT_V4 = HERE / "liboracle_v4.so"


def _build_one(out: pathlib.Path, march: str, verbose: bool) -> None:
    tmp = out.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", f"-march={march}", "-ffp-contract=off", "-std=c++20", "-fPIC", "-shared",
           "-fopenmp", "-o", str(tmp), str(SRC)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = False) -> pathlib.Path:
    _build_one(OUT, "x86-64-v3", verbose)
    _build_one(OUT_V4, "x86-64-v4", verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
