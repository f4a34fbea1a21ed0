"""Build libowtk_hip.so (gfx950) in-tree.

Usage: python owt-krylov_amd/build.py [--jobs N] [--verbose]

Every HIP and host C++ source under csrc/ is compiled with hipcc for gfx950
and linked into owt-krylov_amd/owtk/libowtk_hip.so, next to the Python
package that loads it.  Objects go to owt-krylov_amd/_build/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
OUT = HERE / "owtk" / "libowtk_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every fused multiply-add in the kernels is an explicit
# fma(), so results round exactly like the CPU oracle's explicit std::fma.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
COMMON = [
    "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-I", str(HERE.parent / "include"),
]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip")) + sorted((CSRC / "host").glob("*.cpp"))


def compile_one(src: pathlib.Path, verbose: bool) -> pathlib.Path:
    obj = BUILD / (src.relative_to(CSRC).as_posix().replace("/", "_") + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip() and verbose:
        print(res.stderr, file=sys.stderr)
    return obj


def build(jobs: int = 8, verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(lambda s: compile_one(s, verbose), srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    out = build(args.jobs, args.verbose)
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
