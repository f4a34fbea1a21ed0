"""The header-only C++ host layer (include/owtk/hip_policy.hpp) on the GPU:
HipExecutionPolicy::apply bitwise against a host loop, and the device solvers
called unqualified in the reference's argument order (FGMRES + block-Jacobi
ILU(0), refined FGMRES, compensated BiCGSTAB, an application convergence
test, pipelined BiCGSTAB + SSOR), through stand-in types with the reference's
names (tests/cpp/test_hip_policy.cpp)."""
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = pathlib.Path(__file__).resolve().parent / "cpp" / "test_hip_policy"


@pytest.mark.skipif(not BIN.exists(), reason="C++ test not built (run __graft_entry__.build())")
def test_cpp_policy_layer():
    out = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "apply bitwise ok" in out.stdout
    assert "reused address and in-place value change ok" in out.stdout
    assert "all C++ checks passed" in out.stdout
