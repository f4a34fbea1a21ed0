"""The C ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import pathlib
import re

import owtk
from owtk import _lib

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "owtk_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(owtk_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert len(names) > 40
    assert "owtk_bicgstab" in names and "owtk_spmv" in names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.PROTOTYPES)


def test_no_device_is_reported_cleanly():
    if owtk.device_count() == 0:
        try:
            owtk.Context(0)
        except owtk.NoDeviceError:
            pass
        else:
            raise AssertionError("context creation without a device must fail")


def test_options_defaults_mirror_reference():
    # SolverOptions<double> defaults (core.hpp:88-111)
    o = owtk.SolverOptions()
    assert o.relative_tolerance == 1e-8 and o.absolute_tolerance == 0.0
    assert o.maximum_iterations == 1000 and o.convergence_check_interval == 10
    assert o.restart == 30 and o.residual_replacement_interval == 50
    assert o.gmres_orthogonalization == 0
    assert o.breakdown_tolerance == 64 * 2.220446049250313e-16
    assert o.idr_shadow_space == 4 and o.relaxation == 1.0 and o.verify_true_residual == 1
    assert o.reduction == owtk.REDUCTION_SERIAL and not o.convergence_test


def test_cpp_header_compiles():
    """include/owtk/hip_policy.hpp and the C header compile as C++20 and C."""
    import subprocess
    root = pathlib.Path(__file__).resolve().parent.parent
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I", str(root / "include"),
                    str(root / "tests" / "cpp" / "test_hip_policy.cpp")], check=True)
    subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", str(root / "include" / "owtk_hip.h")],
                   check=True)
