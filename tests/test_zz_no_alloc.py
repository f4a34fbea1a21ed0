"""owtk_multidot_noalloc (the multi-dot for callers that cannot allocate)
against owtk_multidot through the C ABI: bitwise the same results on two
calls in a row, the second with a smaller input, plus the argument checks.
The entry point runs on the null stream of the current device, so the test
waits with hipDeviceSynchronize from the same HIP runtime the library uses
(the context's own stream does not order against the null stream)."""
import ctypes as C

import numpy as np
import pytest

import owtk
from owtk import _lib

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipDeviceSynchronize.restype = C.c_int
    return hip


def _noalloc(n, basis, w, out):
    k = len(basis)
    ptrs = (C.c_void_p * max(k, 1))(*[b.ptr for b in basis])
    return _lib.lib().owtk_multidot_noalloc(None, n, k, ptrs, w.ptr if w is not None else None,
                                            out.ptr if out is not None else None)


def _case(ctx, hip, rng, n, k):
    basis = [ctx.upload(rng.uniform(-1.0, 1.0, n)) for _ in range(k)]
    w = ctx.upload(rng.uniform(-1.0, 1.0, n))
    expected = ctx.multidot(basis, w)          # synchronises the context stream
    out = ctx.empty(k + 1)
    ctx.synchronize()
    print(f"noalloc multidot n={n} k={k}", flush=True)
    rc = _noalloc(n, basis, w, out)
    assert rc == 0, _lib.lib().owtk_last_error().decode()
    assert hip.hipDeviceSynchronize() == 0
    got = out.download()
    np.testing.assert_array_equal(got, expected)
    # and against long-double sums: a thread adds a few products, then the
    # block and finisher trees add about 30 levels, so 64 eps * sum |w v|
    # bounds the FP64 rounding
    host_w = w.download().astype(np.longdouble)
    vecs = [b.download().astype(np.longdouble) for b in basis] + [host_w]
    exact = np.array([float(np.sum(host_w * v)) for v in vecs])
    bound = np.array([float(np.sum(np.abs(host_w * v))) for v in vecs]) * 64 * np.finfo(np.float64).eps
    assert np.all(np.abs(got - exact) <= bound)


def test_noalloc_multidot_matches_context_multidot_twice():
    hip = _hip()
    rng = np.random.default_rng(7)
    with owtk.Context(0) as ctx:
        # first call fills all 1024 blocks of the static scratch and 31 values
        _case(ctx, hip, rng, 1_500_003, 30)
        # second call: smaller input, fewer blocks and values; stale partials of
        # the first call stay in the array and must not leak into the result
        _case(ctx, hip, rng, 70_001, 5)


def test_noalloc_multidot_empty_and_argument_checks():
    hip = _hip()
    with owtk.Context(0) as ctx:
        out = ctx.upload(np.full(4, 7.0))
        ctx.synchronize()
        basis = [ctx.empty(1) for _ in range(3)]
        assert _noalloc(0, basis, None, out) == 0
        assert hip.hipDeviceSynchronize() == 0
        np.testing.assert_array_equal(out.download(), np.zeros(4))
        w = ctx.empty(10)
        many = [ctx.empty(10) for _ in range(63)]
        assert _noalloc(10, many, w, ctx.empty(64)) == _lib.ERROR_INVALID_ARGUMENT
        assert _noalloc(10, basis, w, None) == _lib.ERROR_INVALID_ARGUMENT
