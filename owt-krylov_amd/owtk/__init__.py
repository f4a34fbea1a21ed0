"""owtk - MI355X (gfx950) execution policy for the OWT-Krylov block-CSR solve path.

Thin Python host layer over the C ABI of libowtk_hip.so (include/owtk_hip.h):
device context and buffers, the three block-CSR matrix forms, preconditioners
and device-resident solvers, plus the synthetic-grid generators.  The compute
path is the HIP library; nothing here computes on the CPU besides set-up.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (BF16, F32, F64, LAW_EDGE_CONTRAST, LAW_LOGNORMAL_FIELD, LAW_UNIFORM, MATRIX_BSR,
                   MATRIX_CSR, MATRIX_SPLIT, ORDER_HILBERT, ORDER_IDENTITY, ORDER_RCM, PC_IDENTITY, PC_ILU,
                   PC_JACOBI, PC_LEARNED, PC_SSOR, REDUCTION_COMPENSATED, REDUCTION_MIXED_PRECISION,
                   REDUCTION_SERIAL, InvalidArgumentError, NoDeviceError, OwtkError,
                   UnsupportedError, ZeroPivotError, check, lib)

__all__ = [
    "Context", "DeviceBuffer", "Matrix", "Preconditioner", "SolverOptions", "bicgstab", "pipelined_bicgstab",
    "communication_hiding_bicgstab", "idrs", "Multigrid", "RecycleSpace", "recycled_fgmres", "fgmres",
    "ArnoldiSnapshot", "gcrodr", "extract_harmonic_ritz", "harmonic_ggev", "fgmres_refined", "SplitSystem", "generate_split", "generate_bsr", "split_to_csr", "read_state", "device_count",
    "F64", "F32", "BF16", "OwtkError", "ZeroPivotError", "InvalidArgumentError", "NoDeviceError",
    "UnsupportedError", "lib", "STATUS", "REASON",
]

STATUS = ["converged", "maximum_iterations", "breakdown", "diverged", "invalid_input"]
REASON = ["none", "arnoldi_invariant_subspace", "singular_projected_system", "biorthogonality_loss",
          "alpha_denominator", "omega_denominator", "omega_zero", "idr_shadow_space", "idr_small_system",
          "non_finite_scalar"]

_NP = {F64: np.float64, F32: np.float32, BF16: np.float32}


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().owtk_device_count(C.byref(n)))
    return n.value


class Context:
    """One GPU, one HIP stream (owtk_context)."""

    def __init__(self, device: int = 0):
        self.handle = C.c_void_p()
        check(lib().owtk_context_create(device, C.byref(self.handle)))

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_context_destroy(self.handle))
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self) -> None:
        check(lib().owtk_context_synchronize(self.handle))

    # live kernel timing (HIP events on this context's stream)
    PROF_SPMV_F64, PROF_SPMV_F32, PROF_SPMV_BF16, PROF_PRECOND_F64, PROF_PRECOND_LOW, PROF_LEARNED, \
        PROF_LEARNED_ENDS = range(7)

    def profile(self, enable: bool = True) -> None:
        check(lib().owtk_profile_enable(self.handle, int(enable)))

    def profile_reset(self) -> None:
        check(lib().owtk_profile_reset(self.handle))

    def region_begin(self) -> None:
        check(lib().owtk_profile_region(self.handle, 1, None))

    def region_end(self) -> float:
        """ms on the context stream since region_begin (HIP events)."""
        ms = C.c_double()
        check(lib().owtk_profile_region(self.handle, 0, C.byref(ms)))
        return ms.value

    def profile_read(self, category: int) -> dict:
        n, ms, b = C.c_uint64(), C.c_double(), C.c_int64()
        check(lib().owtk_profile_read(self.handle, category, C.byref(n), C.byref(ms), C.byref(b)))
        return {"launches": n.value, "total_ms": ms.value, "bytes": b.value,
                "avg_ms": ms.value / n.value if n.value else float("nan"),
                "gbps": (b.value / (ms.value * 1e-3) / 1e9) if ms.value > 0 else float("nan")}

    def empty(self, count: int, dtype=np.float64) -> "DeviceBuffer":
        return DeviceBuffer(self, count, dtype)

    def upload(self, array: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(array)
        buf = DeviceBuffer(self, a.size, a.dtype)
        buf.upload(a)
        return buf

    def dot(self, a: "DeviceBuffer", b: "DeviceBuffer", count: int | None = None, reduction: int | None = None) -> float:
        """<a, b>; reduction (REDUCTION_*) selects owtk_reduction_dot."""
        out = C.c_double()
        n = a.count if count is None else count
        if reduction is None:
            check(lib().owtk_dot(self.handle, a.owtk_dtype, n, a.ptr, b.ptr, C.byref(out)))
        else:
            check(lib().owtk_reduction_dot(self.handle, reduction, a.owtk_dtype, n, a.ptr, b.ptr, C.byref(out)))
        return out.value

    def norm(self, a: "DeviceBuffer", count: int | None = None, reduction: int | None = None) -> float:
        out = C.c_double()
        n = a.count if count is None else count
        if reduction is None:
            check(lib().owtk_norm(self.handle, a.owtk_dtype, n, a.ptr, C.byref(out)))
        else:
            check(lib().owtk_reduction_norm(self.handle, reduction, a.owtk_dtype, n, a.ptr, C.byref(out)))
        return out.value

    def axpy(self, alpha: float, x: "DeviceBuffer", y: "DeviceBuffer") -> None:
        check(lib().owtk_axpy(self.handle, x.count, float(alpha), x.ptr, y.ptr))

    def axpy_dot(self, alpha: float, x: "DeviceBuffer", y: "DeviceBuffer", w: "DeviceBuffer") -> float:
        out = C.c_double()
        check(lib().owtk_axpy_dot(self.handle, x.count, float(alpha), x.ptr, y.ptr, w.ptr, C.byref(out)))
        return out.value

    def multidot(self, basis: list["DeviceBuffer"], w: "DeviceBuffer") -> np.ndarray:
        k = len(basis)
        ptrs = (C.c_void_p * max(k, 1))(*[b.ptr for b in basis])
        out = np.zeros(k + 1)
        check(lib().owtk_multidot(self.handle, w.count, k, ptrs, w.ptr, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out


class DeviceBuffer:
    """Device array owned by a context."""

    def __init__(self, ctx: Context, count: int, dtype=np.float64):
        self.ctx = ctx
        self.count = int(count)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.count * self.dtype.itemsize
        self._ptr = C.c_void_p()
        check(lib().owtk_malloc(ctx.handle, max(self.nbytes, 1), C.byref(self._ptr)))
        if self.nbytes:
            check(lib().owtk_fill_zero(ctx.handle, self._ptr, self.nbytes))

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    @property
    def owtk_dtype(self) -> int:
        return F64 if self.dtype == np.float64 else F32

    def upload(self, array: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(array, dtype=self.dtype)
        if a.size != self.count:
            raise ValueError(f"size mismatch: {a.size} != {self.count}")
        check(lib().owtk_memcpy_htod(self.ctx.handle, self._ptr, _ptr(a), self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        check(lib().owtk_memcpy_dtoh(self.ctx.handle, _ptr(out), self._ptr, self.nbytes))
        return out

    def zero(self) -> None:
        check(lib().owtk_fill_zero(self.ctx.handle, self._ptr, self.nbytes))

    def free(self) -> None:
        if self._ptr:
            check(lib().owtk_free(self.ctx.handle, self._ptr))
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            if self._ptr and self.ctx.handle:
                lib().owtk_free(self.ctx.handle, self._ptr)
        except Exception:
            pass


class Matrix:
    """Device copy of one of the reference's three block-CSR forms."""

    def __init__(self, ctx: Context, handle: C.c_void_p, kind: int, owned: int, ghost: int, block: int,
                 keep: tuple):
        self.ctx, self.handle, self.kind = ctx, handle, kind
        self.owned, self.ghost, self.block = owned, ghost, block
        self._keep = keep

    @classmethod
    def split(cls, ctx: Context, owned: int, ghost: int, block: int, offsets, columns, diagonal, off_diagonal):
        o = np.ascontiguousarray(offsets, np.int32)
        c = np.ascontiguousarray(columns, np.int32)
        d = np.ascontiguousarray(diagonal, np.float64)
        v = np.ascontiguousarray(off_diagonal, np.float64)
        h = C.c_void_p()
        check(lib().owtk_matrix_create_split(ctx.handle, owned, ghost, block, _ptr(o), _ptr(c), _ptr(d), _ptr(v),
                                             C.byref(h)))
        return cls(ctx, h, MATRIX_SPLIT, owned, ghost, block, (o, c))

    @classmethod
    def csr(cls, ctx: Context, owned: int, ghost: int, block: int, offsets, columns, values):
        o = np.ascontiguousarray(offsets, np.int32)
        c = np.ascontiguousarray(columns, np.int32)
        v = np.ascontiguousarray(values, np.float64)
        h = C.c_void_p()
        check(lib().owtk_matrix_create_csr(ctx.handle, owned, ghost, block, _ptr(o), _ptr(c), _ptr(v), C.byref(h)))
        return cls(ctx, h, MATRIX_CSR, owned, ghost, block, (o, c))

    @classmethod
    def bsr(cls, ctx: Context, owned: int, ghost: int, block: int, offsets, columns, values):
        o = np.ascontiguousarray(offsets, np.int32)
        c = np.ascontiguousarray(columns, np.int32)
        v = np.ascontiguousarray(values, np.float64)
        h = C.c_void_p()
        check(lib().owtk_matrix_create_bsr(ctx.handle, owned, ghost, block, _ptr(o), _ptr(c), _ptr(v), C.byref(h)))
        return cls(ctx, h, MATRIX_BSR, owned, ghost, block, (o, c))

    @property
    def owned_size(self) -> int:
        return self.owned * self.block

    @property
    def local_size(self) -> int:
        return (self.owned + self.ghost) * self.block

    def info(self) -> _lib.MatrixInfo:
        info = _lib.MatrixInfo()
        check(lib().owtk_matrix_info_get(self.handle, C.byref(info)))
        return info

    def add_storage(self, dtype: int) -> None:
        check(lib().owtk_matrix_add_storage(self.handle, dtype))

    def update_values(self, values, off_diagonal=None) -> None:
        v = np.ascontiguousarray(values, np.float64)
        if off_diagonal is None:
            check(lib().owtk_matrix_update_values(self.handle, _ptr(v), None))
        else:
            e = np.ascontiguousarray(off_diagonal, np.float64)
            check(lib().owtk_matrix_update_values(self.handle, _ptr(v), _ptr(e)))

    def spmv(self, x: DeviceBuffer, y: DeviceBuffer, dtype: int = F64) -> None:
        check(lib().owtk_spmv(self.handle, dtype, x.ptr, y.ptr))

    def residual(self, b: DeviceBuffer, x: DeviceBuffer, r: DeviceBuffer) -> None:
        check(lib().owtk_residual(self.handle, b.ptr, x.ptr, r.ptr))

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_matrix_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            if self.handle:
                lib().owtk_matrix_destroy(self.handle)
        except Exception:
            pass


class RecycleSpace:
    """RecycleSpace<double> (recycling.hpp:24-187) on the device: U and
    C = A U, C orthonormal, at most `capacity` pairs, kept across solves."""

    def __init__(self, layout: Matrix, capacity: int):
        self.matrix = layout
        self.handle = C.c_void_p()
        check(lib().owtk_recycle_space_create(layout.handle, capacity, C.byref(self.handle)))

    def size(self) -> int:
        n = C.c_int64()
        check(lib().owtk_recycle_space_size(self.handle, C.byref(n)))
        return n.value

    def clear(self) -> None:
        check(lib().owtk_recycle_space_clear(self.handle))

    def get(self, ctx: "Context", i: int):
        """(U_i, C_i) as host arrays."""
        n = self.matrix.owned_size
        u, c = ctx.empty(n), ctx.empty(n)
        check(lib().owtk_recycle_space_get(self.handle, i, u.ptr, c.ptr))
        return u.download(), c.download()

    def add_candidate(self, candidate: DeviceBuffer, tol: float = 128 * np.finfo(np.float64).eps) -> tuple:
        added = C.c_int32()
        tel = _lib.SolverResult()
        check(lib().owtk_recycle_space_add_candidate(self.handle, self.matrix.handle, candidate.ptr, tol,
                                                     C.byref(added), C.byref(tel)))
        return bool(added.value), tel.as_dict()

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_recycle_space_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def harmonic_ggev(harmonic, metric):
    """harmonic y = theta metric y (harmonic SPD) in LAPACK dggev's output
    convention: (real, imag, beta, right) with theta = (real + i imag) / beta
    and conjugate pairs as consecutive (real, imaginary) columns of right."""
    g = np.asfortranarray(harmonic, np.float64)
    mm = np.asfortranarray(metric, np.float64)
    n = g.shape[0]
    re, im, be = np.zeros(n), np.zeros(n), np.zeros(n)
    right = np.zeros((n, n), order="F")
    dp = lambda a: a.ctypes.data_as(_lib._DP)
    check(lib().owtk_harmonic_ggev(n, dp(g), dp(mm), dp(re), dp(im), dp(be), dp(right)))
    return re, im, be, right


def extract_harmonic_ritz(matrix: Matrix, snapshot: "ArnoldiSnapshot", space: "RecycleSpace") -> tuple:
    """extract_harmonic_ritz (gcrodr.hpp:86-181); returns (size, telemetry)."""
    tel = _lib.SolverResult()
    k = C.c_uint64()
    check(lib().owtk_extract_harmonic_ritz(matrix.handle, snapshot.handle, space.handle, C.byref(tel), C.byref(k)))
    return k.value, tel.as_dict()


def gcrodr(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer, space: "RecycleSpace",
           options=None, snapshot: "ArnoldiSnapshot | None" = None) -> dict:
    """gcrodr (gcrodr.hpp:185-212) on the device."""
    opts = options if options is not None else SolverOptions()
    res = _lib.SolverResult()
    if snapshot is not None:
        snapshot.local = matrix.local_size
    check(lib().owtk_gcrodr(matrix.handle, pc.handle if pc is not None else None, b.ptr, x.ptr, space.handle,
                            C.byref(opts), snapshot.handle if snapshot is not None else None, C.byref(res)))
    d = res.as_dict()
    d["status_name"] = STATUS[d["status"]] if 0 <= d["status"] < len(STATUS) else str(d["status"])
    d["reason_name"] = REASON[d["breakdown_reason"]] if 0 <= d["breakdown_reason"] < len(REASON) else "?"
    return d


def recycled_fgmres(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer,
                    space: RecycleSpace, options=None, update: bool = True,
                    snapshot: "ArnoldiSnapshot | None" = None) -> dict:
    """recycled_fgmres (recycling.hpp:291-360) on the device; `snapshot`
    receives the last FGMRES cycle."""
    opts = options if options is not None else SolverOptions()
    res = _lib.SolverResult()
    if snapshot is not None:
        snapshot.local = matrix.local_size
    check(lib().owtk_recycled_fgmres_with_snapshot(matrix.handle, pc.handle if pc is not None else None, b.ptr, x.ptr,
                                                   space.handle, C.byref(opts), 1 if update else 0,
                                                   snapshot.handle if snapshot is not None else None, C.byref(res)))
    d = res.as_dict()
    d["status_name"] = STATUS[d["status"]] if 0 <= d["status"] < len(STATUS) else str(d["status"])
    d["reason_name"] = REASON[d["breakdown_reason"]] if 0 <= d["breakdown_reason"] < len(REASON) else "?"
    return d


class Multigrid:
    """AggregationMultigrid (multigrid.hpp:41-421) on the device over a
    CSR-kind Matrix.  smoother: MG_JACOBI, MG_RED_BLACK, MG_CHEBYSHEV; solve
    cycles MG_V, MG_W, MG_FULL."""

    def __init__(self, matrix: Matrix, **kw):
        self.matrix = matrix
        o = _lib.MultigridOptions()
        lib().owtk_multigrid_options_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        self.handle = C.c_void_p()
        check(lib().owtk_multigrid_create(matrix.handle, C.byref(o), C.byref(self.handle)))

    def coarse(self):
        """(nodes, row_offsets, columns, values, fine_to_coarse) of the coarse level."""
        n, e = C.c_int64(), C.c_int64()
        check(lib().owtk_multigrid_coarse(self.handle, C.byref(n), C.byref(e), None, None, None, None))
        off = np.zeros(n.value + 1, np.int32)
        col = np.zeros(e.value, np.int32)
        val = np.zeros(e.value * self.matrix.block)
        f2c = np.zeros(self.matrix.owned, np.int64)
        check(lib().owtk_multigrid_coarse(self.handle, C.byref(n), C.byref(e), _ptr(off), _ptr(col), _ptr(val),
                                          _ptr(f2c)))
        return n.value, off, col, val, f2c

    def solve(self, b: DeviceBuffer, x: DeviceBuffer, options=None, cycle: int = 0) -> dict:
        opts = options if options is not None else SolverOptions()
        res = _lib.SolverResult()
        check(lib().owtk_multigrid_solve(self.handle, b.ptr, x.ptr, C.byref(opts), cycle, C.byref(res)))
        d = res.as_dict()
        d["status_name"] = STATUS[d["status"]] if 0 <= d["status"] < len(STATUS) else str(d["status"])
        d["reason_name"] = REASON[d["breakdown_reason"]] if 0 <= d["breakdown_reason"] < len(REASON) else "?"
        return d

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_multigrid_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MG_JACOBI, MG_RED_BLACK, MG_CHEBYSHEV = 0, 1, 2
MG_V, MG_W, MG_FULL = 0, 1, 2


class Batch:
    """Many independent component-diagonal CSR systems of uneven size solved
    together by batched BiCGSTAB + Jacobi (owtk_batch_*).  systems: list of
    (nodes, row_offsets, columns, values) in the reference's BlockCsrMatrix
    form (diagonal entry included)."""

    def __init__(self, ctx: Context, block: int, systems):
        self.ctx, self.block = ctx, block
        nodes = [int(s[0]) for s in systems]
        self.vertex_offsets = np.concatenate([[0], np.cumsum(nodes)]).astype(np.int64)
        offs, cols, vals, base_e = [], [], [], 0
        for k, (n, o, c, v) in enumerate(systems):
            o = np.asarray(o, np.int64)
            offs.append(o[:-1] + base_e)
            cols.append(np.asarray(c, np.int64) + self.vertex_offsets[k])
            vals.append(np.asarray(v, np.float64))
            base_e += int(o[-1])
        self.row_offsets = np.concatenate(offs + [np.array([base_e])]).astype(np.int64)
        self.columns = np.concatenate(cols).astype(np.int32)
        self.values = np.concatenate(vals)
        self.systems = len(systems)
        self.unknowns = int(self.vertex_offsets[-1]) * block
        self.handle = C.c_void_p()
        check(lib().owtk_batch_create(ctx.handle, self.systems, block, _ptr(self.vertex_offsets),
                                      _ptr(self.row_offsets), _ptr(self.columns), _ptr(self.values),
                                      C.byref(self.handle)))

    def set_tile(self, rows: int) -> None:
        check(lib().owtk_batch_set_tile(self.handle, rows))

    def autotune(self, b: DeviceBuffer, iterations: int = 8):
        tile, ms = C.c_int32(), C.c_double()
        check(lib().owtk_batch_autotune(self.handle, b.ptr, iterations, C.byref(tile), C.byref(ms)))
        return tile.value, ms.value

    def set_group_limit(self, max_vertices: int, threads: int = 512) -> None:
        """Systems of at most max_vertices vertices are solved whole by one
        workgroup each, beside the graph replays of the others (0: none)."""
        check(lib().owtk_batch_set_group_limit(self.handle, max_vertices, threads))

    def autotune_groups(self, b: DeviceBuffer, options=None, graph_iterations: int = 8):
        """Times whole solves over group limits and workgroup sizes and keeps
        the fastest; returns (max_vertices, threads, ms)."""
        opts = options if options is not None else SolverOptions()
        limit, threads, ms = C.c_int64(), C.c_int32(), C.c_double()
        check(lib().owtk_batch_autotune_groups(self.handle, b.ptr, C.byref(opts), graph_iterations, C.byref(limit),
                                               C.byref(threads), C.byref(ms)))
        return limit.value, threads.value, ms.value

    def bicgstab(self, b: DeviceBuffer, x: DeviceBuffer, options=None, graph_iterations: int = 8) -> list[dict]:
        opts = options if options is not None else SolverOptions()
        res = (_lib.BatchResult * self.systems)()
        check(lib().owtk_batch_bicgstab(self.handle, b.ptr, x.ptr, C.byref(opts), graph_iterations, res))
        out = []
        for r in res:
            d = {name: getattr(r, name) for name, _ in r._fields_}
            d["status_name"] = STATUS[d["status"]] if 0 <= d["status"] < len(STATUS) else str(d["status"])
            out.append(d)
        return out

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_batch_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GnnModel:
    """Vertex-block graph network of the learned preconditioner
    (owtk_gnn_*).  A fresh model (Glorot-initialised network, zero decoder,
    g = 1) is the plain truncated Neumann series."""

    def __init__(self, ctx: Context, block: int, degree: int):
        self.ctx = ctx
        self.handle = C.c_void_p()
        check(lib().owtk_gnn_create(ctx.handle, block, degree, C.byref(self.handle)))
        b, d, n = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().owtk_gnn_info(self.handle, C.byref(b), C.byref(d), C.byref(n)))
        self.block, self.degree, self.parameter_count = b.value, d.value, n.value

    def get_parameters(self) -> np.ndarray:
        out = np.zeros(self.parameter_count, np.float32)
        check(lib().owtk_gnn_get_parameters(self.handle, _ptr(out)))
        return out

    def set_parameters(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.parameter_count:
            raise ValueError(f"expected {self.parameter_count} parameters, got {p.size}")
        check(lib().owtk_gnn_set_parameters(self.handle, _ptr(p)))

    def coefficients(self, matrix: "Matrix") -> np.ndarray:
        out = np.zeros((self.degree + 1) * matrix.owned * matrix.block, np.float32)
        check(lib().owtk_gnn_coefficients(self.handle, matrix.handle, _ptr(out)))
        return out.reshape(self.degree + 1, matrix.owned, matrix.block)

    def problem(self, matrix: "Matrix", residuals: DeviceBuffer, errors: DeviceBuffer, pairs: int) -> "GnnProblem":
        return GnnProblem(self, matrix, residuals, errors, pairs)

    def zero_grad(self) -> None:
        check(lib().owtk_gnn_zero_grad(self.handle))

    def accumulate(self, problem: "GnnProblem", weight: float = 1.0) -> None:
        check(lib().owtk_gnn_accumulate(self.handle, problem.handle, weight))

    def read(self):
        loss = C.c_double()
        grad = np.zeros(self.parameter_count, np.float32)
        check(lib().owtk_gnn_read(self.handle, C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> float:
        loss = C.c_double()
        check(lib().owtk_gnn_adam_step(self.handle, lr, beta1, beta2, eps, C.byref(loss)))
        return loss.value

    def save(self, path) -> None:
        np.savez(path, block=self.block, degree=self.degree, params=self.get_parameters())

    def decoder_slice(self) -> slice:
        """Position of W_out and b_out in the flat parameter vector (the
        vector ends with W_out[(degree+1)b][32], b_out, g)."""
        outs = (self.degree + 1) * self.block
        end = self.parameter_count - outs
        return slice(end - outs * 33, end)

    def constant_polynomial(self, matrix: "Matrix") -> "GnnModel":
        """A model whose coefficients are this model's gamma on `matrix`
        averaged over the vertices (per degree and component): the best
        single polynomial the network's output implies on that matrix, the
        ablation that isolates the per-vertex modulation."""
        gam = self.coefficients(matrix).astype(np.float64).mean(axis=1).ravel()
        p = self.get_parameters()
        p[self.decoder_slice()] = 0.0
        p[-gam.size:] = gam.astype(np.float32)
        out = GnnModel(self.ctx, self.block, self.degree)
        out.set_parameters(p)
        return out

    def without_network(self) -> "GnnModel":
        """The same model with its decoder zeroed: gamma = g at every vertex,
        the learned global polynomial alone (the ablation of the network)."""
        p = self.get_parameters()
        p[self.decoder_slice()] = 0.0
        out = GnnModel(self.ctx, self.block, self.degree)
        out.set_parameters(p)
        return out

    @classmethod
    def load(cls, ctx: Context, path) -> "GnnModel":
        with np.load(path, allow_pickle=False) as f:
            model = cls(ctx, int(f["block"]), int(f["degree"]))
            model.set_parameters(f["params"])
        return model

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_gnn_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GnnProblem:
    """One training instance: a split matrix and its (residual, error) pairs,
    resident on the device (owtk_gnn_problem_*)."""

    def __init__(self, model: GnnModel, matrix: "Matrix", residuals: DeviceBuffer, errors: DeviceBuffer,
                 pairs: int):
        self.model, self.matrix = model, matrix
        self.handle = C.c_void_p()
        check(lib().owtk_gnn_problem_create(model.handle, matrix.handle, pairs, residuals.ptr, errors.ptr,
                                            C.byref(self.handle)))

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_gnn_problem_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Preconditioner:
    def __init__(self, matrix: Matrix, kind: int, *, storage: int = F64, omega: float = 1.0,
                 minimum_diagonal: float = 0.0, block_nodes: int = 0, fill_level: int = 0, model=None,
                 schedule: int = 0):
        self.matrix = matrix
        p = _lib.PrecondParams()
        lib().owtk_precond_params_default(C.byref(p), kind)
        p.storage, p.omega, p.minimum_diagonal = storage, omega, minimum_diagonal
        p.block_nodes, p.fill_level, p.schedule = block_nodes, fill_level, schedule
        if model is not None:
            p.learned_model = model.handle
        self._model = model
        self.handle = C.c_void_p()
        check(lib().owtk_precond_create(matrix.handle, C.byref(p), C.byref(self.handle)))
        self.kind, self.storage = kind, storage

    def apply(self, x: DeviceBuffer, y: DeviceBuffer) -> None:
        check(lib().owtk_precond_apply(self.handle, x.ptr, y.ptr))

    def update_values(self) -> None:
        check(lib().owtk_precond_update_values(self.handle))

    def numeric_updates(self) -> int:
        n = C.c_uint64()
        check(lib().owtk_precond_numeric_updates(self.handle, C.byref(n)))
        return n.value

    def info(self) -> _lib.PrecondInfo:
        info = _lib.PrecondInfo()
        check(lib().owtk_precond_info_get(self.handle, C.byref(info)))
        return info

    def write_factors(self, path) -> None:
        """ILU factors in the reference's ilu_fill.cpp output format."""
        check(lib().owtk_write_ilu_factors(self.handle, str(path).encode()))

    def ilu_factors(self):
        n = C.c_int64()
        check(lib().owtk_precond_ilu_factors(self.handle, C.byref(n), None, None, None, None))
        rows = self.matrix.owned
        bs = self.matrix.block
        off = np.zeros(rows + 1, np.int64)
        col = np.zeros(n.value, np.int64)
        val = np.zeros(n.value * bs)
        inv = np.zeros(rows * bs)
        check(lib().owtk_precond_ilu_factors(self.handle, C.byref(n), _ptr(off), _ptr(col), _ptr(val), _ptr(inv)))
        return off, col, val, inv

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_precond_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            if self.handle:
                lib().owtk_precond_destroy(self.handle)
        except Exception:
            pass


def SolverOptions(**kw) -> _lib.SolverOptions:
    """owtk_solver_options with the reference's defaults.  convergence_test
    may be a Python callable test(iteration, device_pointer, count) -> bool
    (SolverOptions::convergence_test); the ctypes callback is kept on the
    options object."""
    o = _lib.SolverOptions()
    lib().owtk_solver_options_default(C.byref(o))
    for k, v in kw.items():
        if k == "convergence_test" and v is not None and not isinstance(v, _lib.CONVERGENCE_TEST):
            test = v

            def bridge(_user, iteration, ptr, count, _test=test):
                return 1 if _test(int(iteration), ptr, int(count)) else 0

            v = _lib.CONVERGENCE_TEST(bridge)
            o._keep_test = v
        setattr(o, k, v)
    return o


def _solve(fn, matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer, options) -> dict:
    opts = options if options is not None else SolverOptions()
    res = _lib.SolverResult()
    check(fn(matrix.handle, pc.handle if pc is not None else None, b.ptr, x.ptr, C.byref(opts), C.byref(res)))
    d = res.as_dict()
    d["status_name"] = STATUS[d["status"]] if 0 <= d["status"] < len(STATUS) else str(d["status"])
    d["reason_name"] = REASON[d["breakdown_reason"]] if 0 <= d["breakdown_reason"] < len(REASON) else "?"
    return d


def bicgstab(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer, options=None) -> dict:
    return _solve(lib().owtk_bicgstab, matrix, pc, b, x, options)


def pipelined_bicgstab(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer,
                       options=None) -> dict:
    """pipelined_bicgstab (pipelined_bicgstab.hpp:20-244) on the device."""
    return _solve(lib().owtk_pipelined_bicgstab, matrix, pc, b, x, options)


def idrs(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer, options=None) -> dict:
    """idrs (idrs.hpp:128-505) on the device; s = options.idr_shadow_space."""
    return _solve(lib().owtk_idrs, matrix, pc, b, x, options)


def communication_hiding_bicgstab(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer,
                                  options=None) -> dict:
    """communication_hiding_bicgstab (pipelined_bicgstab.hpp:317-597) on the device."""
    return _solve(lib().owtk_communication_hiding_bicgstab, matrix, pc, b, x, options)


def fgmres(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer, options=None,
           snapshot: "ArnoldiSnapshot | None" = None) -> dict:
    """fgmres (krylov_solvers.hpp:159-486) on the device; `snapshot` receives
    the last cycle's Arnoldi relation (the arnoldi_snapshot argument)."""
    if snapshot is None:
        return _solve(lib().owtk_fgmres, matrix, pc, b, x, options)
    snapshot.local = matrix.local_size

    def with_snapshot(m, p, bp, xp, o, r):
        return lib().owtk_fgmres_with_snapshot(m, p, bp, xp, o, snapshot.handle, r)

    return _solve(with_snapshot, matrix, pc, b, x, options)


class ArnoldiSnapshot:
    """ArnoldiSnapshot<double> (core.hpp:462-534): the last FGMRES cycle's
    unrotated Hessenberg columns and preconditioned basis vectors, kept on
    the device."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.local = 0
        self.handle = C.c_void_p()
        check(lib().owtk_arnoldi_snapshot_create(ctx.handle, C.byref(self.handle)))

    def columns(self) -> int:
        n = C.c_uint64()
        check(lib().owtk_arnoldi_snapshot_columns(self.handle, C.byref(n)))
        return n.value

    def h(self, row: int, column: int) -> float:
        v = C.c_double()
        check(lib().owtk_arnoldi_snapshot_h(self.handle, row, column, C.byref(v)))
        return v.value

    def hessenberg(self) -> np.ndarray:
        """The (columns + 1) x columns unrotated Hessenberg matrix."""
        k = self.columns()
        return np.array([[self.h(i, j) for j in range(k)] for i in range(k + 1)]).reshape(k + 1, k)

    def basis_pointer(self, column: int) -> int:
        p = C.c_void_p()
        check(lib().owtk_arnoldi_snapshot_basis(self.handle, column, C.byref(p)))
        return p.value

    def preconditioned_basis(self, column: int) -> np.ndarray:
        """Z_column (owned + ghost values) copied to the host."""
        out = np.empty(self.local)
        check(lib().owtk_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(self.basis_pointer(column)), out.nbytes))
        return out

    def clear(self) -> None:
        check(lib().owtk_arnoldi_snapshot_clear(self.handle))

    def begin_cycle(self, layout: "Matrix", maximum_columns: int) -> None:
        """ArnoldiSnapshot::begin_cycle (core.hpp:465-480) for the vector
        layout of `layout`."""
        self.local = layout.local_size
        check(lib().owtk_arnoldi_snapshot_begin_cycle(self.handle, layout.handle, maximum_columns))

    def record_column(self, column: int, values) -> None:
        v = np.ascontiguousarray(values, np.float64)
        check(lib().owtk_arnoldi_snapshot_record_column(self.handle, column, v.ctypes.data_as(_lib._DP), v.size))

    def finish_cycle(self, basis: list["DeviceBuffer"], columns: int | None = None) -> None:
        k = len(basis) if columns is None else columns
        arr = (C.c_void_p * max(k, 1))(*[b.ptr for b in basis[:k]])
        check(lib().owtk_arnoldi_snapshot_finish_cycle(self.handle, arr, k))

    def close(self) -> None:
        if self.handle:
            check(lib().owtk_arnoldi_snapshot_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def refinement_options(**kw) -> _lib.RefinementOptions:
    o = _lib.RefinementOptions()
    lib().owtk_refinement_options_default(C.byref(o))
    for k, v in kw.items():
        if k in ("relative_tolerance", "absolute_tolerance", "maximum_iterations"):
            setattr(o.outer, k, v)
        else:
            setattr(o, k, v)
    return o


def fgmres_refined(matrix: Matrix, pc: Preconditioner | None, b: DeviceBuffer, x: DeviceBuffer,
                   options=None, fallback: Preconditioner | None = None) -> dict:
    opts = options if options is not None else refinement_options()
    res = _lib.SolverResult()
    if fallback is None:
        check(lib().owtk_fgmres_refined(matrix.handle, pc.handle if pc is not None else None, b.ptr, x.ptr,
                                        C.byref(opts), C.byref(res)))
    else:
        check(lib().owtk_fgmres_refined_fallback(matrix.handle, pc.handle if pc is not None else None,
                                                 fallback.handle, b.ptr, x.ptr, C.byref(opts), C.byref(res)))
    d = res.as_dict()
    d["status_name"] = STATUS[d["status"]] if 0 <= d["status"] < len(STATUS) else str(d["status"])
    d["reason_name"] = REASON[d["breakdown_reason"]] if 0 <= d["breakdown_reason"] < len(REASON) else "?"
    return d


# ----------------------------------------------------------------- generators
@dataclass
class SplitSystem:
    nodes: int
    block: int
    offsets: np.ndarray
    columns: np.ndarray
    diagonal: np.ndarray
    off_diagonal: np.ndarray
    rhs: np.ndarray
    coordinates: np.ndarray
    final_to_natural: np.ndarray

    def csr(self):
        return split_to_csr(self.nodes, self.block, self.offsets, self.columns, self.diagonal, self.off_diagonal)


def grid_spec(nx: int, ny: int, block: int, seed: int = 0, ordering: int = ORDER_RCM, permute: bool = True,
              law: int = LAW_UNIFORM, sigma: float = 0.0, correlation: float = 16.0, jitter: float = 0.2,
              shift: float = 0.05) -> _lib.GridSpec:
    s = _lib.GridSpec()
    lib().owtk_grid_spec_default(C.byref(s), nx, ny, block, seed)
    s.ordering, s.permute, s.coefficient_law = ordering, int(permute), law
    s.field_sigma, s.field_correlation, s.jitter, s.diagonal_shift = sigma, correlation, jitter, shift
    return s


def generate_split(nx: int, ny: int, block: int, seed: int = 0, **kw) -> SplitSystem:
    spec = grid_spec(nx, ny, block, seed, **kw)
    n, o = C.c_int64(), C.c_int64()
    check(lib().owtk_grid_sizes(C.byref(spec), C.byref(n), C.byref(o)))
    n, o = n.value, o.value
    offsets = np.zeros(n + 1, np.int32)
    columns = np.zeros(o, np.int32)
    diag = np.zeros(n * block)
    off = np.zeros(o * block)
    rhs = np.zeros(n * block)
    coords = np.zeros(2 * n)
    f2n = np.zeros(n, np.int64)
    check(lib().owtk_grid_generate_split(C.byref(spec), _ptr(offsets), _ptr(columns), _ptr(diag), _ptr(off),
                                         _ptr(rhs), _ptr(coords), _ptr(f2n)))
    return SplitSystem(n, block, offsets, columns, diag, off, rhs, coords.reshape(n, 2), f2n)


def generate_bsr(nx: int, ny: int, block: int, seed: int = 0, **kw):
    spec = grid_spec(nx, ny, block, seed, **kw)
    n, o = C.c_int64(), C.c_int64()
    check(lib().owtk_grid_sizes(C.byref(spec), C.byref(n), C.byref(o)))
    n, o = n.value, o.value
    offsets = np.zeros(n + 1, np.int32)
    columns = np.zeros(o + n, np.int32)
    values = np.zeros((o + n) * block * block)
    rhs = np.zeros(n * block)
    check(lib().owtk_grid_generate_bsr(C.byref(spec), _ptr(offsets), _ptr(columns), _ptr(values), _ptr(rhs)))
    return n, offsets, columns, values, rhs


def read_block_csr(path) -> tuple:
    """(nodes, block, row_offsets, columns, values) of a system file in the
    reference's binary fixture format (owtk_read_block_csr)."""
    n, b, e = C.c_int64(), C.c_int64(), C.c_int64()
    p = str(path).encode()
    check(lib().owtk_read_block_csr(p, C.byref(n), C.byref(b), C.byref(e), None, None, None))
    off = np.zeros(n.value + 1, np.int32)
    col = np.zeros(e.value, np.int32)
    val = np.zeros(e.value * b.value)
    check(lib().owtk_read_block_csr(p, C.byref(n), C.byref(b), C.byref(e), _ptr(off), _ptr(col), _ptr(val)))
    return n.value, b.value, off, col, val


def write_block_csr(path, nodes: int, block: int, offsets, columns, values) -> None:
    o = np.ascontiguousarray(offsets, np.int32)
    c = np.ascontiguousarray(columns, np.int32)
    v = np.ascontiguousarray(values, np.float64)
    check(lib().owtk_write_block_csr(str(path).encode(), nodes, block, _ptr(o), _ptr(c), _ptr(v)))


def read_state(paths, ns: int, nd: int) -> np.ndarray:
    """The application's per-rank final state files (SWSTATE1,
    specwave_limon.py:66-100) as one [vertices, ns*nd] FP64 array ordered by
    global vertex id (owtk_read_state)."""
    paths = [str(p).encode() for p in paths]
    arr = (C.c_char_p * max(1, len(paths)))(*paths)
    n = C.c_int64()
    check(lib().owtk_read_state(arr, len(paths), ns, nd, C.byref(n), None))
    out = np.zeros((n.value, ns * nd))
    check(lib().owtk_read_state(arr, len(paths), ns, nd, C.byref(n), _ptr(out)))
    return out


def split_to_csr(nodes: int, block: int, offsets, columns, diagonal, off_diagonal):
    o = np.ascontiguousarray(offsets, np.int32)
    c = np.ascontiguousarray(columns, np.int32)
    d = np.ascontiguousarray(diagonal, np.float64)
    v = np.ascontiguousarray(off_diagonal, np.float64)
    entries = int(o[nodes]) + nodes
    co = np.zeros(nodes + 1, np.int32)
    cc = np.zeros(entries, np.int32)
    cv = np.zeros(entries * block)
    check(lib().owtk_split_to_csr(nodes, block, _ptr(o), _ptr(c), _ptr(d), _ptr(v), _ptr(co), _ptr(cc), _ptr(cv)))
    return co, cc, cv


def order_rcm(n: int, offsets, columns) -> np.ndarray:
    o = np.ascontiguousarray(offsets, np.int32)
    c = np.ascontiguousarray(columns, np.int32)
    out = np.zeros(n, np.int64)
    check(lib().owtk_order_rcm(n, _ptr(o), _ptr(c), _ptr(out)))
    return out


def order_hilbert(coordinates) -> np.ndarray:
    xy = np.ascontiguousarray(coordinates, np.float64).reshape(-1)
    n = xy.size // 2
    out = np.zeros(n, np.int64)
    check(lib().owtk_order_hilbert(n, _ptr(xy), _ptr(out)))
    return out
