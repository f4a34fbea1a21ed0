"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
it.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB = pathlib.Path(__file__).resolve().parent / "liboracle.so"
# The same restatement built for x86-64-v4 (AVX-512), used by the CPU
# baseline on hosts whose CPU has the v4 instructions (select_build).
LIB_V4 = pathlib.Path(__file__).resolve().parent / "liboracle_v4.so"
V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def host_has_v4() -> bool:
    try:
        flags = set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
        return all(f in flags for f in V4_FLAGS)
    except OSError:
        return False


def select_build(prefer_v4: bool = True) -> str:
    """Choose the library before the first call: the x86-64-v4 build when
    the host has AVX-512 and it exists, else the x86-64-v3 one.  Returns the
    -march level used."""
    global LIB
    if _lib is not None:
        return "x86-64-v4" if LIB == LIB_V4 else "x86-64-v3"
    if prefer_v4 and LIB_V4.exists() and host_has_v4():
        LIB = LIB_V4
        return "x86-64-v4"
    return "x86-64-v3"


def set_threads(threads: int) -> None:
    """Threads of the all-cores baseline (1: the serial restatement)."""
    lib().oracle_set_threads(int(threads))

KIND_CSR, KIND_SPLIT, KIND_BSR = 0, 1, 2


class Matrix(C.Structure):
    _fields_ = [("kind", C.c_int32), ("block_size", C.c_int32), ("owned_nodes", C.c_int64),
                ("ghost_nodes", C.c_int64), ("offsets", C.c_void_p), ("columns", C.c_void_p),
                ("values", C.c_void_p), ("diagonal", C.c_void_p)]


class PcParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("fill_level", C.c_int32), ("omega", C.c_double),
                ("minimum_diagonal", C.c_double), ("block_nodes", C.c_int64)]


class Options(C.Structure):
    _fields_ = [("relative_tolerance", C.c_double), ("absolute_tolerance", C.c_double),
                ("breakdown_tolerance", C.c_double), ("maximum_iterations", C.c_uint64),
                ("convergence_check_interval", C.c_uint64), ("restart", C.c_uint64),
                ("residual_replacement_interval", C.c_uint64), ("gmres_orthogonalization", C.c_int32),
                ("collect_timings", C.c_int32), ("idr_shadow_space", C.c_uint64)]


class MgOptions(C.Structure):
    _fields_ = [("pre_smoothing_steps", C.c_uint64), ("post_smoothing_steps", C.c_uint64),
                ("coarse_smoothing_steps", C.c_uint64), ("target_aggregate_size", C.c_uint64),
                ("jacobi_weight", C.c_double), ("smoother", C.c_int32), ("chebyshev_levels", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("breakdown_reason", C.c_int32), ("iterations", C.c_uint64),
                ("operator_applications", C.c_uint64), ("preconditioner_applications", C.c_uint64),
                ("global_reductions", C.c_uint64), ("reorthogonalizations", C.c_uint64),
                ("arnoldi_norm_verifications", C.c_uint64), ("initial_residual_norm", C.c_double),
                ("recursive_residual_norm", C.c_double), ("true_residual_norm", C.c_double),
                ("relative_residual_norm", C.c_double), ("solve_seconds", C.c_double),
                ("refinement_steps", C.c_uint64), ("inner_iterations", C.c_uint64)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise OSError(f"{LIB} missing: run `python oracle/build.py`")
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_get_threads.restype = C.c_int
        L.oracle_spmv.argtypes = [C.POINTER(Matrix), P, P]
        L.oracle_spmv_f32.argtypes = [C.POINTER(Matrix), P, P, P, P]
        L.oracle_dot.restype = C.c_double
        L.oracle_dot.argtypes = [C.c_int64, P, P]
        L.oracle_norm.restype = C.c_double
        L.oracle_norm.argtypes = [C.c_int64, P]
        L.oracle_pc_create.argtypes = [C.POINTER(Matrix), C.POINTER(PcParams), C.POINTER(P)]
        L.oracle_pc_update.argtypes = [P]
        L.oracle_pc_apply.argtypes = [P, P, P]
        L.oracle_pc_destroy.argtypes = [P]
        L.oracle_pc_destroy.restype = None
        L.oracle_ilu_factors.argtypes = [P, C.POINTER(C.c_int64), P, P, P, P]
        L.oracle_bicgstab.argtypes = [C.POINTER(Matrix), P, P, P, C.POINTER(Options), C.POINTER(Result)]
        L.oracle_fgmres.argtypes = [C.POINTER(Matrix), P, P, P, C.POINTER(Options), C.POINTER(Result)]
        L.oracle_fgmres_snapshot.argtypes = [C.POINTER(Matrix), P, P, P, C.POINTER(Options),
                                             C.POINTER(C.c_uint64), P, P, C.POINTER(Result)]
        L.oracle_recycled_fgmres_snapshot.argtypes = [C.POINTER(Matrix), P, P, P, P, C.POINTER(Options), C.c_int,
                                                      C.POINTER(C.c_uint64), P, P, C.POINTER(Result)]
        L.oracle_pipelined_bicgstab.argtypes = [C.POINTER(Matrix), P, P, P, C.POINTER(Options), C.POINTER(Result)]
        L.oracle_communication_hiding_bicgstab.argtypes = [C.POINTER(Matrix), P, P, P, C.POINTER(Options),
                                                           C.POINTER(Result)]
        L.oracle_idrs.argtypes = [C.POINTER(Matrix), P, P, P, C.POINTER(Options), C.POINTER(Result)]
        L.oracle_recycle_create.argtypes = [C.c_uint64, C.POINTER(P)]
        L.oracle_recycle_destroy.argtypes = [P]
        L.oracle_recycle_destroy.restype = None
        L.oracle_recycle_size.argtypes = [P]
        L.oracle_recycle_size.restype = C.c_int64
        L.oracle_recycle_get.argtypes = [P, C.c_int64, P, P]
        L.oracle_recycle_add.argtypes = [P, C.POINTER(Matrix), P, C.c_double, C.POINTER(C.c_int),
                                         C.POINTER(Result)]
        L.oracle_recycled_fgmres.argtypes = [C.POINTER(Matrix), P, P, P, P, C.POINTER(Options), C.c_int,
                                             C.POINTER(Result)]
        L.oracle_mg_create.argtypes = [C.POINTER(Matrix), C.POINTER(MgOptions), C.POINTER(P)]
        L.oracle_mg_solve.argtypes = [P, P, P, C.POINTER(Options), C.c_int, C.POINTER(Result)]
        L.oracle_mg_coarse.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), P, P, P, P]
        L.oracle_mg_destroy.argtypes = [P]
        L.oracle_mg_destroy.restype = None
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def _check(code):
    if code != 0:
        raise OracleError(code, lib().oracle_last_error().decode())


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class OMatrix:
    """Oracle view of a block-CSR matrix (keeps its numpy arrays alive)."""

    def __init__(self, kind, owned, ghost, block, offsets, columns, values, diagonal=None):
        self.kind, self.owned, self.ghost, self.block = kind, owned, ghost, block
        self.offsets = np.ascontiguousarray(offsets, np.int32)
        self.columns = np.ascontiguousarray(columns, np.int32)
        self.values = np.ascontiguousarray(values, np.float64)
        self.diagonal = None if diagonal is None else np.ascontiguousarray(diagonal, np.float64)
        self.s = Matrix(kind, block, owned, ghost, self.offsets.ctypes.data, self.columns.ctypes.data,
                        self.values.ctypes.data, 0 if self.diagonal is None else self.diagonal.ctypes.data)

    @classmethod
    def split(cls, owned, ghost, block, offsets, columns, diagonal, off_diagonal):
        return cls(KIND_SPLIT, owned, ghost, block, offsets, columns, off_diagonal, diagonal)

    @classmethod
    def csr(cls, owned, ghost, block, offsets, columns, values):
        return cls(KIND_CSR, owned, ghost, block, offsets, columns, values)

    @classmethod
    def bsr(cls, owned, ghost, block, offsets, columns, values):
        return cls(KIND_BSR, owned, ghost, block, offsets, columns, values)

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.owned * self.block)
        _check(lib().oracle_spmv(C.byref(self.s), _p(x), _p(y)))
        return y

    def spmv_f32(self, values32, diag32, x32):
        v = np.ascontiguousarray(values32, np.float32)
        d = None if diag32 is None else np.ascontiguousarray(diag32, np.float32)
        x = np.ascontiguousarray(x32, np.float32)
        y = np.zeros(self.owned * self.block, np.float32)
        _check(lib().oracle_spmv_f32(C.byref(self.s), _p(v), _p(d), _p(x), _p(y)))
        return y


def dot(a, b) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().oracle_dot(a.size, _p(a), _p(b))


def norm(a) -> float:
    a = np.ascontiguousarray(a, np.float64)
    return lib().oracle_norm(a.size, _p(a))


class OPreconditioner:
    IDENTITY, JACOBI, SSOR, ILU = 0, 1, 2, 3

    def __init__(self, matrix: OMatrix, kind, omega=1.0, block_nodes=0, fill_level=0, minimum_diagonal=0.0):
        self.matrix = matrix
        p = PcParams(kind, fill_level, omega, minimum_diagonal, block_nodes)
        self.h = C.c_void_p()
        _check(lib().oracle_pc_create(C.byref(matrix.s), C.byref(p), C.byref(self.h)))

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        out = np.zeros(self.matrix.owned * self.matrix.block)
        _check(lib().oracle_pc_apply(self.h, _p(r), _p(out)))
        return out

    def apply_inplace(self, r):
        a = np.array(r, np.float64)
        _check(lib().oracle_pc_apply(self.h, _p(a), _p(a)))
        return a

    def update(self):
        _check(lib().oracle_pc_update(self.h))

    def ilu_factors(self):
        n = C.c_int64()
        _check(lib().oracle_ilu_factors(self.h, C.byref(n), None, None, None, None))
        rows, bs = self.matrix.owned, self.matrix.block
        off = np.zeros(rows + 1, np.int64)
        col = np.zeros(n.value, np.int64)
        val = np.zeros(n.value * bs)
        inv = np.zeros(rows * bs)
        _check(lib().oracle_ilu_factors(self.h, C.byref(n), _p(off), _p(col), _p(val), _p(inv)))
        return off, col, val, inv

    def __del__(self):
        try:
            if self.h:
                lib().oracle_pc_destroy(self.h)
        except Exception:
            pass


def options(**kw) -> Options:
    o = Options(1e-8, 0.0, 64 * np.finfo(np.float64).eps, 1000, 10, 30, 50, 0, 0, 4)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def bicgstab(matrix: OMatrix, pc: OPreconditioner | None, b, x0=None, **kw):
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    r = Result()
    _check(lib().oracle_bicgstab(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), C.byref(o), C.byref(r)))
    return x, r.as_dict()


def pipelined_bicgstab(matrix: OMatrix, pc: OPreconditioner | None, b, x0=None, **kw):
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    r = Result()
    _check(lib().oracle_pipelined_bicgstab(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), C.byref(o),
                                           C.byref(r)))
    return x, r.as_dict()


def communication_hiding_bicgstab(matrix: OMatrix, pc: OPreconditioner | None, b, x0=None, **kw):
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    r = Result()
    _check(lib().oracle_communication_hiding_bicgstab(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x),
                                                      C.byref(o), C.byref(r)))
    return x, r.as_dict()


def idrs(matrix: OMatrix, pc: OPreconditioner | None, b, x0=None, **kw):
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    r = Result()
    _check(lib().oracle_idrs(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), C.byref(o), C.byref(r)))
    return x, r.as_dict()


def fgmres(matrix: OMatrix, pc: OPreconditioner | None, b, x0=None, **kw):
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    r = Result()
    _check(lib().oracle_fgmres(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), C.byref(o), C.byref(r)))
    return x, r.as_dict()


def fgmres_snapshot(matrix: OMatrix, pc: OPreconditioner | None, b, x0=None, **kw):
    """fgmres with an ArnoldiSnapshot (core.hpp:462-534): returns x, the
    result, and the snapshot as (H, Z) with H the (columns + 1) x columns
    unrotated Hessenberg matrix and Z the columns preconditioned basis
    vectors (local length each)."""
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    m = int(o.restart)
    h = np.zeros((m, m + 1))  # column-major storage: row j is column j
    z = np.zeros((m, n_local))
    cols = C.c_uint64()
    r = Result()
    _check(lib().oracle_fgmres_snapshot(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), C.byref(o),
                                        C.byref(cols), _p(h), _p(z), C.byref(r)))
    k = cols.value
    return x, r.as_dict(), (h[:k, :k + 1].T.copy(), z[:k].copy())


class OMultigrid:
    """AggregationMultigrid (multigrid.hpp:41-421) over a CSR-kind OMatrix.
    smoother: 0 Jacobi, 1 red-black Gauss-Seidel, 2 Chebyshev; cycle: 0 V,
    1 W, 2 full."""
    JACOBI, RED_BLACK, CHEBYSHEV = 0, 1, 2
    V, W, FULL = 0, 1, 2

    def __init__(self, matrix: OMatrix, pre=2, post=2, coarse_steps=24, aggregate=4, weight=0.8, smoother=0,
                 chebyshev_levels=8):
        self.matrix = matrix
        o = MgOptions(pre, post, coarse_steps, aggregate, weight, smoother, chebyshev_levels)
        self.h = C.c_void_p()
        _check(lib().oracle_mg_create(C.byref(matrix.s), C.byref(o), C.byref(self.h)))

    def coarse(self):
        nodes, entries = C.c_int64(), C.c_int64()
        _check(lib().oracle_mg_coarse(self.h, C.byref(nodes), C.byref(entries), None, None, None, None))
        off = np.zeros(nodes.value + 1, np.int32)
        col = np.zeros(entries.value, np.int32)
        val = np.zeros(entries.value * self.matrix.block)
        f2c = np.zeros(self.matrix.owned, np.int64)
        _check(lib().oracle_mg_coarse(self.h, C.byref(nodes), C.byref(entries), _p(off), _p(col), _p(val), _p(f2c)))
        return nodes.value, off, col, val, f2c

    def solve(self, b, x0=None, cycle=0, **kw):
        b = np.ascontiguousarray(b, np.float64)
        n_local = (self.matrix.owned + self.matrix.ghost) * self.matrix.block
        x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
        o = options(**kw)
        r = Result()
        _check(lib().oracle_mg_solve(self.h, _p(b), _p(x), C.byref(o), cycle, C.byref(r)))
        return x, r.as_dict()

    def __del__(self):
        try:
            if self.h:
                lib().oracle_mg_destroy(self.h)
        except Exception:
            pass


class ORecycleSpace:
    """RecycleSpace (recycling.hpp:24-187): U, C = A U, C orthonormal."""

    def __init__(self, capacity: int):
        self.h = C.c_void_p()
        _check(lib().oracle_recycle_create(capacity, C.byref(self.h)))

    def size(self) -> int:
        return lib().oracle_recycle_size(self.h)

    def get(self, i: int, n: int):
        u, c = np.zeros(n), np.zeros(n)
        _check(lib().oracle_recycle_get(self.h, i, _p(u), _p(c)))
        return u, c

    def add_candidate(self, matrix: OMatrix, candidate, tol=128 * np.finfo(np.float64).eps):
        cand = np.ascontiguousarray(candidate, np.float64)
        added = C.c_int()
        r = Result()
        _check(lib().oracle_recycle_add(self.h, C.byref(matrix.s), _p(cand), tol, C.byref(added), C.byref(r)))
        return bool(added.value), r.as_dict()

    def __del__(self):
        try:
            if self.h:
                lib().oracle_recycle_destroy(self.h)
        except Exception:
            pass


def recycled_fgmres(matrix: OMatrix, pc: OPreconditioner | None, b, space: ORecycleSpace, x0=None, update=True,
                    **kw):
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    r = Result()
    _check(lib().oracle_recycled_fgmres(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), space.h, C.byref(o),
                                        1 if update else 0, C.byref(r)))
    return x, r.as_dict()


def extract_harmonic_ritz(matrix: OMatrix, Z, capacity: int, eig=None):
    """extract_harmonic_ritz (gcrodr.hpp:86-181) over the snapshot basis Z
    (columns x local): images A Z, G = (AZ)^T AZ, M = (AZ)^T Z, the pencil
    G y = theta M y by LAPACK ggev (scipy.linalg.eig with two matrices, the
    routine the reference calls; `eig` may substitute another solver with the
    same (alpha, beta, vr) output), groups in ascending |theta|, each a real
    vector or a conjugate pair's (real, imaginary) parts of Z y, committed to
    a new ORecycleSpace(capacity) only if all its directions pass
    add_candidate.  Returns the new space, its candidates and the telemetry
    the reference adds (operator applications and global reductions of the
    images, the projected matrices and every trial add_candidate)."""
    import scipy.linalg as sla

    n_owned = matrix.owned * matrix.block
    Z = np.asarray(Z, np.float64)
    k = Z.shape[0]
    space = ORecycleSpace(capacity)
    tel = {"operator_applications": 0, "global_reductions": 0}
    if k == 0:
        return space, [], tel
    AZ = np.array([matrix.spmv(z)[:n_owned] for z in Z])
    Zo = Z[:, :n_owned]
    G = AZ @ AZ.T
    M = AZ @ Zo.T
    tel["operator_applications"] += k
    tel["global_reductions"] += 1
    if eig is None:
        w, vr = sla.eig(G, M, homogeneous_eigvals=True)
        alpha, beta = w[0], w[1]
    else:
        alpha, beta, vr = eig(G, M)
    # gcrodr.hpp:137-155 over dggev's (alphar + i alphai, beta)
    groups = []
    j = 0
    while j < k:
        width = 2 if alpha[j].imag > 0 else 1
        bj = abs(beta[j])
        mag = abs(alpha[j]) / bj if bj != 0 else np.inf
        if alpha[j].imag >= 0 and j + width <= k and bj != 0 and np.isfinite(mag):
            groups.append((mag, j, width))
        j += width
    groups.sort(key=lambda g: g[0])  # stable
    accepted = []
    for _, first, width in groups:
        if len(accepted) + width > capacity:
            continue
        y = vr[:, first]
        parts = [np.real(y)] if width == 1 else [np.real(y), np.imag(y)]
        cands = [np.asarray(part @ Z, np.float64) for part in parts]
        trial = ORecycleSpace(capacity)
        for c in accepted:
            trial.add_candidate(matrix, c)
        ok = True
        for c in cands:  # every part is tried, as gcrodr.hpp:169-170
            added, t = trial.add_candidate(matrix, c)
            ok = added and ok
            for key in tel:
                tel[key] += t[key]
        if ok:
            accepted.extend(cands)
    for c in accepted:
        space.add_candidate(matrix, c)
    return space, accepted, tel


def gcrodr(matrix: OMatrix, pc: OPreconditioner | None, b, space: ORecycleSpace, capacity: int, x0=None, **kw):
    """gcrodr (gcrodr.hpp:185-212): recycled_fgmres without the update,
    recording the last cycle, then extract_harmonic_ritz.  Returns x, the
    result, the new space and the snapshot (H, Z)."""
    b = np.ascontiguousarray(b, np.float64)
    n_local = (matrix.owned + matrix.ghost) * matrix.block
    x = np.zeros(n_local) if x0 is None else np.array(x0, np.float64)
    o = options(**kw)
    m = int(o.restart)
    h = np.zeros((m, m + 1))
    z = np.zeros((m, n_local))
    cols = C.c_uint64()
    r = Result()
    _check(lib().oracle_recycled_fgmres_snapshot(C.byref(matrix.s), pc.h if pc else None, _p(b), _p(x), space.h,
                                                 C.byref(o), 0, C.byref(cols), _p(h), _p(z), C.byref(r)))
    kc = cols.value
    new_space = space
    res = r.as_dict()
    if kc:
        new_space, _, tel = extract_harmonic_ritz(matrix, z[:kc], capacity)
        for key in tel:
            res[key] += tel[key]
    return x, res, new_space, (h[:kc, :kc + 1].T.copy(), z[:kc].copy())

