"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the CPU
baseline timed by bench.py.  The product package (magla_amd) never imports
this module.  Each wrapped C function cites the MAGLA reference lines it
restates (see oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("colidx", C.POINTER(C.c_int)), ("values", C.POINTER(C.c_double)),
                ("diag", C.POINTER(C.c_int))]


class _Opts(C.Structure):
    _fields_ = [("ordering", C.c_int), ("amalgamation_limit", C.c_int), ("panel_width", C.c_int),
                ("allow_regularization", C.c_int), ("pivot_tolerance", C.c_double),
                ("regularization", C.c_double), ("time_budget_s", C.c_double)]


class _Sym(C.Structure):
    _fields_ = [("n", C.c_int), ("matrix_nnz", C.c_int),
                ("amd", C.POINTER(C.c_int)), ("perm", C.POINTER(C.c_int)), ("iperm", C.POINTER(C.c_int)),
                ("etree", C.POINTER(C.c_int)), ("colcount", C.POINTER(C.c_int)),
                ("pptr", C.POINTER(C.c_int)), ("prows", C.POINTER(C.c_int)),
                ("lower_start", C.POINTER(C.c_int)), ("lvptr", C.POINTER(C.c_int)),
                ("lvidx", C.POINTER(C.c_int)), ("dvidx", C.POINTER(C.c_int)),
                ("nsuper", C.c_int), ("first", C.POINTER(C.c_int)), ("super_of", C.POINTER(C.c_int)),
                ("sparent", C.POINTER(C.c_int)), ("child_ptr", C.POINTER(C.c_int)),
                ("children", C.POINTER(C.c_int)), ("pat_ptr", C.POINTER(C.c_int)),
                ("pat_rows", C.POINTER(C.c_int)), ("block_off", C.POINTER(C.c_int64)),
                ("factor_entries", C.c_int64), ("strict_nonzeros", C.c_int64),
                ("factor_ops", C.c_int64), ("max_front", C.c_int), ("max_width", C.c_int),
                ("peak_contribution", C.c_int64)]


class _Num(C.Structure):
    _fields_ = [("sym", C.c_void_p), ("factor", C.POINTER(C.c_double)), ("diag", C.POINTER(C.c_double)),
                ("regularized", C.c_int), ("negative", C.c_int), ("smallest", C.c_double),
                ("largest", C.c_double), ("supernodes_done", C.c_int), ("ops_done", C.c_int64)]


class _PcgRes(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("residual_norm", C.c_double),
                ("relative_residual", C.c_double)]


class _GCsr(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("colidx", C.POINTER(C.c_int)), ("values", C.POINTER(C.c_double)), ("diag", C.POINTER(C.c_int))]


class _Ilu0(C.Structure):
    _fields_ = [("a", C.POINTER(_GCsr)), ("pivot_tolerance", C.c_double), ("factors", C.POINTER(C.c_double)),
                ("diag", C.POINTER(C.c_int))]


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: oracle status {code}")
        self.code = code


_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_csr_from_coo.argtypes = [C.c_int, C.c_int64, _i32p, _i32p, _f64p, C.POINTER(C.POINTER(_Csr))]
        L.or_csr_create.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.POINTER(C.POINTER(_Csr))]
        L.or_csr_free.argtypes = [C.POINTER(_Csr)]
        L.or_spmv.argtypes = [C.POINTER(_Csr), _f64p, _f64p]
        L.or_linecross.argtypes = [C.POINTER(_Csr), _f64p, C.c_double, C.c_int]
        L.or_amd.argtypes = [C.c_int, _i32p, _i32p, _i32p]
        L.or_options_default.argtypes = [C.POINTER(_Opts)]
        L.or_analyze.argtypes = [C.POINTER(_Csr), C.POINTER(_Opts), C.POINTER(C.POINTER(_Sym))]
        L.or_symbolic_free.argtypes = [C.POINTER(_Sym)]
        L.or_factorize.argtypes = [C.POINTER(_Sym), C.POINTER(_Csr), C.POINTER(_Opts), C.POINTER(C.POINTER(_Num))]
        L.or_numeric_free.argtypes = [C.POINTER(_Num)]
        L.or_solve.argtypes = [C.POINTER(_Num), _f64p, _f64p]
        L.or_factorize_mt.argtypes = [C.POINTER(_Sym), C.POINTER(_Csr), C.POINTER(_Opts), C.c_int,
                                      C.POINTER(C.POINTER(_Num))]
        L.or_solve_mt.argtypes = [C.POINTER(_Num), C.c_int, _f64p, _f64p, C.c_int]
        L.or_pcg.argtypes = [C.POINTER(_Csr), _f64p, C.c_void_p, C.c_double, C.c_int, C.c_int, _f64p,
                             C.POINTER(_PcgRes)]
        L.or_mpnn_apply.argtypes = [C.POINTER(_Csr), _f64p, _f64p, _f64p]
        L.or_mpnn_loss_grad.argtypes = [C.POINTER(_Csr), _f64p, C.c_int, _f64p, _f64p, C.c_double,
                                        C.POINTER(C.c_double), _f64p]
        L.or_gcsr_from_coo.argtypes = [C.c_int, C.c_int, C.c_int64, _i32p, _i32p, _f64p,
                                       C.POINTER(C.POINTER(_GCsr))]
        L.or_gcsr_create.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.POINTER(C.POINTER(_GCsr))]
        L.or_gcsr_free.argtypes = [C.POINTER(_GCsr)]
        L.or_gspmv.argtypes = [C.POINTER(_GCsr), _f64p, _f64p]
        L.or_gcsr_linecross.argtypes = [C.POINTER(_GCsr), _f64p, C.c_double, C.c_int]
        L.or_ilu0_create.argtypes = [C.POINTER(_GCsr), C.c_double, C.POINTER(C.POINTER(_Ilu0)),
                                     C.POINTER(C.c_int)]
        L.or_ilu0_apply.argtypes = [C.POINTER(_Ilu0), _f64p, _f64p]
        L.or_ilu0_free.argtypes = [C.POINTER(_Ilu0)]
        L.or_bicgstab.argtypes = [C.POINTER(_GCsr), C.POINTER(_Ilu0), C.c_int, _f64p, C.c_void_p, C.c_double,
                                  C.c_int, C.c_double, _f64p, C.POINTER(_PcgRes), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _arr(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


def _check(rc, what):
    if rc != 0:
        raise OracleError(rc, what)


class Csr:
    """Owned oracle CSR (upper triangle + diagonal)."""

    def __init__(self, ptr):
        self._p = ptr

    @classmethod
    def from_coo(cls, n, rows, cols, vals):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        p = C.POINTER(_Csr)()
        _check(lib().or_csr_from_coo(n, len(rows), rows, cols, vals, C.byref(p)), "csr_from_coo")
        return cls(p)

    @classmethod
    def from_csr(cls, n, rowptr, colidx, values):
        p = C.POINTER(_Csr)()
        _check(lib().or_csr_create(n, np.ascontiguousarray(rowptr, np.int32),
                                   np.ascontiguousarray(colidx, np.int32),
                                   np.ascontiguousarray(values, np.float64), C.byref(p)), "csr_create")
        return cls(p)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_csr_free(self._p)
            self._p = None

    @property
    def n(self):
        return self._p.contents.n

    @property
    def nnz(self):
        return self._p.contents.nnz

    def arrays(self):
        c = self._p.contents
        return (_arr(c.rowptr, c.n + 1, np.int32), _arr(c.colidx, c.nnz, np.int32),
                _arr(c.values, c.nnz, np.float64))

    def diag_index(self):
        c = self._p.contents
        return _arr(c.diag, c.n, np.int32)

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n, dtype=np.float64)
        lib().or_spmv(self._p, x, y)
        return y

    def linecross(self, rhs, prescribed, index):
        """In place on both the matrix values and rhs (a numpy float64 array)."""
        _check(lib().or_linecross(self._p, rhs, float(prescribed), int(index)), "linecross")

    def dense(self):
        rp, ci, v = self.arrays()
        a = np.zeros((self.n, self.n))
        for r in range(self.n):
            for k in range(rp[r], rp[r + 1]):
                a[r, ci[k]] = v[k]
                a[ci[k], r] = v[k]
        return a


def options(ordering=1, amalgamation_limit=64, panel_width=32, allow_regularization=True,
            pivot_tolerance=1e-13, regularization=1e-10, time_budget_s=0.0):
    o = _Opts()
    lib().or_options_default(C.byref(o))
    o.ordering = ordering
    o.amalgamation_limit = amalgamation_limit
    o.panel_width = panel_width
    o.allow_regularization = 1 if allow_regularization else 0
    o.pivot_tolerance = pivot_tolerance
    o.regularization = regularization
    o.time_budget_s = time_budget_s
    return o


def amd(n, rowptr, colidx):
    perm = np.zeros(n, dtype=np.int32)
    _check(lib().or_amd(n, np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(colidx, np.int32), perm),
           "amd")
    return perm


class Symbolic:
    def __init__(self, csr: Csr, opts=None):
        self.opts = opts or options()
        p = C.POINTER(_Sym)()
        _check(lib().or_analyze(csr._p, C.byref(self.opts), C.byref(p)), "analyze")
        self._p = p
        s = p.contents
        n, ns = s.n, s.nsuper
        self.n, self.nsuper = n, ns
        self.amd = _arr(s.amd, n, np.int32)
        self.perm = _arr(s.perm, n, np.int32)
        self.iperm = _arr(s.iperm, n, np.int32)
        self.etree = _arr(s.etree, n, np.int32)
        self.colcount = _arr(s.colcount, n, np.int32)
        self.first = _arr(s.first, ns + 1, np.int32)
        self.sparent = _arr(s.sparent, ns, np.int32)
        self.child_ptr = _arr(s.child_ptr, ns + 1, np.int32)
        self.children = _arr(s.children, int(self.child_ptr[-1]), np.int32)
        self.pat_ptr = _arr(s.pat_ptr, ns + 1, np.int32)
        self.pat_rows = _arr(s.pat_rows, int(self.pat_ptr[-1]), np.int32)
        self.block_off = _arr(s.block_off, ns + 1, np.int64)
        self.factor_entries = s.factor_entries
        self.strict_nonzeros = s.strict_nonzeros
        self.factor_ops = s.factor_ops
        self.max_front = s.max_front
        self.max_width = s.max_width
        self.peak_contribution = s.peak_contribution

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_symbolic_free(self._p)
            self._p = None


class Numeric:
    def __init__(self, sym: Symbolic, csr: Csr, threads=None):
        """threads=None: the serial restatement; threads=T: the reference's threaded schedule
        over T workers (the same factor bit for bit)."""
        self.sym = sym
        p = C.POINTER(_Num)()
        if threads is None:
            _check(lib().or_factorize(sym._p, csr._p, C.byref(sym.opts), C.byref(p)), "factorize")
        else:
            _check(lib().or_factorize_mt(sym._p, csr._p, C.byref(sym.opts), int(threads), C.byref(p)), "factorize")
        self._p = p
        f = p.contents
        self.diag = _arr(f.diag, sym.n, np.float64)
        self.factor = _arr(f.factor, int(sym.factor_entries), np.float64)
        self.regularized = f.regularized
        self.negative = f.negative
        self.smallest = f.smallest
        self.largest = f.largest
        self.supernodes_done = f.supernodes_done
        self.ops_done = f.ops_done

    @property
    def is_positive_definite(self):
        return self.regularized == 0 and self.negative == 0

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        _check(lib().or_solve(self._p, b, x), "solve")
        return x

    def solve_many(self, B, threads=1):
        """columns of B (n x k) solved independently over `threads` threads"""
        Bt = np.ascontiguousarray(np.asarray(B, dtype=np.float64).T)
        Xt = np.zeros_like(Bt)
        _check(lib().or_solve_mt(self._p, Bt.shape[0], Bt, Xt, int(threads)), "solve")
        return Xt.T

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_numeric_free(self._p)
            self._p = None


def pcg(csr: Csr, b, x0=None, tol=1e-8, max_iterations=0, use_jacobi=True):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(csr.n, dtype=np.float64)
    res = _PcgRes()
    x0p = None
    if x0 is not None:
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x0p = x0.ctypes.data_as(C.c_void_p)
    _check(lib().or_pcg(csr._p, b, x0p, tol, max_iterations, 1 if use_jacobi else 0, x, C.byref(res)), "pcg")
    return x, dict(iterations=res.iterations, converged=bool(res.converged),
                   residual_norm=res.residual_norm, relative_residual=res.relative_residual)


# ---- learned preconditioner network (checker; no reference counterpart) ----
MP_W = 32
MP_P = 3 * MP_W + 3 * 2 * MP_W * MP_W + 7 * MP_W


def mpnn_apply(csr: Csr, params, r):
    """z = G'G r with the network's factor G, in binary64 (see oracle.h for the network)."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    assert params.size == MP_P
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros(csr.n, dtype=np.float64)
    _check(lib().or_mpnn_apply(csr._p, params, r, z), "mpnn_apply")
    return z


def mpnn_loss_grad(csr: Csr, params, R, E, kappa=0.0):
    """Mean A-norm loss over the rows of R (residuals) and E (exact errors), plus kappa times
    the log Kaporin condition number of G A G', and its gradient."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    R = np.ascontiguousarray(np.atleast_2d(R), dtype=np.float64)
    E = np.ascontiguousarray(np.atleast_2d(E), dtype=np.float64)
    assert R.shape == E.shape and R.shape[1] == csr.n
    loss = C.c_double()
    grad = np.zeros(MP_P, dtype=np.float64)
    _check(lib().or_mpnn_loss_grad(csr._p, params, R.shape[0], R, E, float(kappa), C.byref(loss), grad),
           "mpnn_loss_grad")
    return loss.value, grad


# ---- general CSR, ILU(0), BiCGSTAB (CSRMatrix.cs, BiCGSTAB/*.cs) ----
class GCsr:
    """Owned oracle CSRMatrix (general, rows x cols)."""

    def __init__(self, ptr):
        self._p = ptr

    @classmethod
    def from_coo(cls, rows, cols, ri, cj, vals):
        ri = np.ascontiguousarray(ri, dtype=np.int32)
        cj = np.ascontiguousarray(cj, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        p = C.POINTER(_GCsr)()
        _check(lib().or_gcsr_from_coo(rows, cols, len(ri), ri, cj, vals, C.byref(p)), "gcsr_from_coo")
        return cls(p)

    @classmethod
    def from_dense(cls, A):
        A = np.asarray(A, dtype=np.float64)
        ri, cj = np.nonzero(A)
        return cls.from_coo(A.shape[0], A.shape[1], ri, cj, A[ri, cj])

    @classmethod
    def from_csr(cls, rows, cols, rowptr, colidx, values):
        p = C.POINTER(_GCsr)()
        _check(lib().or_gcsr_create(rows, cols, np.ascontiguousarray(rowptr, np.int32),
                                    np.ascontiguousarray(colidx, np.int32),
                                    np.ascontiguousarray(values, np.float64), C.byref(p)), "gcsr_create")
        return cls(p)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_gcsr_free(self._p)
            self._p = None

    @property
    def shape(self):
        return self._p.contents.rows, self._p.contents.cols

    def arrays(self):
        a = self._p.contents
        return (_arr(a.rowptr, a.rows + 1, np.int32), _arr(a.colidx, a.nnz, np.int32),
                _arr(a.values, a.nnz, np.float64))

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self._p.contents.rows)
        lib().or_gspmv(self._p, x, y)
        return y

    def linecross(self, rhs, prescribed, index):
        _check(lib().or_gcsr_linecross(self._p, rhs, prescribed, index), "gcsr_linecross")

    def dense(self):
        rp, ci, v = self.arrays()
        r, c = self.shape
        A = np.zeros((r, c))
        for i in range(r):
            A[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
        return A


class Ilu0:
    """Oracle Ilu0Preconditioner; raises OracleError (with .row) on a missing diagonal / bad pivot."""

    def __init__(self, a: GCsr, pivot_tolerance=1e-30):
        self.a = a
        p = C.POINTER(_Ilu0)()
        bad = C.c_int(-1)
        rc = lib().or_ilu0_create(a._p, pivot_tolerance, C.byref(p), C.byref(bad))
        if rc != 0:
            e = OracleError(rc, "ilu0")
            e.row = bad.value
            raise e
        self._p = p

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_ilu0_free(self._p)
            self._p = None

    @property
    def factors(self):
        return _arr(self._p.contents.factors, self.a._p.contents.nnz, np.float64)

    def apply(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().or_ilu0_apply(self._p, b, x)
        return x


def bicgstab(a: GCsr, b, x0=None, tol=1e-8, max_iterations=0, use_preconditioner=True,
             breakdown_tolerance=1e-30, preconditioner: Ilu0 | None = None):
    """BiCgStabSolver.SolveWithInitialGuess: (x, result dict); raises OracleError with .fail_code and
    .fail_it on a breakdown (1 rho, 2 omega, 3 (r0,Ap), 4 (t,t), 5 omega; 6/7 ILU(0) diagonal/pivot)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    res = _PcgRes()
    code, it = C.c_int(0), C.c_int(0)
    x0p = None
    if x0 is not None:
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x0p = x0.ctypes.data_as(C.c_void_p)
    pre = preconditioner._p if preconditioner is not None else None
    rc = lib().or_bicgstab(a._p, pre, 1 if use_preconditioner else 0, b, x0p, tol, max_iterations,
                           breakdown_tolerance, x, C.byref(res), C.byref(code), C.byref(it))
    if rc != 0:
        e = OracleError(rc, "bicgstab")
        e.fail_code, e.fail_it = code.value, it.value
        raise e
    return x, dict(iterations=res.iterations, converged=bool(res.converged), residual_norm=res.residual_norm,
                   relative_residual=res.relative_residual)
