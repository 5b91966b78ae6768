"""MI355X UCHOL / CGGP solver library: Python binding of libmagla.so.

The classes mirror MAGLA's C# surface (CAESolvers) so that tests and the
benchmark read like the reference's own code:

  SymmetricCSRMatrix          SymmetricCSRMatrix.cs (upper triangle + diagonal)
  UtduSolverOptions           UCHOL/UtduSymbolicFactorization.cs:31-98
  SymmetricUtduSolver         UCHOL/SymmetricUtduSolver.cs:45-109
  UtduSymbolicFactorization   UCHOL/UtduSymbolicFactorization.cs:109-320
  UtduNumericFactorization    UCHOL/UtduNumericFactorization.cs:24-235
  ConjugateGradientGaussPreSolver  CGGP/ConjugateGradientGaussPreSolver.cs:40-196
  CSRMatrix                   CSRMatrix.cs (general square CSR)
  Ilu0Preconditioner          BiCGSTAB/Ilu0Preconditioner.cs
  BiCgStabSolver              BiCGSTAB/BiCgStabSolver.cs

All numeric work runs in libmagla.so on the GPU; there is no CPU fallback:
loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmagla.so")

MG_OK, MG_E_NULL, MG_E_ARG, MG_E_RANGE, MG_E_NOT_PD, MG_E_TOO_DENSE, MG_E_DEVICE, MG_E_INTERNAL, MG_E_NOMEM = range(9)
NATURAL, AMD = 0, 1
FP64, FP32, BF16 = 0, 1, 2

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _Options(C.Structure):
    _fields_ = [("ordering", C.c_int32), ("max_degree_of_parallelism", C.c_int32),
                ("supernode_amalgamation_limit", C.c_int32), ("panel_width", C.c_int32),
                ("allow_diagonal_regularization", C.c_int32), ("precision", C.c_int32),
                ("pivot_tolerance", C.c_double), ("diagonal_regularization", C.c_double),
                ("parallel_update_threshold", C.c_int64), ("device_graph", C.c_int32), ("reserved", C.c_int32)]


class _SymInfo(C.Structure):
    _fields_ = [("size", C.c_int32), ("supernode_count", C.c_int32), ("max_front_size", C.c_int32),
                ("max_supernode_width", C.c_int32), ("factor_entry_count", C.c_int64),
                ("strict_factor_nonzero_count", C.c_int64), ("factor_operation_count", C.c_int64),
                ("estimated_factor_bytes", C.c_int64), ("max_front_bytes", C.c_int64),
                ("peak_contribution_bytes", C.c_int64), ("structure_bytes", C.c_int64),
                ("level_count", C.c_int32), ("matrix_nonzero_count", C.c_int32)]


class _NumInfo(C.Structure):
    _fields_ = [("regularized_pivot_count", C.c_int32), ("negative_pivot_count", C.c_int32),
                ("smallest_pivot_magnitude", C.c_double), ("largest_pivot_magnitude", C.c_double),
                ("is_positive_definite", C.c_int32), ("precision", C.c_int32)]


class _SolveReport(C.Structure):
    _fields_ = [("refinement_steps", C.c_int32), ("converged", C.c_int32), ("relative_residual", C.c_double)]


class _PcgResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("residual_norm", C.c_double),
                ("relative_residual", C.c_double)]


class MaglaError(RuntimeError):
    """Status code + message; the exception class mirrors the C# one."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class ArgumentNullError(MaglaError, ValueError):
    pass


class ArgumentError(MaglaError, ValueError):
    pass


class ArgumentOutOfRangeError(MaglaError, ValueError):
    pass


class InvalidOperationError(MaglaError):
    pass


_EXC = {MG_E_NULL: ArgumentNullError, MG_E_ARG: ArgumentError, MG_E_RANGE: ArgumentOutOfRangeError}

_lib = None


def _init_torch_runtime_first():
    """PyTorch-ROCm ships its own HIP runtime (a libamdhip64.so without the SONAME of ROCm's
    libamdhip64.so.7, so the loader keeps both).  Two runtimes can share one process only if
    torch's initializes the device first; the other order leaves torch with no devices.  When
    torch is importable its runtime is therefore started before libmagla.so is loaded."""
    if os.environ.get("MAGLA_SKIP_TORCH_INIT"):
        return
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def lib():
    """Load libmagla.so (built by magla_amd.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libmagla.so not built: run `python -m magla_amd.build` ({LIB_PATH})")
    _init_torch_runtime_first()
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.mg_last_error.restype = C.c_char_p
    L.mg_device_count.argtypes = [C.POINTER(C.c_int)]
    L.mg_options_default.argtypes = [C.POINTER(_Options)]
    L.mg_options_default.restype = None
    L.mg_csr_from_coo.argtypes = [C.c_int32, C.c_int64, _i32p, _i32p, _f64p, C.POINTER(C.c_int32), vp, vp, vp]
    L.mg_csr_line_cross.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, C.c_int32, _i32p, _f64p]
    L.mg_analyze.argtypes = [C.c_int32, _i32p, vp, C.POINTER(_Options), C.POINTER(vp)]
    L.mg_symbolic_info_get.argtypes = [vp, C.POINTER(_SymInfo), vp, vp]
    L.mg_symbolic_structure.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.mg_symbolic_estimate_peak_bytes.argtypes = [vp, C.c_int32]
    L.mg_symbolic_estimate_peak_bytes.restype = C.c_int64
    L.mg_symbolic_release.argtypes = [vp]
    L.mg_symbolic_release.restype = None
    L.mg_factorize.argtypes = [vp, C.c_int32, C.c_int32, vp, C.POINTER(_Options), C.POINTER(vp)]
    L.mg_factorize_dev.argtypes = [vp, C.c_int32, vp, C.POINTER(_Options), vp, C.POINTER(vp)]
    L.mg_numeric_info_get.argtypes = [vp, C.POINTER(_NumInfo), vp]
    L.mg_numeric_release.argtypes = [vp]
    L.mg_numeric_release.restype = None
    L.mg_solve.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, C.c_double, C.c_int32,
                           C.POINTER(_SolveReport)]
    L.mg_solve_dev.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, C.c_double, C.c_int32, vp,
                               C.POINTER(_SolveReport)]
    L.mg_solve_dev_noalloc.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, C.c_double, C.c_int32, vp,
                                       C.POINTER(_SolveReport)]
    L.mg_matrix_create.argtypes = [C.c_int32, _i32p, vp, vp, C.POINTER(vp)]
    L.mg_matrix_release.argtypes = [vp]
    L.mg_matrix_release.restype = None
    L.mg_spmv.argtypes = [vp, _f64p, _f64p]
    L.mg_spmv_dev.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, vp]
    L.mg_pcg_solve.argtypes = [vp, _f64p, vp, C.c_double, C.c_int32, C.c_int32, _f64p, C.POINTER(_PcgResult)]
    L.mg_pcg_solve_dev.argtypes = [vp, vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp, C.POINTER(_PcgResult)]
    L.mg_batch_analyze.argtypes = [C.c_int32, vp, vp, vp, C.POINTER(_Options), C.POINTER(vp)]
    L.mg_batch_layout.argtypes = [vp, vp, vp, C.POINTER(_SymInfo)]
    L.mg_batch_factorize_dev.argtypes = [vp, vp, C.POINTER(_Options), vp]
    L.mg_batch_solve_dev.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, C.c_double, C.c_int32, vp,
                                     C.POINTER(_SolveReport)]
    L.mg_batch_factorize.argtypes = [vp, vp, C.POINTER(_Options)]
    L.mg_batch_solve.argtypes = [vp, vp, vp, vp, C.c_double, C.c_int32, C.POINTER(_SolveReport)]
    L.mg_batch_numeric_info.argtypes = [vp, C.POINTER(_NumInfo)]
    L.mg_autotune_dev.argtypes = [vp, vp, C.POINTER(_Options), vp, vp, C.POINTER(C.c_double)]
    L.mg_batch_autotune_dev.argtypes = [vp, vp, C.POINTER(_Options), vp, vp, C.POINTER(C.c_double)]
    L.mg_test_set_schur_tiles.argtypes = [vp, vp]
    L.mg_gcsr_from_coo.argtypes = [C.c_int32, C.c_int32, C.c_int64, _i32p, _i32p, _f64p, C.POINTER(C.c_int32),
                                   vp, vp, vp]
    L.mg_gcsr_line_cross.argtypes = [C.c_int32, C.c_int32, _i32p, vp, vp, _f64p, C.c_int32, _i32p, _f64p]
    L.mg_gmatrix_create.argtypes = [C.c_int32, C.c_int32, _i32p, vp, vp, C.POINTER(vp)]
    L.mg_gmatrix_release.argtypes = [vp]
    L.mg_gmatrix_release.restype = None
    L.mg_gspmv.argtypes = [vp, _f64p, _f64p]
    L.mg_gspmv_dev.argtypes = [vp, vp, vp, vp]
    L.mg_ilu0_create.argtypes = [vp, C.c_double, C.POINTER(vp)]
    L.mg_ilu0_release.argtypes = [vp]
    L.mg_ilu0_release.restype = None
    L.mg_ilu0_apply.argtypes = [vp, _f64p, _f64p]
    L.mg_ilu0_apply_dev.argtypes = [vp, vp, vp, vp]
    L.mg_bicgstab_solve.argtypes = [vp, vp, C.c_int32, _f64p, vp, C.c_double, C.c_int32, C.c_double, _f64p,
                                    C.POINTER(_PcgResult)]
    L.mg_bicgstab_solve_dev.argtypes = [vp, vp, C.c_int32, vp, vp, C.c_double, C.c_int32, C.c_double, vp, vp,
                                        C.POINTER(_PcgResult)]
    L.mg_test_ilu0_factors.argtypes = [vp, _f64p]
    L.mg_batch_release.argtypes = [vp]
    L.mg_batch_release.restype = None
    L.mg_mpnn_param_count.restype = C.c_int32
    L.mg_mpnn_create.argtypes = [C.c_uint64, C.POINTER(vp)]
    L.mg_mpnn_release.argtypes = [vp]
    L.mg_mpnn_release.restype = None
    L.mg_mpnn_get_params.argtypes = [vp, vp]
    L.mg_mpnn_set_params.argtypes = [vp, vp]
    L.mg_mpnn_apply_dev.argtypes = [vp, vp, vp, vp, C.c_int32, vp]
    L.mg_mpnn_loss_grad_dev.argtypes = [vp, vp, C.c_int32, vp, vp, C.c_int64, C.c_double, vp,
                                        C.POINTER(C.c_double), vp]
    L.mg_mpnn_train_dev.argtypes = [vp, vp, C.c_int32, vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                    C.c_double, C.c_uint64, vp, vp]
    L.mg_pcg_solve_learned_dev.argtypes = [vp, vp, vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp,
                                           C.POINTER(_PcgResult), C.POINTER(C.c_int32)]
    L.mg_pcg_solve_learned.argtypes = [vp, vp, _f64p, vp, C.c_double, C.c_int32, C.c_int32, _f64p,
                                       C.POINTER(_PcgResult), C.POINTER(C.c_int32)]
    L.mg_test_mfma_bf16.argtypes = [vp, vp, vp]
    L.mg_test_numeric_front.argtypes = [vp, C.c_int32, vp]
    L.mg_test_mfma_layout.argtypes = [C.c_int32, vp]
    L.mg_test_plan_stats.argtypes = [vp, vp]
    L.mg_test_plan_layout.argtypes = [vp, vp]
    L.mg_test_front_addr.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int32, vp, vp, vp]
    L.mg_profile_enable.argtypes = [C.c_int32]
    L.mg_profile_enable.restype = None
    L.mg_profile_read.argtypes = [C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


def test_mfma_layout(precision):
    """Probe MFMA product placed by the kernel's layout rule (include/magla_testing.h): (16, 16) array
    that must equal i + 16 j."""
    out = np.zeros(256)
    _check(lib().mg_test_mfma_layout(precision, _ptr(out)))
    return out.reshape(16, 16)


def plan_stats(symbolic):
    """Device schedule summary (host only): fronts, levels, arena elements, max front, max width,
    launches per factorization, Schur tiles of the lookahead part, factor-store elements,
    contribution-pool elements, largest sum of contribution elements live at once."""
    out = np.zeros(10, dtype=np.int64)
    _check(lib().mg_test_plan_stats(symbolic._h, _ptr(out)))
    return dict(zip(["fronts", "levels", "arena_elems", "max_front", "max_width", "launches", "lookahead_tiles",
                     "factor_elems", "pool_elems", "pool_live_elems"], out.tolist()))


def plan_layout(symbolic):
    """Placement of every device front (host only): (fronts, 6) int64 array of size, width, level,
    parent front (-1: root), factor-store offset, contribution-pool offset."""
    out = np.zeros((plan_stats(symbolic)["fronts"], 6), dtype=np.int64)
    _check(lib().mg_test_plan_layout(symbolic._h, _ptr(out)))
    return out


def front_addr(f, w, off, coff, i, j):
    """Arena offsets of the entries (i[k], j[k]), i >= j, of a front placed at (off, coff)."""
    i = np.ascontiguousarray(i, dtype=np.int32)
    j = np.ascontiguousarray(j, dtype=np.int32)
    out = np.zeros(len(i), dtype=np.int64)
    _check(lib().mg_test_front_addr(f, w, off, coff, len(i), _ptr(i), _ptr(j), _ptr(out)))
    return out


def test_factor_block(numeric, supernode, front_size, width):
    """Factor block of one supernode, (f, w): unit diagonal, L below (include/magla_testing.h)."""
    out = np.zeros(front_size * width)
    _check(lib().mg_test_numeric_front(numeric._h, supernode, _ptr(out)))
    return out.reshape(front_size, width, order="F")


SCHUR_BUCKETS = 6  # front sizes < 128, < 256, < 512, < 1024, < 2048, >= 2048


def test_set_schur_tiles(numeric, tiles):
    """Force the Schur tile edge (64 or 128) per front-size bucket for the handle's precision."""
    t = np.ascontiguousarray(tiles, dtype=np.int32)
    if t.shape != (SCHUR_BUCKETS,):
        raise ArgumentError(MG_E_ARG, f"tiles must have {SCHUR_BUCKETS} entries")
    _check(lib().mg_test_set_schur_tiles(numeric._h, _ptr(t)))


def profile_enable(on=True):
    """Bracket every factorization launch with HIP events (see include/magla_testing.h)."""
    lib().mg_profile_enable(1 if on else 0)


def profile_read():
    """{kernel: dict(launches, ms, flops, bytes)} accumulated since profile_enable()."""
    out = {}
    k = 0
    while True:
        name, n, ms, fl, by = C.c_char_p(), C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        rc = lib().mg_profile_read(k, C.byref(name), C.byref(n), C.byref(ms), C.byref(fl), C.byref(by))
        if rc < 0:
            _check(-rc)
        if rc == 0:
            return out
        out[name.value.decode()] = dict(launches=n.value, ms=ms.value, flops=fl.value, bytes=by.value)
        k += 1


def _check(rc):
    if rc != MG_OK:
        msg = lib().mg_last_error().decode(errors="replace")
        raise _EXC.get(rc, InvalidOperationError)(rc, msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count():
    c = C.c_int(0)
    _check(lib().mg_device_count(C.byref(c)))
    return c.value


# ---------------------------------------------------------------------------
# SymmetricCSRMatrix
# ---------------------------------------------------------------------------
class SymmetricCSRMatrix:
    """Upper-triangle CSR with the diagonal (SymmetricCSRMatrix.cs:19-90)."""

    def __init__(self, size, rowptr, colidx, values):
        self.size = int(size)
        self.row_pointers = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.column_indices = np.ascontiguousarray(colidx, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)

    @classmethod
    def from_coo(cls, size, rows, cols, vals):
        """SymmetricCSRMatrix(int, IEnumerable<MatrixEntry>): normalise, sum duplicates in input order,
        drop |v| <= 1e-15, sort columns (SymmetricCSRMatrix.cs:36-90)."""
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        nnz = C.c_int32(0)
        _check(lib().mg_csr_from_coo(size, len(rows), rows, cols, vals, C.byref(nnz), None, None, None))
        rp = np.zeros(size + 1, dtype=np.int32)
        ci = np.zeros(nnz.value, dtype=np.int32)
        va = np.zeros(nnz.value, dtype=np.float64)
        _check(lib().mg_csr_from_coo(size, len(rows), rows, cols, vals, C.byref(nnz), _ptr(rp), _ptr(ci), _ptr(va)))
        return cls(size, rp, ci, va)

    @property
    def nonzero_count(self):
        return int(self.row_pointers[-1])

    def line_cross(self, rhs, prescribed, indices):
        """LineCross (SymmetricCSRMatrix.cs:125-167) for each index in order; in place."""
        idx = np.ascontiguousarray(np.atleast_1d(indices), dtype=np.int32)
        pv = np.ascontiguousarray(np.broadcast_to(np.asarray(prescribed, dtype=np.float64), idx.shape))
        _check(lib().mg_csr_line_cross(self.size, self.row_pointers, self.column_indices, self.values, rhs,
                                       len(idx), idx, pv))

    def diagonal(self):
        d = np.zeros(self.size)
        rp, ci = self.row_pointers, self.column_indices
        rows = np.repeat(np.arange(self.size), np.diff(rp))
        m = rows == ci
        d[rows[m]] = self.values[m]
        return d


class DeviceMatrix:
    """Device-resident matrix for SpMV and PCG (full CSR in HBM)."""

    def __init__(self, a: SymmetricCSRMatrix):
        self.size = a.size
        h = C.c_void_p()
        _check(lib().mg_matrix_create(a.size, a.row_pointers, _ptr(a.column_indices), _ptr(a.values), C.byref(h)))
        self._h = h

    def multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.size)
        _check(lib().mg_spmv(self._h, x, y))
        return y

    def multiply_dev(self, x_ptr, y_ptr, nrhs=1, ldx=None, ldy=None, stream=None):
        _check(lib().mg_spmv_dev(self._h, nrhs, x_ptr, ldx or self.size, y_ptr, ldy or self.size, stream))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mg_matrix_release(self._h)
            self._h = None


# ---------------------------------------------------------------------------
# UCHOL
# ---------------------------------------------------------------------------
@dataclass
class UtduSolverOptions:
    """UtduSolverOptions (UtduSymbolicFactorization.cs:31-98) plus the factor precision."""
    ordering: int = AMD
    max_degree_of_parallelism: int = 0
    supernode_amalgamation_limit: int = 64
    panel_width: int = 32
    allow_diagonal_regularization: bool = True
    pivot_tolerance: float = 1e-13
    diagonal_regularization: float = 1e-10
    parallel_update_threshold: int = 1 << 20
    precision: int = FP64
    # hipGraph replay: True or 1 = the factorization launches and the solves, 2 = the solves only
    device_graph: int = 0

    def _c(self):
        o = _Options()
        lib().mg_options_default(C.byref(o))
        o.ordering = self.ordering
        o.max_degree_of_parallelism = self.max_degree_of_parallelism
        o.supernode_amalgamation_limit = self.supernode_amalgamation_limit
        o.panel_width = self.panel_width
        o.allow_diagonal_regularization = 1 if self.allow_diagonal_regularization else 0
        o.pivot_tolerance = self.pivot_tolerance
        o.diagonal_regularization = self.diagonal_regularization
        o.parallel_update_threshold = self.parallel_update_threshold
        o.precision = self.precision
        o.device_graph = int(self.device_graph)
        return o


class UtduSymbolicFactorization:
    def __init__(self, handle, options):
        self._h = handle
        self.options = options
        info = _SymInfo()
        n = self._info_raw(info)
        self.size = n
        self.permutation = np.zeros(n, dtype=np.int32)
        self.inverse_permutation = np.zeros(n, dtype=np.int32)
        _check(lib().mg_symbolic_info_get(self._h, C.byref(info), _ptr(self.permutation),
                                          _ptr(self.inverse_permutation)))
        self._info = info

    def _info_raw(self, info):
        _check(lib().mg_symbolic_info_get(self._h, C.byref(info), None, None))
        return info.size

    def info(self):
        info = _SymInfo()
        self._info_raw(info)
        return {k: getattr(info, k) for k, _ in _SymInfo._fields_}

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        keys = {"supernode_count", "max_front_size", "factor_entry_count", "strict_factor_nonzero_count",
                "factor_operation_count", "estimated_factor_bytes", "max_front_bytes", "peak_contribution_bytes",
                "structure_bytes", "max_supernode_width", "level_count"}
        if name in keys:
            return self.info()[name]
        raise AttributeError(name)

    def estimate_peak_bytes(self, worker_count):
        return lib().mg_symbolic_estimate_peak_bytes(self._h, worker_count)

    def structure(self):
        """etree, column counts, fill order, supernode boundaries / parents / row patterns."""
        inf = self.info()
        n, ns = self.size, inf["supernode_count"]
        etree = np.zeros(n, np.int32)
        cc = np.zeros(n, np.int32)
        fo = np.zeros(n, np.int32)
        first = np.zeros(ns + 1, np.int32)
        sp = np.zeros(ns, np.int32)
        pp = np.zeros(ns + 1, np.int32)
        _check(lib().mg_symbolic_structure(self._h, _ptr(etree), _ptr(cc), _ptr(fo), _ptr(first), _ptr(sp),
                                           _ptr(pp), None))
        pr = np.zeros(int(pp[-1]), np.int32)
        _check(lib().mg_symbolic_structure(self._h, None, None, None, None, None, None, _ptr(pr)))
        return dict(etree=etree, colcount=cc, fill_order=fo, first=first, sparent=sp, pat_ptr=pp, pat_rows=pr)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mg_symbolic_release(self._h)
            self._h = None


class UtduNumericFactorization:
    def __init__(self, handle, symbolic):
        self._h = handle
        self.symbolic = symbolic
        self.size = symbolic.size

    def _stats(self, with_diag=False):
        info = _NumInfo()
        diag = np.zeros(self.size) if with_diag else None
        _check(lib().mg_numeric_info_get(self._h, C.byref(info), _ptr(diag)))
        return info, diag

    @property
    def regularized_pivot_count(self):
        return self._stats()[0].regularized_pivot_count

    @property
    def negative_pivot_count(self):
        return self._stats()[0].negative_pivot_count

    @property
    def smallest_pivot_magnitude(self):
        return self._stats()[0].smallest_pivot_magnitude

    @property
    def largest_pivot_magnitude(self):
        return self._stats()[0].largest_pivot_magnitude

    @property
    def is_positive_definite(self):
        return bool(self._stats()[0].is_positive_definite)

    @property
    def diagonal(self):
        return self._stats(True)[1]

    def solve(self, b, tol=0.0, max_refine=0, report=None):
        """Solve(double[]) (UtduNumericFactorization.cs:90-112); b may be (n,) or (n, k)."""
        b = np.asarray(b, dtype=np.float64)
        if b.ndim == 1:
            if b.shape[0] != self.size:
                raise ArgumentError(MG_E_ARG, f"The right-hand side length {b.shape[0]} does not match the matrix "
                                              f"size {self.size}.")
            bb = np.ascontiguousarray(b)
            k = 1
        else:
            if b.shape[0] != self.size:
                raise ArgumentError(MG_E_ARG, "The right-hand side length does not match the matrix size.")
            bb = np.asfortranarray(b)
            k = b.shape[1]
        x = np.zeros_like(bb, order="F") if b.ndim == 2 else np.zeros(self.size)
        rep = _SolveReport()
        _check(lib().mg_solve(self._h, k, _ptr(bb), max(self.size, 1), _ptr(x), max(self.size, 1), tol, max_refine,
                              C.byref(rep)))
        if report is not None:
            report.update(refinement_steps=rep.refinement_steps, converged=bool(rep.converged),
                          relative_residual=rep.relative_residual)
        return x

    def autotune_dev(self, values_ptr, options=None, stream=None):
        """Schur tile autotuner (mg_autotune_dev): returns (tiles per front-size bucket, tuned
        factorization ms); the handle is left factorized from values_ptr."""
        tiles = np.zeros(SCHUR_BUCKETS, dtype=np.int32)
        ms = C.c_double()
        o = (options or self.symbolic.options)._c()
        _check(lib().mg_autotune_dev(self._h, C.c_void_p(values_ptr), C.byref(o), C.c_void_p(stream), _ptr(tiles),
                                     C.byref(ms)))
        return tiles, ms.value

    def solve_dev(self, b_ptr, x_ptr, nrhs, tol=0.0, max_refine=0, stream=None, ldb=None, ldx=None):
        rep = _SolveReport()
        _check(lib().mg_solve_dev(self._h, nrhs, b_ptr, ldb or self.size, x_ptr, ldx or self.size, tol, max_refine,
                                  stream, C.byref(rep)))
        return dict(refinement_steps=rep.refinement_steps, converged=bool(rep.converged),
                    relative_residual=rep.relative_residual)

    def solve_dev_noalloc(self, b_ptr, x_ptr, nrhs, tol=0.0, max_refine=0, stream=None, ldb=None, ldx=None):
        """solve_dev with the workspace in the library's static device scratch (no allocation;
        calls overlapping on one GPU race)."""
        rep = _SolveReport()
        _check(lib().mg_solve_dev_noalloc(self._h, nrhs, b_ptr, ldb or self.size, x_ptr, ldx or self.size, tol,
                                          max_refine, stream, C.byref(rep)))
        return dict(refinement_steps=rep.refinement_steps, converged=bool(rep.converged),
                    relative_residual=rep.relative_residual)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mg_numeric_release(self._h)
            self._h = None


class SymmetricUtduSolver:
    """SymmetricUtduSolver (UCHOL/SymmetricUtduSolver.cs:45-109) on the GPU."""

    def __init__(self, options: UtduSolverOptions | None = None):
        self.options = options or UtduSolverOptions()

    def analyze(self, matrix: SymmetricCSRMatrix) -> UtduSymbolicFactorization:
        if matrix is None:
            raise ArgumentNullError(MG_E_NULL, "matrix")
        h = C.c_void_p()
        _check(lib().mg_analyze(matrix.size, matrix.row_pointers, _ptr(matrix.column_indices),
                                C.byref(self.options._c()), C.byref(h)))
        return UtduSymbolicFactorization(h, self.options)

    def factorize(self, matrix: SymmetricCSRMatrix, symbolic: UtduSymbolicFactorization | None = None):
        if matrix is None:
            raise ArgumentNullError(MG_E_NULL, "matrix")
        if symbolic is None:
            symbolic = self.analyze(matrix)
        h = C.c_void_p()
        _check(lib().mg_factorize(symbolic._h, matrix.size, matrix.nonzero_count, _ptr(matrix.values),
                                  C.byref(self.options._c()), C.byref(h)))
        return UtduNumericFactorization(h, symbolic)

    def factorize_dev(self, symbolic, values_ptr, nnz, numeric=None, stream=None):
        """Values already in HBM; refactorizes in place into `numeric` when given."""
        h = C.c_void_p(numeric._h.value) if numeric is not None else C.c_void_p()
        _check(lib().mg_factorize_dev(symbolic._h, nnz, values_ptr, C.byref(self.options._c()), stream, C.byref(h)))
        return numeric if numeric is not None else UtduNumericFactorization(h, symbolic)

    def solve(self, matrix: SymmetricCSRMatrix, rhs):
        """ILinearSolver.Solve via SolveCore (SymmetricUtduSolver.cs:105-108)."""
        return self.factorize(matrix).solve(rhs)


# ---------------------------------------------------------------------------
# CGGP
# ---------------------------------------------------------------------------
@dataclass
class IterativeSolverResult:
    """CGGP/IterativeSolverResult.cs:8-39."""
    solution: np.ndarray
    iterations: int
    converged: bool
    residual_norm: float
    relative_residual: float


class ConjugateGradientGaussPreSolver:
    """Jacobi-preconditioned CG (CGGP/ConjugateGradientGaussPreSolver.cs:40-196) on the GPU."""

    def __init__(self, relative_tolerance=1e-8, max_iterations=0, use_preconditioner=True):
        self.relative_tolerance = relative_tolerance
        self.max_iterations = max_iterations
        self.use_preconditioner = use_preconditioner
        self.last_result = None

    def solve(self, matrix: SymmetricCSRMatrix | DeviceMatrix, rhs, initial_guess=None):
        self.last_result = None
        dm = matrix if isinstance(matrix, DeviceMatrix) else DeviceMatrix(matrix)
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        if b.shape[0] != dm.size:
            raise ArgumentError(MG_E_ARG, "The right-hand side length does not match the matrix row count.")
        x0 = None
        if initial_guess is not None:
            x0 = np.ascontiguousarray(initial_guess, dtype=np.float64)
            if x0.shape[0] != dm.size:
                raise ArgumentError(MG_E_ARG, f"The initial guess length {x0.shape[0]} does not match the matrix "
                                              f"size {dm.size}.")
        x = np.zeros(dm.size)
        res = _PcgResult()
        _check(lib().mg_pcg_solve(dm._h, b, _ptr(x0), self.relative_tolerance, self.max_iterations,
                                  1 if self.use_preconditioner else 0, x, C.byref(res)))
        self.last_result = IterativeSolverResult(x, res.iterations, bool(res.converged), res.residual_norm,
                                                 res.relative_residual)
        return x

    def solve_dev(self, matrix: "DeviceMatrix", b_ptr, x_ptr, x0_ptr=None, stream=None):
        """Same on HBM vectors; returns the result record (solution left in x_ptr)."""
        res = _PcgResult()
        _check(lib().mg_pcg_solve_dev(matrix._h, C.c_void_p(b_ptr), C.c_void_p(x0_ptr), self.relative_tolerance,
                                      self.max_iterations, 1 if self.use_preconditioner else 0, C.c_void_p(x_ptr),
                                      C.c_void_p(stream), C.byref(res)))
        self.last_result = IterativeSolverResult(None, res.iterations, bool(res.converged), res.residual_norm,
                                                 res.relative_residual)
        return self.last_result


def test_mfma_bf16(a, b):
    """One bf16 MFMA with the learned preconditioner's operand layout: returns a @ b (FP32)."""
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(16, 16)
    b = np.ascontiguousarray(b, dtype=np.float32).reshape(16, 16)
    d = np.zeros((16, 16), dtype=np.float32)
    _check(lib().mg_test_mfma_bf16(_ptr(a), _ptr(b), _ptr(d)))
    return d


class LearnedPreconditioner:
    """Learned correction on top of Jacobi (no counterpart in MAGLA): a three-layer, width-32
    message-passing network on the matrix graph predicts a sparse factor G (two-hop lower
    pattern) and z = G'G r.  Trained on the GPU (FP32, Adam) from (residual, exact error) pairs
    with a Kaporin condition-number term; the network runs once per matrix in BF16 on MFMA."""

    def __init__(self, seed=0):
        h = C.c_void_p()
        _check(lib().mg_mpnn_create(C.c_uint64(seed), C.byref(h)))
        self._h = h

    @staticmethod
    def param_count():
        return int(lib().mg_mpnn_param_count())

    @property
    def params(self):
        p = np.zeros(self.param_count(), dtype=np.float32)
        _check(lib().mg_mpnn_get_params(self._h, _ptr(p)))
        return p

    @params.setter
    def params(self, values):
        p = np.ascontiguousarray(values, dtype=np.float32)
        assert p.size == self.param_count()
        _check(lib().mg_mpnn_set_params(self._h, _ptr(p)))

    def apply_dev(self, matrix: "DeviceMatrix", r_ptr, z_ptr, precision=BF16, stream=None):
        _check(lib().mg_mpnn_apply_dev(self._h, matrix._h, C.c_void_p(r_ptr), C.c_void_p(z_ptr), precision,
                                       C.c_void_p(stream)))

    def loss_grad_dev(self, matrix: "DeviceMatrix", R_ptr, E_ptr, count, ld, kappa=0.0, stream=None):
        loss = C.c_double()
        grad = np.zeros(self.param_count(), dtype=np.float32)
        _check(lib().mg_mpnn_loss_grad_dev(self._h, matrix._h, count, C.c_void_p(R_ptr), C.c_void_p(E_ptr), ld,
                                           float(kappa), C.c_void_p(stream), C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def train_dev(self, matrix: "DeviceMatrix", R_ptr, E_ptr, count, ld, epochs=10, batch=16, learning_rate=1e-3,
                  kappa=1.0, seed=0, stream=None):
        losses = np.zeros(max(epochs, 1))
        _check(lib().mg_mpnn_train_dev(self._h, matrix._h, count, C.c_void_p(R_ptr), C.c_void_p(E_ptr), ld, epochs,
                                       batch, C.c_float(learning_rate), float(kappa), C.c_uint64(seed),
                                       C.c_void_p(stream), _ptr(losses)))
        return losses[:epochs]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mg_mpnn_release(self._h)
            self._h = None


class LearnedConjugateGradientSolver(ConjugateGradientGaussPreSolver):
    """ConjugateGradientGaussPreSolver with the learned preconditioner in a flexible PCG; falls
    back to plain Jacobi-PCG when the learned correction stalls (`fell_back`)."""

    def __init__(self, network: LearnedPreconditioner, relative_tolerance=1e-8, max_iterations=0,
                 stall_window=50):
        super().__init__(relative_tolerance, max_iterations, True)
        self.network = network
        self.stall_window = stall_window
        self.fell_back = False

    def solve(self, matrix, rhs, initial_guess=None):
        self.last_result = None
        dm = matrix if isinstance(matrix, DeviceMatrix) else DeviceMatrix(matrix)
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        if b.shape[0] != dm.size:
            raise ArgumentError(MG_E_ARG, "The right-hand side length does not match the matrix row count.")
        x0 = None if initial_guess is None else np.ascontiguousarray(initial_guess, dtype=np.float64)
        x = np.zeros(dm.size)
        res = _PcgResult()
        fb = C.c_int32()
        _check(lib().mg_pcg_solve_learned(dm._h, self.network._h, b, _ptr(x0), self.relative_tolerance,
                                          self.max_iterations, self.stall_window, x, C.byref(res), C.byref(fb)))
        self.fell_back = bool(fb.value)
        self.last_result = IterativeSolverResult(x, res.iterations, bool(res.converged), res.residual_norm,
                                                 res.relative_residual)
        return x

    def solve_dev(self, matrix: "DeviceMatrix", b_ptr, x_ptr, x0_ptr=None, stream=None):
        res = _PcgResult()
        fb = C.c_int32()
        _check(lib().mg_pcg_solve_learned_dev(matrix._h, self.network._h, C.c_void_p(b_ptr), C.c_void_p(x0_ptr),
                                              self.relative_tolerance, self.max_iterations, self.stall_window,
                                              C.c_void_p(x_ptr), C.c_void_p(stream), C.byref(res), C.byref(fb)))
        self.fell_back = bool(fb.value)
        self.last_result = IterativeSolverResult(None, res.iterations, bool(res.converged), res.residual_norm,
                                                 res.relative_residual)
        return self.last_result


class UtduBatch:
    """Many independent systems factored and solved together (mg_batch_*): each system's
    symbolic phase is the reference's; the numeric phase runs every tree level of all systems
    in the same launches, with per-system pivot scale and per-system residual norms."""

    def __init__(self, matrices, options: UtduSolverOptions | None = None):
        self.options = options or UtduSolverOptions()
        self.matrices = list(matrices)
        self._keep = [(np.ascontiguousarray(m.row_pointers, dtype=np.int32),
                       np.ascontiguousarray(m.column_indices, dtype=np.int32)) for m in self.matrices]
        count = len(self.matrices)
        sizes = np.array([m.size for m in self.matrices], dtype=np.int32)
        rps = (C.c_void_p * max(count, 1))(*[rp.ctypes.data for rp, _ in self._keep])
        cis = (C.c_void_p * max(count, 1))(*[ci.ctypes.data for _, ci in self._keep])
        o = self.options._c()
        h = C.c_void_p()
        _check(lib().mg_batch_analyze(count, _ptr(sizes), rps, cis, C.byref(o), C.byref(h)))
        self._h = h
        self.count = count
        self.row_begin = np.zeros(count + 1, dtype=np.int32)
        self.nnz_begin = np.zeros(count + 1, dtype=np.int32)
        info = _SymInfo()
        _check(lib().mg_batch_layout(h, _ptr(self.row_begin), _ptr(self.nnz_begin), C.byref(info)))
        self._info = {name: getattr(info, name) for name, _ in _SymInfo._fields_}

    def info(self):
        info = _SymInfo()
        _check(lib().mg_batch_layout(self._h, None, None, C.byref(info)))
        return {name: getattr(info, name) for name, _ in _SymInfo._fields_}

    def values(self):
        """Concatenated upper-CSR values of all systems (system order)."""
        return np.concatenate([m.values for m in self.matrices]) if self.count else np.zeros(0)

    def factorize(self, precision=None):
        o = self.options._c()
        if precision is not None:
            o.precision = precision
        vals = [np.ascontiguousarray(m.values, dtype=np.float64) for m in self.matrices]
        arr = (C.c_void_p * max(self.count, 1))(*[v.ctypes.data for v in vals])
        _check(lib().mg_batch_factorize(self._h, arr, C.byref(o)))

    def factorize_dev(self, values_ptr, precision=None, stream=None, device_graph=None):
        o = self.options._c()
        if precision is not None:
            o.precision = precision
        if device_graph is not None:
            o.device_graph = int(device_graph)
        _check(lib().mg_batch_factorize_dev(self._h, C.c_void_p(values_ptr), C.byref(o), C.c_void_p(stream)))

    def autotune_dev(self, values_ptr, precision=None, stream=None):
        """Schur tile autotuner over the merged batch (mg_batch_autotune_dev): returns (tiles per
        front-size bucket, tuned factorization ms); the batch is left factorized."""
        o = self.options._c()
        if precision is not None:
            o.precision = precision
        tiles = np.zeros(SCHUR_BUCKETS, dtype=np.int32)
        ms = C.c_double()
        _check(lib().mg_batch_autotune_dev(self._h, C.c_void_p(values_ptr), C.byref(o), C.c_void_p(stream),
                                           _ptr(tiles), C.byref(ms)))
        return tiles, ms.value

    def solve(self, rhs, tol=1e-10, max_refine=30):
        """rhs: list of (n_t,) or (n_t, k_t) arrays; returns the solutions in the same shapes."""
        bs = [np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(m.size, -1)) for b, m in zip(rhs, self.matrices)]
        xs = [np.zeros_like(b, order="F") for b in bs]
        k = np.array([b.shape[1] for b in bs], dtype=np.int32)
        barr = (C.c_void_p * max(self.count, 1))(*[b.ctypes.data for b in bs])
        xarr = (C.c_void_p * max(self.count, 1))(*[x.ctypes.data for x in xs])
        rep = _SolveReport()
        _check(lib().mg_batch_solve(self._h, _ptr(k), barr, xarr, tol, max_refine, C.byref(rep)))
        self.last_report = dict(refinement_steps=rep.refinement_steps, converged=bool(rep.converged),
                                relative_residual=rep.relative_residual)
        return [x.reshape(np.asarray(b).shape) for x, b in zip(xs, rhs)]

    def solve_dev(self, b_ptr, x_ptr, nrhs, tol=1e-10, max_refine=30, stream=None, ldb=None, ldx=None):
        n = int(self.row_begin[-1])
        rep = _SolveReport()
        _check(lib().mg_batch_solve_dev(self._h, nrhs, C.c_void_p(b_ptr), ldb or n, C.c_void_p(x_ptr), ldx or n, tol,
                                        max_refine, C.c_void_p(stream), C.byref(rep)))
        return dict(refinement_steps=rep.refinement_steps, converged=bool(rep.converged),
                    relative_residual=rep.relative_residual)

    def numeric_info(self):
        info = _NumInfo()
        _check(lib().mg_batch_numeric_info(self._h, C.byref(info)))
        return {name: getattr(info, name) for name, _ in _NumInfo._fields_}

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mg_batch_release(self._h)
            self._h = None


# ---------------------------------------------------------------------------
# BiCGSTAB with ILU(0) on general square CSR matrices
# ---------------------------------------------------------------------------
class CSRMatrix:
    """CSRMatrix (CAESolvers/CSRMatrix.cs): general rows x cols CSR, columns ascending per row.
    The device copy (for SpMV, ILU(0) and BiCGSTAB) is made on first use and dropped when the
    values change."""

    def __init__(self, rows, cols, row_pointers, column_indices, values):
        self.row_count, self.column_count = int(rows), int(cols)
        self.row_pointers = np.ascontiguousarray(row_pointers, dtype=np.int32)
        self.column_indices = np.ascontiguousarray(column_indices, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self._h = None

    @classmethod
    def from_coo(cls, rows, cols, row_idx, col_idx, vals):
        """CSRMatrix(int, int, IEnumerable<MatrixEntry>) (CSRMatrix.cs:34-80)."""
        ri = np.ascontiguousarray(row_idx, dtype=np.int32)
        cj = np.ascontiguousarray(col_idx, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        nnz = C.c_int32()
        _check(lib().mg_gcsr_from_coo(rows, cols, len(ri), ri, cj, v, C.byref(nnz), None, None, None))
        rp = np.zeros(rows + 1, dtype=np.int32)
        ci = np.zeros(nnz.value, dtype=np.int32)
        vv = np.zeros(nnz.value)
        _check(lib().mg_gcsr_from_coo(rows, cols, len(ri), ri, cj, v, C.byref(nnz), _ptr(rp), _ptr(ci), _ptr(vv)))
        return cls(rows, cols, rp, ci, vv)

    @classmethod
    def from_dense(cls, A):
        A = np.asarray(A, dtype=np.float64)
        ri, cj = np.nonzero(A)
        return cls.from_coo(A.shape[0], A.shape[1], ri, cj, A[ri, cj])

    @property
    def nonzero_count(self):
        return int(self.row_pointers[-1])

    def _dev(self):
        if self._h is None:
            h = C.c_void_p()
            _check(lib().mg_gmatrix_create(self.row_count, self.column_count, self.row_pointers,
                                           _ptr(self.column_indices), _ptr(self.values), C.byref(h)))
            self._h = h
        return self._h

    def multiply(self, x):
        """CSRMatrix.Multiply (CSRMatrix.cs:230-248), on the GPU."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape[0] != self.column_count:
            raise ArgumentError(MG_E_ARG, f"Vector length {x.shape[0]} does not match the matrix column count "
                                          f"{self.column_count}.")
        y = np.zeros(self.row_count)
        _check(lib().mg_gspmv(self._dev(), x, y))
        return y

    def line_cross(self, rhs, prescribed_value, index):
        """CSRMatrix.LineCross (CSRMatrix.cs:110-164): rhs is updated in place."""
        if rhs is None:
            raise ArgumentNullError(MG_E_NULL, "rightHandSide")
        if len(rhs) != self.row_count:
            raise ArgumentError(MG_E_ARG, f"The right-hand side length {len(rhs)} does not match the matrix row "
                                          f"count {self.row_count}.")
        idx = np.array([index], dtype=np.int32)
        pv = np.array([prescribed_value], dtype=np.float64)
        _check(lib().mg_gcsr_line_cross(self.row_count, self.column_count, self.row_pointers,
                                        _ptr(self.column_indices), _ptr(self.values), rhs, 1, idx, pv))
        self._release()

    def _release(self):
        if self._h is not None and _lib is not None:
            _lib.mg_gmatrix_release(self._h)
        self._h = None

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            self._release()


class Ilu0Preconditioner:
    """Ilu0Preconditioner (BiCGSTAB/Ilu0Preconditioner.cs): ILU(0) built on the GPU."""

    def __init__(self, matrix: CSRMatrix, pivot_tolerance=1e-30):
        if matrix is None:
            raise ArgumentNullError(MG_E_NULL, "matrix")
        self.matrix = matrix
        self.size = matrix.row_count
        self.pivot_tolerance = pivot_tolerance
        h = C.c_void_p()
        _check(lib().mg_ilu0_create(matrix._dev(), pivot_tolerance, C.byref(h)))
        self._h = h

    def apply(self, right_hand_side):
        """Apply (:44-84): (L U)^-1 right_hand_side."""
        b = np.ascontiguousarray(right_hand_side, dtype=np.float64)
        if b.shape[0] != self.size:
            raise ArgumentError(MG_E_ARG, "The right-hand side length does not match the preconditioner size.")
        x = np.zeros(self.size)
        _check(lib().mg_ilu0_apply(self._h, b, x))
        return x

    def factors(self):
        """Test hook: L (strictly lower, unit diagonal) and U in the matrix's pattern, CSR order."""
        out = np.zeros(self.matrix.nonzero_count)
        _check(lib().mg_test_ilu0_factors(self._h, out))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mg_ilu0_release(self._h)
            self._h = None


class BiCgStabSolver:
    """BiCgStabSolver (BiCGSTAB/BiCgStabSolver.cs) on the GPU."""

    def __init__(self, relative_tolerance=1e-8, max_iterations=0, use_preconditioner=True,
                 breakdown_tolerance=1e-30):
        self.relative_tolerance = relative_tolerance
        self.max_iterations = max_iterations
        self.use_preconditioner = use_preconditioner
        self.breakdown_tolerance = breakdown_tolerance
        self.last_result = None

    def solve(self, matrix: CSRMatrix, right_hand_side):
        return self.solve_with_initial_guess(matrix, right_hand_side, None)

    def solve_with_initial_guess(self, matrix: CSRMatrix, right_hand_side, initial_guess,
                                 preconditioner: Ilu0Preconditioner | None = None):
        self.last_result = None
        if matrix is None:
            raise ArgumentNullError(MG_E_NULL, "matrix")
        b = np.ascontiguousarray(right_hand_side, dtype=np.float64)
        if b.shape[0] != matrix.row_count:
            raise ArgumentError(MG_E_ARG, "The right-hand side length does not match the matrix size.")
        x0 = None
        if initial_guess is not None:
            x0 = np.ascontiguousarray(initial_guess, dtype=np.float64)
            if x0.shape[0] != matrix.row_count:
                raise ArgumentError(MG_E_ARG, "The initial guess length does not match the matrix size.")
        if preconditioner is not None and preconditioner.matrix is not matrix:
            raise ArgumentError(MG_E_ARG, "The preconditioner was built for another matrix.")
        x = np.zeros(matrix.row_count)
        res = _PcgResult()
        _check(lib().mg_bicgstab_solve(matrix._dev(), preconditioner._h if preconditioner is not None else None,
                                       1 if self.use_preconditioner else 0, b, _ptr(x0), self.relative_tolerance,
                                       self.max_iterations, self.breakdown_tolerance, x, C.byref(res)))
        self.last_result = IterativeSolverResult(x, res.iterations, bool(res.converged), res.residual_norm,
                                                 res.relative_residual)
        return x

    def solve_dev(self, matrix: CSRMatrix, b_ptr, x_ptr, x0_ptr=None, preconditioner=None, stream=None):
        res = _PcgResult()
        _check(lib().mg_bicgstab_solve_dev(matrix._dev(), preconditioner._h if preconditioner is not None else None,
                                           1 if self.use_preconditioner else 0, C.c_void_p(b_ptr),
                                           C.c_void_p(x0_ptr) if x0_ptr else None, self.relative_tolerance,
                                           self.max_iterations, self.breakdown_tolerance, C.c_void_p(x_ptr),
                                           C.c_void_p(stream) if stream else None, C.byref(res)))
        self.last_result = IterativeSolverResult(None, res.iterations, bool(res.converged), res.residual_norm,
                                                 res.relative_residual)
        return self.last_result
