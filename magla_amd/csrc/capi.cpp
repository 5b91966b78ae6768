// capi.cpp -- extern "C" entry points of libmagla.so (include/magla.h).
// Every call converts mg::Error / std::bad_alloc into an mg_status and keeps
// the message for mg_last_error() on the calling thread.
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>

#include "device.h"
#include "../../include/magla_testing.h"

namespace mg {
void numeric_front(const Numeric *N, int s, double *out);
void selftest_mfma(int precision, double *host_out);
void profile_enable(int on);
void plan_stats(const Symbolic &S, int64_t *out);
void plan_layout(const Symbolic &S, int64_t *out);
int64_t front_addr(int64_t off, int64_t coff, int f, int w, int i, int j);
int profile_read(int kind, const char **name, int64_t *launches, double *ms, double *flops, double *bytes);
void csr_from_coo(int n, int64_t count, const int *rows, const int *cols, const double *vals, int *nnz_out,
                  int *rowptr, int *colidx, double *values);
void csr_line_cross(int n, const int *rowptr, const int *colidx, double *values, double *rhs, int count,
                    const int *indices, const double *prescribed);
void selftest_mfma_bf16(const float *a, const float *b, float *d);
void gcsr_from_coo(int rows, int cols, int64_t count, const int *ri, const int *cj, const double *vals,
                   int *nnz_out, int *rowptr, int *colidx, double *values);
void gcsr_line_cross(int rows, int cols, const int *rowptr, const int *colidx, double *values, double *rhs,
                     int count, const int *indices, const double *prescribed);
}  // namespace mg

struct mg_gmatrix {
    mg::GMatrix g;
};
struct mg_ilu0 {
    mg::Ilu0 p;
};

struct mg_symbolic {
    mg::Symbolic s;
};
struct mg_numeric {
    mg::Numeric *num = nullptr;
    mg_symbolic *owner = nullptr;
};
struct mg_batch {
    mg::Symbolic s;  // merged block-diagonal symbolic factorization
    int count = 0;
    std::vector<int> row_begin, nnz_begin;
    mg::Numeric *num = nullptr;
};

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F &&fn) {
    try {
        fn();
        g_last_error.clear();
        return MG_OK;
    } catch (const mg::Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "Out of host memory.";
        return MG_E_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return MG_E_INTERNAL;
    }
}

void require(bool ok, int code, const char *msg) {
    if (!ok) throw mg::Error(code, msg);
}

void fill_symbolic_info(const mg::Symbolic &S, mg_symbolic_info &info) {
    std::memset(&info, 0, sizeof(info));
    info.size = S.n;
    info.supernode_count = S.nsuper;
    info.max_front_size = S.max_front;
    info.max_supernode_width = S.max_width;
    info.factor_entry_count = S.factor_entries;
    info.strict_factor_nonzero_count = S.strict_nonzeros;
    info.factor_operation_count = S.factor_ops;
    info.estimated_factor_bytes = S.factor_entries * 8;
    info.max_front_bytes = (int64_t)S.max_front * (S.max_front + 1) / 2 * 8;
    info.peak_contribution_bytes = S.peak_contribution * 8;
    info.structure_bytes = ((int64_t)S.prows.size() + (int64_t)S.lvidx.size() + (int64_t)S.pat_rows.size()) * 4;
    info.level_count = S.plan ? mg::plan_level_count(*S.plan) : -1;
    info.matrix_nonzero_count = S.matrix_nnz;
}

// The device plan (index maps, level schedule) is built with the analysis when a device is
// visible, so factorizations never mutate the symbolic handle; without a device (host-only
// use of the symbolic phase) the first factorization builds it.
void plan_if_device(mg::Symbolic &S) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return;
    }
    mg::ensure_plan(S);
}

// Structural checks the C# types guarantee by construction.
void check_csr(int n, const int32_t *rowptr, const int32_t *colidx) {
    require(n >= 0, MG_E_ARG, "Matrix size cannot be negative.");
    require(rowptr != nullptr, MG_E_NULL, "rowptr is null.");
    require(rowptr[0] == 0, MG_E_ARG, "rowptr[0] must be 0.");
    for (int r = 0; r < n; ++r) {
        require(rowptr[r + 1] >= rowptr[r], MG_E_ARG, "rowptr must be non-decreasing.");
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            require(colidx[k] >= r && colidx[k] < n, MG_E_ARG, "Column indices must describe the upper triangle.");
            require(k == rowptr[r] || colidx[k] > colidx[k - 1], MG_E_ARG, "Column indices must ascend within a row.");
        }
    }
}
}  // namespace

extern "C" {

const char *mg_last_error(void) { return g_last_error.c_str(); }

int mg_device_count(int *count) {
    return guarded([&] {
        require(count != nullptr, MG_E_NULL, "count is null.");
        MG_HIP(hipGetDeviceCount(count));
    });
}

void mg_options_default(mg_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->ordering = MG_ORDER_AMD;
    o->max_degree_of_parallelism = 0;
    o->supernode_amalgamation_limit = 64;
    o->panel_width = 32;
    o->allow_diagonal_regularization = 1;
    o->precision = MG_FP64;
    o->pivot_tolerance = 1e-13;
    o->diagonal_regularization = 1e-10;
    o->parallel_update_threshold = 1LL << 20;
}

int mg_csr_from_coo(int32_t n, int64_t count, const int32_t *rows, const int32_t *cols, const double *vals,
                    int32_t *nnz_out, int32_t *rowptr, int32_t *colidx, double *values) {
    return guarded([&] {
        require(nnz_out != nullptr, MG_E_NULL, "nnz_out is null.");
        require(count == 0 || (rows && cols && vals), MG_E_NULL, "elements are null.");
        require(n >= 0 && count >= 0, MG_E_ARG, "Negative size.");
        mg::csr_from_coo(n, count, rows, cols, vals, nnz_out, rowptr, colidx, values);
    });
}

int mg_csr_line_cross(int32_t n, const int32_t *rowptr, const int32_t *colidx, double *values, double *rhs,
                      int32_t count, const int32_t *indices, const double *prescribed) {
    return guarded([&] {
        require(rhs != nullptr, MG_E_NULL, "rightHandSide is null.");
        require(count == 0 || indices, MG_E_NULL, "indices are null.");
        check_csr(n, rowptr, colidx);
        mg::csr_line_cross(n, rowptr, colidx, values, rhs, count, indices, prescribed);
    });
}

int mg_gcsr_from_coo(int32_t rows, int32_t cols, int64_t count, const int32_t *row_idx, const int32_t *col_idx,
                     const double *vals, int32_t *nnz_out, int32_t *rowptr, int32_t *colidx, double *values) {
    return guarded([&] {
        require(nnz_out != nullptr, MG_E_NULL, "nnz_out is null.");
        require(count == 0 || (row_idx && col_idx && vals), MG_E_NULL, "elements are null.");
        require(rows >= 0 && cols >= 0 && count >= 0, MG_E_ARG, "Negative size.");
        mg::gcsr_from_coo(rows, cols, count, row_idx, col_idx, vals, nnz_out, rowptr, colidx, values);
    });
}

int mg_gcsr_line_cross(int32_t rows, int32_t cols, const int32_t *rowptr, const int32_t *colidx, double *values,
                       double *rhs, int32_t count, const int32_t *indices, const double *prescribed) {
    return guarded([&] {
        require(rhs != nullptr, MG_E_NULL, "rightHandSide is null.");
        require(rowptr != nullptr && (rowptr[rows] == 0 || (colidx && values)), MG_E_NULL, "matrix is null.");
        require(count == 0 || indices, MG_E_NULL, "indices are null.");
        mg::gcsr_line_cross(rows, cols, rowptr, colidx, values, rhs, count, indices, prescribed);
    });
}

int mg_gmatrix_create(int32_t rows, int32_t cols, const int32_t *rowptr, const int32_t *colidx,
                      const double *values, mg_gmatrix **out) {
    return guarded([&] {
        require(out != nullptr && rowptr != nullptr, MG_E_NULL, "matrix is null.");
        require(rows >= 0 && cols >= 0, MG_E_ARG, "Matrix dimensions cannot be negative.");
        require(rowptr[rows] == 0 || (colidx && values), MG_E_NULL, "matrix is null.");
        auto *h = new mg_gmatrix();
        try {
            mg::gmatrix_build(h->g, rows, cols, rowptr, colidx, values);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void mg_gmatrix_release(mg_gmatrix *m) { delete m; }

int mg_gspmv_dev(mg_gmatrix *m, const double *d_x, double *d_y, void *stream) {
    return guarded([&] {
        require(m != nullptr && (m->g.rows == 0 || (d_x && d_y)), MG_E_NULL, "vector is null.");
        mg::gspmv_device(m->g, d_x, d_y, (hipStream_t)stream);
    });
}

int mg_gspmv(mg_gmatrix *m, const double *x, double *y) {
    return guarded([&] {
        require(m != nullptr && (m->g.rows == 0 || (x && y)), MG_E_NULL, "vector is null.");
        const mg::GMatrix &G = m->g;
        mg::DevArray<double> dx(std::max(G.cols, 1)), dy(std::max(G.rows, 1));
        if (G.cols) MG_HIP(hipMemcpy(dx.p, x, sizeof(double) * G.cols, hipMemcpyHostToDevice));
        mg::gspmv_device(G, dx.p, dy.p, 0);
        if (G.rows) MG_HIP(hipMemcpy(y, dy.p, sizeof(double) * G.rows, hipMemcpyDeviceToHost));
    });
}

int mg_ilu0_create(mg_gmatrix *m, double pivot_tolerance, mg_ilu0 **out) {
    return guarded([&] {
        require(m != nullptr && out != nullptr, MG_E_NULL, "matrix is null.");
        auto *h = new mg_ilu0();
        try {
            mg::ilu0_build(h->p, m->g, pivot_tolerance, 0);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void mg_ilu0_release(mg_ilu0 *p) { delete p; }

int mg_ilu0_apply_dev(mg_ilu0 *p, const double *d_b, double *d_x, void *stream) {
    return guarded([&] {
        require(p != nullptr && (p->p.n == 0 || (d_b && d_x)), MG_E_NULL, "rightHandSide is null.");
        mg::ilu0_apply_device(p->p, d_b, d_x, (hipStream_t)stream);
    });
}

int mg_ilu0_apply(mg_ilu0 *p, const double *b, double *x) {
    return guarded([&] {
        require(p != nullptr && (p->p.n == 0 || (b && x)), MG_E_NULL, "rightHandSide is null.");
        const int n = p->p.n;
        mg::DevArray<double> db(std::max(n, 1)), dx(std::max(n, 1));
        if (n) MG_HIP(hipMemcpy(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice));
        mg::ilu0_apply_device(p->p, db.p, dx.p, 0);
        if (n) MG_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}

int mg_bicgstab_solve_dev(mg_gmatrix *m, mg_ilu0 *pre, int32_t use_preconditioner, const double *d_b,
                          const double *d_x0, double relative_tolerance, int32_t max_iterations,
                          double breakdown_tolerance, double *d_x, void *stream, mg_pcg_result *result) {
    return guarded([&] {
        require(m != nullptr && result != nullptr, MG_E_NULL, "matrix is null.");
        require(m->g.rows == 0 || (d_b && d_x), MG_E_NULL, "rightHandSide is null.");
        std::unique_ptr<mg::Ilu0> owned;
        mg::bicgstab_device(m->g, pre ? &pre->p : nullptr, use_preconditioner, d_b, d_x0, relative_tolerance,
                            max_iterations, breakdown_tolerance, d_x, (hipStream_t)stream, *result, owned);
        MG_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

int mg_bicgstab_solve(mg_gmatrix *m, mg_ilu0 *pre, int32_t use_preconditioner, const double *b, const double *x0,
                      double relative_tolerance, int32_t max_iterations, double breakdown_tolerance, double *x,
                      mg_pcg_result *result) {
    return guarded([&] {
        require(m != nullptr && result != nullptr, MG_E_NULL, "matrix is null.");
        const int n = m->g.rows;
        require(n == 0 || (b && x), MG_E_NULL, "rightHandSide is null.");
        mg::DevArray<double> db(std::max(n, 1)), dx0(std::max(n, 1)), dx(std::max(n, 1));
        if (n) MG_HIP(hipMemcpy(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice));
        if (n && x0) MG_HIP(hipMemcpy(dx0.p, x0, sizeof(double) * n, hipMemcpyHostToDevice));
        std::unique_ptr<mg::Ilu0> owned;
        mg_pcg_result r{};
        std::exception_ptr err;
        try {
            mg::bicgstab_device(m->g, pre ? &pre->p : nullptr, use_preconditioner, db.p, x0 ? dx0.p : nullptr,
                                relative_tolerance, max_iterations, breakdown_tolerance, dx.p, 0, r, owned);
        } catch (...) {
            err = std::current_exception();
        }
        // the last iterate is returned also when the iteration breaks down
        if (n) MG_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (err) std::rethrow_exception(err);
        *result = r;
    });
}

int mg_analyze(int32_t n, const int32_t *rowptr, const int32_t *colidx, const mg_options *opts, mg_symbolic **out) {
    return guarded([&] {
        require(out != nullptr && opts != nullptr, MG_E_NULL, "matrix or options is null.");
        require(n == 0 || colidx != nullptr || rowptr[n] == 0, MG_E_NULL, "colidx is null.");
        check_csr(n, rowptr, colidx);
        auto h = std::make_unique<mg_symbolic>();
        mg::analyze(n, rowptr, colidx, *opts, h->s);
        plan_if_device(h->s);
        *out = h.release();
    });
}

int mg_symbolic_info_get(const mg_symbolic *sym, mg_symbolic_info *info, int32_t *perm, int32_t *iperm) {
    return guarded([&] {
        require(sym != nullptr && info != nullptr, MG_E_NULL, "symbolic or info is null.");
        const mg::Symbolic &S = sym->s;
        fill_symbolic_info(S, *info);
        if (perm && S.n) std::memcpy(perm, S.perm.data(), sizeof(int) * S.n);
        if (iperm && S.n) std::memcpy(iperm, S.iperm.data(), sizeof(int) * S.n);
    });
}

int mg_symbolic_structure(const mg_symbolic *sym, int32_t *etree, int32_t *colcount, int32_t *fill_order,
                          int32_t *first_column, int32_t *super_parent, int32_t *pattern_pointers,
                          int32_t *pattern_rows) {
    return guarded([&] {
        require(sym != nullptr, MG_E_NULL, "symbolic is null.");
        const mg::Symbolic &S = sym->s;
        auto put = [](int32_t *dst, const std::vector<int> &v) {
            if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(int) * v.size());
        };
        put(etree, S.etree);
        put(colcount, S.colcount);
        put(fill_order, S.fill_order);
        put(first_column, S.first);
        put(super_parent, S.sparent);
        put(pattern_pointers, S.pat_ptr);
        put(pattern_rows, S.pat_rows);
    });
}

int64_t mg_symbolic_estimate_peak_bytes(const mg_symbolic *sym, int32_t worker_count) {
    if (!sym) return -1;
    const mg::Symbolic &S = sym->s;
    const int64_t workers = worker_count > 1 ? worker_count : 1;
    const int64_t max_front_bytes = (int64_t)S.max_front * (S.max_front + 1) / 2 * 8;
    const int64_t retained = std::min<int64_t>(max_front_bytes, (int64_t)(1 << 25) * 8);
    const int64_t structure = ((int64_t)S.prows.size() + (int64_t)S.lvidx.size() + (int64_t)S.pat_rows.size()) * 4;
    return S.factor_entries * 8 + structure + S.peak_contribution * 8 + workers * retained + max_front_bytes;
}

void mg_symbolic_release(mg_symbolic *sym) { delete sym; }

int mg_factorize_dev(mg_symbolic *sym, int32_t nnz, const double *d_values, const mg_options *opts, void *stream,
                     mg_numeric **num) {
    return guarded([&] {
        require(sym != nullptr && opts != nullptr && num != nullptr, MG_E_NULL, "matrix or symbolic is null.");
        mg::Symbolic &S = sym->s;
        // EnsureCompatible (UtduSymbolicFactorization.cs:306-319)
        require(nnz == S.matrix_nnz, MG_E_ARG,
                "The matrix sparse structure has changed since symbolic analysis. Run Analyze again.");
        require(S.matrix_nnz == 0 || d_values != nullptr, MG_E_NULL, "values are null.");
        mg_numeric *h = *num;
        if (h) require(h->owner == sym, MG_E_ARG, "The numeric handle belongs to another symbolic factorization.");
        mg::Numeric *raw = h ? h->num : nullptr;
        mg::factorize_device(S, d_values, opts->precision, *opts, (hipStream_t)stream, raw);
        if (!h) {
            h = new mg_numeric();
            h->num = raw;
            h->owner = sym;
            *num = h;
        }
    });
}

int mg_factorize(mg_symbolic *sym, int32_t n, int32_t nnz, const double *values, const mg_options *opts,
                 mg_numeric **out) {
    return guarded([&] {
        require(sym != nullptr && opts != nullptr && out != nullptr, MG_E_NULL, "matrix or symbolic is null.");
        mg::Symbolic &S = sym->s;
        require(n == S.n, MG_E_ARG, "Matrix size does not match the symbolic factorization.");
        require(nnz == S.matrix_nnz, MG_E_ARG,
                "The matrix sparse structure has changed since symbolic analysis. Run Analyze again.");
        require(nnz == 0 || values != nullptr, MG_E_NULL, "values are null.");
        mg::DevArray<double> dv(std::max(nnz, 1));
        if (nnz) MG_HIP(hipMemcpy(dv.p, values, sizeof(double) * nnz, hipMemcpyHostToDevice));
        mg::Numeric *raw = nullptr;
        mg::factorize_device(S, dv.p, opts->precision, *opts, 0, raw);
        MG_HIP(hipDeviceSynchronize());
        auto *h = new mg_numeric();
        h->num = raw;
        h->owner = sym;
        *out = h;
    });
}

int mg_autotune_dev(mg_numeric *num, const double *d_values, const mg_options *opts, void *stream, int32_t *tiles,
                    double *ms) {
    return guarded([&] {
        require(num != nullptr && opts != nullptr, MG_E_NULL, "numeric or options is null.");
        mg::Symbolic &S = num->owner->s;
        require(S.matrix_nnz == 0 || d_values != nullptr, MG_E_NULL, "values are null.");
        mg::autotune_device(S, d_values, *opts, (hipStream_t)stream, num->num, tiles, ms);
    });
}

int mg_numeric_info_get(const mg_numeric *num, mg_numeric_info *info, double *diag) {
    return guarded([&] {
        require(num != nullptr && info != nullptr, MG_E_NULL, "numeric or info is null.");
        mg::numeric_info(num->num, *info, diag);
    });
}

void mg_numeric_release(mg_numeric *num) {
    if (!num) return;
    mg::numeric_release(num->num);
    delete num;
}

int mg_solve_dev(mg_numeric *num, int32_t nrhs, const double *d_b, int32_t ldb, double *d_x, int32_t ldx, double tol,
                 int32_t max_refine, void *stream, mg_solve_report *report) {
    return guarded([&] {
        require(num != nullptr, MG_E_NULL, "numeric is null.");
        require(nrhs == 0 || (d_b && d_x), MG_E_NULL, "rightHandSide is null.");
        mg::solve_device(num->num, nrhs, d_b, ldb, d_x, ldx, tol, max_refine, (hipStream_t)stream, report);
    });
}

int mg_solve_dev_noalloc(mg_numeric *num, int32_t nrhs, const double *d_b, int32_t ldb, double *d_x, int32_t ldx,
                         double tol, int32_t max_refine, void *stream, mg_solve_report *report) {
    return guarded([&] {
        require(num != nullptr, MG_E_NULL, "numeric is null.");
        require(nrhs == 0 || (d_b && d_x), MG_E_NULL, "rightHandSide is null.");
        mg::solve_device_noalloc(num->num, nrhs, d_b, ldb, d_x, ldx, tol, max_refine, (hipStream_t)stream, report);
    });
}

int mg_solve(mg_numeric *num, int32_t nrhs, const double *b, int32_t ldb, double *x, int32_t ldx, double tol,
             int32_t max_refine, mg_solve_report *report) {
    return guarded([&] {
        require(num != nullptr, MG_E_NULL, "numeric is null.");
        require(nrhs == 0 || (b && x), MG_E_NULL, "rightHandSide is null.");
        const int n = num->owner->s.n;
        require(nrhs >= 0, MG_E_ARG, "nrhs cannot be negative.");
        require(ldb >= n && ldx >= n, MG_E_ARG, "The right-hand side length does not match the matrix size.");
        if (n == 0 || nrhs == 0) {
            if (report) *report = mg_solve_report{0, 1, 0.0};
            return;
        }
        mg::DevArray<double> db((size_t)n * nrhs), dx((size_t)n * nrhs);
        for (int c = 0; c < nrhs; ++c)
            MG_HIP(hipMemcpy(db.p + (size_t)c * n, b + (size_t)c * ldb, sizeof(double) * n, hipMemcpyHostToDevice));
        mg::solve_device(num->num, nrhs, db.p, n, dx.p, n, tol, max_refine, 0, report);
        MG_HIP(hipDeviceSynchronize());
        for (int c = 0; c < nrhs; ++c)
            MG_HIP(hipMemcpy(x + (size_t)c * ldx, dx.p + (size_t)c * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}

int mg_matrix_create(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *values, mg_matrix **out) {
    return guarded([&] {
        require(out != nullptr, MG_E_NULL, "out is null.");
        check_csr(n, rowptr, colidx);
        require(rowptr[n] == 0 || values != nullptr, MG_E_NULL, "values are null.");
        auto h = std::make_unique<mg_matrix>();
        h->m.build(n, rowptr, colidx);
        h->m.set_values_host(values);
        *out = h.release();
    });
}

void mg_matrix_release(mg_matrix *m) { delete m; }

int mg_spmv_dev(mg_matrix *m, int32_t nrhs, const double *d_x, int32_t ldx, double *d_y, int32_t ldy, void *stream) {
    return guarded([&] {
        require(m != nullptr, MG_E_NULL, "matrix is null.");
        mg::spmv_device(m->m, nrhs, d_x, ldx, d_y, ldy, (hipStream_t)stream);
    });
}

int mg_spmv(mg_matrix *m, const double *x, double *y) {
    return guarded([&] {
        require(m != nullptr && x != nullptr && y != nullptr, MG_E_NULL, "argument is null.");
        const int n = m->m.n;
        mg::DevArray<double> dx(std::max(n, 1)), dy(std::max(n, 1));
        MG_HIP(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
        mg::spmv_device(m->m, 1, dx.p, n, dy.p, n, 0);
        MG_HIP(hipDeviceSynchronize());
        MG_HIP(hipMemcpy(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}

int mg_pcg_solve_dev(mg_matrix *m, const double *d_b, const double *d_x0, double tol, int32_t max_iterations,
                     int32_t use_jacobi, double *d_x, void *stream, mg_pcg_result *result) {
    return guarded([&] {
        require(m != nullptr && result != nullptr, MG_E_NULL, "argument is null.");
        require(m->m.n == 0 || (d_b && d_x), MG_E_NULL, "rightHandSide is null.");
        mg::pcg_device(m->m, d_b, d_x0, tol, max_iterations, use_jacobi, d_x, (hipStream_t)stream, *result);
    });
}

int mg_pcg_solve(mg_matrix *m, const double *b, const double *x0, double tol, int32_t max_iterations,
                 int32_t use_jacobi, double *x, mg_pcg_result *result) {
    return guarded([&] {
        require(m != nullptr && result != nullptr, MG_E_NULL, "argument is null.");
        const int n = m->m.n;
        require(n == 0 || (b && x), MG_E_NULL, "rightHandSide is null.");
        mg::DevArray<double> db(std::max(n, 1)), dx(std::max(n, 1)), dx0;
        if (n) MG_HIP(hipMemcpy(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice));
        if (x0 && n) {
            dx0.alloc(n);
            MG_HIP(hipMemcpy(dx0.p, x0, sizeof(double) * n, hipMemcpyHostToDevice));
        }
        mg::pcg_device(m->m, db.p, x0 ? dx0.p : nullptr, tol, max_iterations, use_jacobi, dx.p, 0, *result);
        MG_HIP(hipDeviceSynchronize());
        if (n) MG_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}

int mg_batch_analyze(int32_t count, const int32_t *sizes, const int32_t *const *rowptrs,
                     const int32_t *const *colidxs, const mg_options *opts, mg_batch **out) {
    return guarded([&] {
        require(out != nullptr, MG_E_NULL, "out is null.");
        require(count >= 0, MG_E_ARG, "The system count cannot be negative.");
        require(count == 0 || (sizes && rowptrs && colidxs), MG_E_NULL, "matrix arrays are null.");
        mg_options o;
        mg_options_default(&o);
        if (opts) o = *opts;
        for (int t = 0; t < count; ++t) check_csr(sizes[t], rowptrs[t], colidxs[t]);
        auto *b = new mg_batch();
        try {
            std::vector<mg::Symbolic> each;
            mg::analyze_batch(count, sizes, rowptrs, colidxs, o, each);
            mg::merge_symbolic(each, b->s);
            b->count = count;
            b->row_begin = b->s.sys_begin;
            b->nnz_begin.assign(count + 1, 0);
            for (int t = 0; t < count; ++t) b->nnz_begin[t + 1] = b->nnz_begin[t] + rowptrs[t][sizes[t]];
            plan_if_device(b->s);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int mg_batch_layout(const mg_batch *b, int32_t *row_begin, int32_t *nnz_begin, mg_symbolic_info *info) {
    return guarded([&] {
        require(b != nullptr, MG_E_NULL, "batch is null.");
        if (row_begin) std::copy(b->row_begin.begin(), b->row_begin.end(), row_begin);
        if (nnz_begin) std::copy(b->nnz_begin.begin(), b->nnz_begin.end(), nnz_begin);
        if (info) fill_symbolic_info(b->s, *info);
    });
}

int mg_batch_factorize_dev(mg_batch *b, const double *d_values, const mg_options *opts, void *stream) {
    return guarded([&] {
        require(b != nullptr, MG_E_NULL, "batch is null.");
        require(b->s.matrix_nnz == 0 || d_values, MG_E_NULL, "values are null.");
        mg_options o = b->s.opts;
        if (opts) o = *opts;
        mg::factorize_device(b->s, d_values, o.precision, o, (hipStream_t)stream, b->num);
    });
}

int mg_batch_autotune_dev(mg_batch *b, const double *d_values, const mg_options *opts, void *stream, int32_t *tiles,
                          double *ms) {
    return guarded([&] {
        require(b != nullptr, MG_E_NULL, "batch is null.");
        require(b->s.matrix_nnz == 0 || d_values, MG_E_NULL, "values are null.");
        mg_options o = b->s.opts;
        if (opts) o = *opts;
        mg::autotune_device(b->s, d_values, o, (hipStream_t)stream, b->num, tiles, ms);
    });
}

int mg_batch_solve_dev(mg_batch *b, int32_t nrhs, const double *d_b, int32_t ldb, double *d_x, int32_t ldx,
                       double tol, int32_t max_refine, void *stream, mg_solve_report *report) {
    return guarded([&] {
        require(b != nullptr, MG_E_NULL, "batch is null.");
        require(b->num != nullptr, MG_E_ARG, "The batch has not been factorized.");
        require(b->s.n == 0 || nrhs == 0 || (d_b && d_x), MG_E_NULL, "rightHandSide is null.");
        mg::solve_device(b->num, nrhs, d_b, ldb, d_x, ldx, tol, max_refine, (hipStream_t)stream, report);
    });
}

int mg_batch_factorize(mg_batch *b, const double *const *values, const mg_options *opts) {
    return guarded([&] {
        require(b != nullptr, MG_E_NULL, "batch is null.");
        require(b->count == 0 || values, MG_E_NULL, "values are null.");
        const int64_t nnz = b->s.matrix_nnz;
        std::vector<double> all((size_t)std::max<int64_t>(nnz, 1));
        for (int t = 0; t < b->count; ++t)
            std::copy(values[t], values[t] + (b->nnz_begin[t + 1] - b->nnz_begin[t]), all.begin() + b->nnz_begin[t]);
        mg::DevArray<double> dv(all.size());
        MG_HIP(hipMemcpy(dv.p, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice));
        mg_options o = b->s.opts;
        if (opts) o = *opts;
        mg::factorize_device(b->s, dv.p, o.precision, o, 0, b->num);
        MG_HIP(hipDeviceSynchronize());
    });
}

int mg_batch_solve(mg_batch *b, const int32_t *nrhs, const double *const *rhs, double *const *x, double tol,
                   int32_t max_refine, mg_solve_report *report) {
    return guarded([&] {
        require(b != nullptr, MG_E_NULL, "batch is null.");
        require(b->num != nullptr, MG_E_ARG, "The batch has not been factorized.");
        require(b->count == 0 || (nrhs && rhs && x), MG_E_NULL, "argument is null.");
        int k = 0;
        for (int t = 0; t < b->count; ++t) {
            require(nrhs[t] >= 0, MG_E_ARG, "The number of right-hand sides cannot be negative.");
            k = std::max(k, nrhs[t]);
        }
        const int n = b->s.n;
        if (k == 0 || n == 0) {
            if (report) *report = mg_solve_report{0, 1, 0.0};
            return;
        }
        std::vector<double> B((size_t)n * k, 0.0), X((size_t)n * k, 0.0);
        for (int t = 0; t < b->count; ++t) {
            const int r0 = b->row_begin[t], nt = b->row_begin[t + 1] - r0;
            for (int c = 0; c < nrhs[t]; ++c)
                std::copy(rhs[t] + (size_t)c * nt, rhs[t] + (size_t)(c + 1) * nt, B.begin() + (size_t)c * n + r0);
        }
        mg::DevArray<double> db(B.size()), dx(X.size());
        MG_HIP(hipMemcpy(db.p, B.data(), sizeof(double) * B.size(), hipMemcpyHostToDevice));
        mg::solve_device(b->num, k, db.p, n, dx.p, n, tol, max_refine, 0, report);
        MG_HIP(hipDeviceSynchronize());
        MG_HIP(hipMemcpy(X.data(), dx.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost));
        for (int t = 0; t < b->count; ++t) {
            const int r0 = b->row_begin[t], nt = b->row_begin[t + 1] - r0;
            for (int c = 0; c < nrhs[t]; ++c)
                std::copy(X.begin() + (size_t)c * n + r0, X.begin() + (size_t)c * n + r0 + nt, x[t] + (size_t)c * nt);
        }
    });
}

int mg_batch_numeric_info(const mg_batch *b, mg_numeric_info *info) {
    return guarded([&] {
        require(b != nullptr && info != nullptr, MG_E_NULL, "argument is null.");
        require(b->num != nullptr, MG_E_ARG, "The batch has not been factorized.");
        mg::numeric_info(b->num, *info, nullptr);
    });
}

void mg_batch_release(mg_batch *b) {
    if (!b) return;
    if (b->num) mg::numeric_release(b->num);
    delete b;
}

int32_t mg_mpnn_param_count(void) { return mg::kMpP; }

int mg_mpnn_create(uint64_t seed, mg_mpnn **out) {
    return guarded([&] {
        require(out != nullptr, MG_E_NULL, "out is null.");
        auto *h = new mg_mpnn();
        try {
            mg::mpnn_init(h->net, seed, 0);
            MG_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void mg_mpnn_release(mg_mpnn *net) { delete net; }

int mg_mpnn_get_params(const mg_mpnn *net, float *params) {
    return guarded([&] {
        require(net != nullptr && params != nullptr, MG_E_NULL, "argument is null.");
        MG_HIP(hipMemcpy(params, net->net.w.p, sizeof(float) * mg::kMpP, hipMemcpyDeviceToHost));
    });
}

int mg_mpnn_set_params(mg_mpnn *net, const float *params) {
    return guarded([&] {
        require(net != nullptr && params != nullptr, MG_E_NULL, "argument is null.");
        MG_HIP(hipMemcpy(net->net.w.p, params, sizeof(float) * mg::kMpP, hipMemcpyHostToDevice));
        mg::mpnn_pack_bf16(net->net, 0);
        MG_HIP(hipDeviceSynchronize());
    });
}

int mg_mpnn_apply_dev(mg_mpnn *net, mg_matrix *m, const double *d_r, double *d_z, int32_t precision,
                      void *stream) {
    return guarded([&] {
        require(net != nullptr && m != nullptr, MG_E_NULL, "argument is null.");
        require(m->m.n == 0 || (d_r && d_z), MG_E_NULL, "vector is null.");
        require(precision == MG_FP32 || precision == MG_BF16, MG_E_ARG, "precision must be MG_FP32 or MG_BF16.");
        const hipStream_t st = (hipStream_t)stream;
        mg::mpnn_apply(net->net, m->m, d_r, d_z, precision, st);
    });
}

int mg_mpnn_loss_grad_dev(mg_mpnn *net, mg_matrix *m, int32_t ns, const double *d_R, const double *d_E, int64_t ld,
                          double kappa, void *stream, double *loss, float *grad) {
    return guarded([&] {
        require(net != nullptr && m != nullptr && loss != nullptr, MG_E_NULL, "argument is null.");
        require(ns >= 0 && ld >= m->m.n, MG_E_ARG, "sample count or leading dimension is invalid.");
        require(ns == 0 || (d_R && d_E), MG_E_NULL, "samples are null.");
        const hipStream_t st = (hipStream_t)stream;
        *loss = mg::mpnn_loss_grad(net->net, m->m, ns, d_R, d_E, ld, kappa, st);
        if (grad) {
            MG_HIP(hipMemcpyAsync(grad, net->net.grad.p, sizeof(float) * mg::kMpP, hipMemcpyDeviceToHost, st));
            MG_HIP(hipStreamSynchronize(st));
        }
    });
}

int mg_mpnn_train_dev(mg_mpnn *net, mg_matrix *m, int32_t ns, const double *d_R, const double *d_E, int64_t ld,
                      int32_t epochs, int32_t batch, float learning_rate, double kappa, uint64_t seed, void *stream,
                      double *epoch_losses) {
    return guarded([&] {
        require(net != nullptr && m != nullptr, MG_E_NULL, "argument is null.");
        require(ns >= 0 && ld >= m->m.n && epochs >= 0, MG_E_ARG, "invalid training arguments.");
        require(ns == 0 || (d_R && d_E), MG_E_NULL, "samples are null.");
        mg::mpnn_train(net->net, m->m, ns, d_R, d_E, ld, epochs, batch, learning_rate, kappa, seed, epoch_losses,
                       (hipStream_t)stream);
    });
}

int mg_pcg_solve_learned_dev(mg_matrix *m, mg_mpnn *net, const double *d_b, const double *d_x0, double tol,
                             int32_t max_iterations, int32_t stall_window, double *d_x, void *stream,
                             mg_pcg_result *result, int32_t *fell_back) {
    return guarded([&] {
        require(m != nullptr && net != nullptr && result != nullptr, MG_E_NULL, "argument is null.");
        require(m->m.n == 0 || (d_b && d_x), MG_E_NULL, "rightHandSide is null.");
        int fb = 0;
        mg::fcg_learned_device(m->m, net->net, d_b, d_x0, tol, max_iterations, stall_window, d_x,
                               (hipStream_t)stream, *result, fb);
        if (fell_back) *fell_back = fb;
    });
}

int mg_pcg_solve_learned(mg_matrix *m, mg_mpnn *net, const double *b, const double *x0, double tol,
                         int32_t max_iterations, int32_t stall_window, double *x, mg_pcg_result *result,
                         int32_t *fell_back) {
    return guarded([&] {
        require(m != nullptr && net != nullptr && result != nullptr, MG_E_NULL, "argument is null.");
        const int n = m->m.n;
        require(n == 0 || (b && x), MG_E_NULL, "rightHandSide is null.");
        mg::DevArray<double> db(std::max(n, 1)), dx(std::max(n, 1)), dx0;
        if (n) MG_HIP(hipMemcpy(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice));
        if (x0 && n) {
            dx0.alloc(n);
            MG_HIP(hipMemcpy(dx0.p, x0, sizeof(double) * n, hipMemcpyHostToDevice));
        }
        int fb = 0;
        mg::fcg_learned_device(m->m, net->net, db.p, x0 ? dx0.p : nullptr, tol, max_iterations, stall_window, dx.p,
                               0, *result, fb);
        MG_HIP(hipDeviceSynchronize());
        if (n) MG_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (fell_back) *fell_back = fb;
    });
}

int mg_test_mfma_bf16(const float *a, const float *b, float *d) {
    return guarded([&] {
        require(a && b && d, MG_E_NULL, "argument is null.");
        mg::selftest_mfma_bf16(a, b, d);
    });
}

int mg_test_numeric_front(const mg_numeric *num, int32_t supernode, double *out) {
    return guarded([&] {
        require(num != nullptr && out != nullptr, MG_E_NULL, "argument is null.");
        mg::numeric_front(num->num, supernode, out);
    });
}

int mg_test_ilu0_factors(const mg_ilu0 *p, double *out) {
    return guarded([&] {
        require(p != nullptr && out != nullptr, MG_E_NULL, "preconditioner or out is null.");
        const int nnz = p->p.A ? p->p.A->nnz : 0;
        if (nnz) MG_HIP(hipMemcpy(out, p->p.f.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    });
}

int mg_test_set_schur_tiles(mg_numeric *num, const int32_t *tiles) {
    return guarded([&] {
        require(num != nullptr && tiles != nullptr, MG_E_NULL, "numeric or tiles is null.");
        mg::set_schur_tiles(num->num, tiles);
    });
}

int mg_test_mfma_layout(int32_t precision, double *out) {
    return guarded([&] {
        require(out != nullptr, MG_E_NULL, "out is null.");
        mg::selftest_mfma(precision, out);
    });
}

int mg_test_plan_stats(const mg_symbolic *sym, int64_t *out) {
    return guarded([&] {
        require(sym != nullptr && out != nullptr, MG_E_NULL, "argument is null.");
        mg::plan_stats(sym->s, out);
    });
}

int mg_test_plan_layout(const mg_symbolic *sym, int64_t *out) {
    return guarded([&] {
        require(sym != nullptr && out != nullptr, MG_E_NULL, "argument is null.");
        mg::plan_layout(sym->s, out);
    });
}

int mg_test_front_addr(int32_t f, int32_t w, int64_t off, int64_t coff, int32_t n, const int32_t *i,
                       const int32_t *j, int64_t *out) {
    return guarded([&] {
        require(n == 0 || (i != nullptr && j != nullptr && out != nullptr), MG_E_NULL, "argument is null.");
        require(f >= 0 && w >= 0 && w <= f && n >= 0, MG_E_RANGE, "f, w or n is out of range.");
        for (int32_t k = 0; k < n; ++k) {
            require(i[k] >= j[k] && j[k] >= 0 && i[k] < f, MG_E_RANGE, "entry is not in the front's lower triangle.");
            out[k] = mg::front_addr(off, coff, f, w, i[k], j[k]);
        }
    });
}

void mg_profile_enable(int32_t on) { mg::profile_enable(on); }

int mg_profile_read(int32_t kind, const char **name, int64_t *launches, double *ms, double *flops, double *bytes) {
    int found = 0;
    const int rc = guarded([&] {
        require(name && launches && ms && flops && bytes, MG_E_NULL, "argument is null.");
        found = mg::profile_read(kind, name, launches, ms, flops, bytes);
    });
    return rc != MG_OK ? -rc : found;
}

}  // extern "C"
