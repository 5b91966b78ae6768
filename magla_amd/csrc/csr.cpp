// csr.cpp -- host assembly of a SymmetricCSRMatrix from coordinate entries and
// the Dirichlet LineCross operation, with the reference's semantics:
//   SymmetricCSRMatrix(int, IEnumerable<MatrixEntry>) (SymmetricCSRMatrix.cs:36-90)
//   SymmetricCSRMatrix.LineCross (SymmetricCSRMatrix.cs:125-167)
// The sort is a stable counting sort by (row, col) so duplicates are summed
// in input order, exactly as Enumerable.Sum over a GroupBy group does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "internal.h"

namespace mg {

struct CooResult {
    std::vector<int> rowptr, colidx;
    std::vector<double> values;
};

static void coo_to_csr(int n, int64_t count, const int *rows, const int *cols, const double *vals, CooResult &out) {
    for (int64_t k = 0; k < count; ++k)
        if (rows[k] < 0 || rows[k] >= n || cols[k] < 0 || cols[k] >= n)
            throw Error(MG_E_RANGE, "Indices are out of range.");
    // bucket by normalised row (stable), then by column within each row (stable)
    std::vector<int64_t> rstart(n + 1, 0);
    for (int64_t k = 0; k < count; ++k) ++rstart[std::min(rows[k], cols[k]) + 1];
    for (int i = 0; i < n; ++i) rstart[i + 1] += rstart[i];
    std::vector<int64_t> byrow(count);
    {
        std::vector<int64_t> at(rstart.begin(), rstart.end() - 1);
        for (int64_t k = 0; k < count; ++k) byrow[at[std::min(rows[k], cols[k])]++] = k;
    }
    out.rowptr.assign(n + 1, 0);
    out.colidx.clear();
    out.values.clear();
    std::vector<int64_t> colhead(n, -1), next(count, -1), coltail(n, -1);
    std::vector<int> touched;
    for (int r = 0; r < n; ++r) {
        touched.clear();
        for (int64_t q = rstart[r]; q < rstart[r + 1]; ++q) {
            const int64_t k = byrow[q];
            const int c = std::max(rows[k], cols[k]);
            if (colhead[c] < 0) { colhead[c] = k; touched.push_back(c); }
            else next[coltail[c]] = k;
            coltail[c] = k;
        }
        std::sort(touched.begin(), touched.end());
        for (int c : touched) {
            double sum = 0.0;
            for (int64_t k = colhead[c]; k >= 0; k = next[k]) sum += vals[k];
            colhead[c] = -1;
            if (std::fabs(sum) > 1e-15) {
                out.colidx.push_back(c);
                out.values.push_back(sum);
            }
        }
        out.rowptr[r + 1] = (int)out.colidx.size();
    }
}


void csr_from_coo(int n, int64_t count, const int *rows, const int *cols, const double *vals, int *nnz_out,
                  int *rowptr, int *colidx, double *values) {
    CooResult r;
    coo_to_csr(n, count, rows, cols, vals, r);
    if (r.colidx.size() > 2147483647ULL) throw Error(MG_E_TOO_DENSE, "Too many nonzeros.");
    *nnz_out = (int)r.colidx.size();
    if (rowptr) std::memcpy(rowptr, r.rowptr.data(), sizeof(int) * (n + 1));
    if (colidx && !r.colidx.empty()) std::memcpy(colidx, r.colidx.data(), sizeof(int) * r.colidx.size());
    if (values && !r.values.empty()) std::memcpy(values, r.values.data(), sizeof(double) * r.values.size());
}

// LineCross for many constrained equations, each with the reference's effect:
// off-diagonal entries of row/column `index` are moved to the right-hand side
// (rhs[other] -= a * prescribed) and zeroed, the diagonal becomes 1 and
// rhs[index] = prescribed.  One pass over the matrix per call instead of one
// per equation; the per-equation order of the rhs updates is preserved
// (equations in the given order, entries in CSR order).
void csr_line_cross(int n, const int *rowptr, const int *colidx, double *values, double *rhs, int count,
                    const int *indices, const double *prescribed) {
    std::vector<int> diag(n, -1);
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
            if (colidx[k] == r) diag[r] = k;
    for (int q = 0; q < count; ++q) {
        const int idx = indices[q];
        if (idx < 0 || idx >= n) throw Error(MG_E_RANGE, "index is out of range.");
        if (diag[idx] < 0) throw Error(MG_E_NOT_PD, "Diagonal position is absent from the sparse structure.");
    }
    // the reference visits every stored entry touching `index` in CSR order:
    // entries of rows < index whose column is index, then row index itself,
    // then rows > index never touch it (upper storage).  Build, per
    // constrained equation, its column list once.
    std::vector<std::vector<int>> bycol(n);
    std::vector<char> constrained(n, 0);
    for (int q = 0; q < count; ++q) constrained[indices[q]] = 1;
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
            if (colidx[k] != r && constrained[colidx[k]]) bycol[colidx[k]].push_back(k);
    std::vector<int> rowof(rowptr[n]);
    for (int r = 0; r < n; ++r)
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) rowof[k] = r;
    for (int q = 0; q < count; ++q) {
        const int idx = indices[q];
        const double pv = prescribed ? prescribed[q] : 0.0;
        // positions in CSR order: column entries (rows < idx) come before row idx's own entries
        for (int k : bycol[idx]) {
            rhs[rowof[k]] -= values[k] * pv;
            values[k] = 0.0;
        }
        for (int k = rowptr[idx]; k < rowptr[idx + 1]; ++k) {
            if (colidx[k] == idx) continue;
            rhs[colidx[k]] -= values[k] * pv;
            values[k] = 0.0;
        }
        values[diag[idx]] = 1.0;
        rhs[idx] = pv;
    }
}

// CSRMatrix(int, int, IEnumerable<MatrixEntry>) (CSRMatrix.cs:34-80): duplicates of one
// (row, col) summed in input order, |sum| <= 1e-15 dropped, columns ascending per row.
// Indices are checked as CSRMatrixBuilder.AddToElement does (CSRMatrixBuilder.cs:74-81).
void gcsr_from_coo(int rows, int cols, int64_t count, const int *ri, const int *cj, const double *vals,
                   int *nnz_out, int *rowptr, int *colidx, double *values) {
    for (int64_t k = 0; k < count; ++k)
        if (ri[k] < 0 || ri[k] >= rows || cj[k] < 0 || cj[k] >= cols)
            throw Error(MG_E_RANGE, "Indices are out of range.");
    std::vector<int64_t> rstart(rows + 1, 0);
    for (int64_t k = 0; k < count; ++k) ++rstart[ri[k] + 1];
    for (int i = 0; i < rows; ++i) rstart[i + 1] += rstart[i];
    std::vector<int64_t> byrow(count);
    {
        std::vector<int64_t> at(rstart.begin(), rstart.end() - 1);
        for (int64_t k = 0; k < count; ++k) byrow[at[ri[k]]++] = k;
    }
    std::vector<int> rp(rows + 1, 0), ci;
    std::vector<double> va;
    std::vector<int64_t> colhead(cols, -1), next(count, -1), coltail(cols, -1);
    std::vector<int> touched;
    for (int r = 0; r < rows; ++r) {
        touched.clear();
        for (int64_t q = rstart[r]; q < rstart[r + 1]; ++q) {
            const int64_t k = byrow[q];
            const int c = cj[k];
            if (colhead[c] < 0) { colhead[c] = k; touched.push_back(c); }
            else next[coltail[c]] = k;
            coltail[c] = k;
        }
        std::sort(touched.begin(), touched.end());
        for (int c : touched) {
            double sum = 0.0;
            for (int64_t k = colhead[c]; k >= 0; k = next[k]) sum += vals[k];
            colhead[c] = -1;
            if (std::fabs(sum) > 1e-15) {
                ci.push_back(c);
                va.push_back(sum);
            }
        }
        rp[r + 1] = (int)ci.size();
    }
    if (ci.size() > 2147483647ULL) throw Error(MG_E_TOO_DENSE, "Too many nonzeros.");
    *nnz_out = (int)ci.size();
    if (rowptr) std::memcpy(rowptr, rp.data(), sizeof(int) * (rows + 1));
    if (colidx && !ci.empty()) std::memcpy(colidx, ci.data(), sizeof(int) * ci.size());
    if (values && !va.empty()) std::memcpy(values, va.data(), sizeof(double) * va.size());
}

// CSRMatrix.LineCross (CSRMatrix.cs:110-164) for each constrained equation in the given order:
// the row's off-diagonal values are zeroed, the column's values move to the right-hand side
// (rhs[row] -= a * prescribed) and are zeroed, the diagonal becomes 1, rhs[index] = prescribed.
void gcsr_line_cross(int rows, int cols, const int *rowptr, const int *colidx, double *values, double *rhs,
                     int count, const int *indices, const double *prescribed) {
    if (rows != cols) throw Error(MG_E_NOT_PD, "A boundary condition can be applied only to a square matrix.");
    for (int q = 0; q < count; ++q) {
        const int idx = indices[q];
        if (idx < 0 || idx >= rows) throw Error(MG_E_RANGE, "index is out of range.");
        int dpos = -1;
        for (int k = rowptr[idx]; k < rowptr[idx + 1]; ++k)
            if (colidx[k] == idx) dpos = k;
        if (dpos < 0)
            throw Error(MG_E_NOT_PD, "Diagonal position (" + std::to_string(idx) + ", " + std::to_string(idx) +
                                         ") is absent from the sparse structure.");
        const double pv = prescribed ? prescribed[q] : 0.0;
        for (int r = 0; r < rows; ++r)
            for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
                const int c = colidx[k];
                if (r == idx) {
                    if (c != idx) values[k] = 0.0;
                } else if (c == idx) {
                    rhs[r] -= values[k] * pv;
                    values[k] = 0.0;
                }
            }
        values[dpos] = 1.0;
        rhs[idx] = pv;
    }
}

}  // namespace mg
