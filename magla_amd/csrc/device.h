// device.h -- device-side objects of libmagla.so (implemented in *.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace mg {

#define MG_HIP(x)                                                                               \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            throw ::mg::Error(MG_E_DEVICE, std::string("HIP error '") + hipGetErrorString(e_) + \
                                               "' in " #x);                                     \
    } while (0)

// Owned device allocation.
template <typename T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    DevArray() = default;
    explicit DevArray(size_t count) { alloc(count); }
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    DevArray(DevArray &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevArray &operator=(DevArray &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevArray() { release(); }
    void alloc(size_t count) {
        release();
        if (count) MG_HIP(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const std::vector<T> &v) {
        alloc(v.size());
        if (!v.empty()) MG_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
};

// Device-resident SymmetricCSRMatrix expanded to both triangles for
// row-parallel products.
struct Matrix {
    int n = 0;
    int nnz_upper = 0;
    int64_t nnz_full = 0;
    std::vector<int> h_rowptr, h_colidx;  // upper structure (host copy)
    DevArray<int> rowptr, colidx;         // full structure
    DevArray<int> src;                    // full entry -> upper value index
    DevArray<int> dpos;                   // full-CSR position of A[i,i] or -1
    DevArray<double> vals;                // full values
    DevArray<double> invdiag;             // Jacobi: 1 / A[i,i]
    int diag_status = 0;                  // 0 ok, else MG_E_NOT_PD (zero/invalid diagonal)
    int lanes = 8;                        // lanes per row in the SpMV kernel
    void build(int n, const int *rowptr, const int *colidx);
    void set_values_host(const double *values);
    void set_values_dev(const double *d_values, hipStream_t st);
};

struct DevicePlan;
struct Numeric;

std::shared_ptr<DevicePlan> build_plan(const Symbolic &S);
int plan_level_count(const DevicePlan &P);

// Factorization: values (upper CSR order, float64) already on the device.
void factorize_device(Symbolic &S, const double *d_values, int precision, const mg_options &opts,
                      hipStream_t st, Numeric *&num);
void numeric_release(Numeric *num);
void numeric_info(const Numeric *num, mg_numeric_info &info, double *diag_host);
void solve_device(Numeric *num, int nrhs, const double *d_b, int ldb, double *d_x, int ldx, double tol,
                  int max_refine, hipStream_t st, mg_solve_report *rep);

// SpMV / PCG
void spmv_device(const Matrix &A, int nrhs, const double *d_x, int ldx, double *d_y, int ldy, hipStream_t st);
void pcg_device(const Matrix &A, const double *d_b, const double *d_x0, double tol, int max_it, int use_jacobi,
                double *d_x, hipStream_t st, mg_pcg_result &res);

}  // namespace mg

struct mg_matrix {
    mg::Matrix m;
};
