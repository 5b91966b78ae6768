// device.h -- device-side objects of libmagla.so (implemented in *.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

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

struct MpGraph;

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
    // sliced ELL copy for the PCG direction kernel: slices of 64 rows, slice s holds
    // sell_ptr[s + 1] - sell_ptr[s] = 64 w_s entries, entry k of row 64 s + l at
    // sell_ptr[s] + 64 k + l (padding: column of the row itself, value 0)
    int nslices = 0;
    bool sell = false;                    // built (padding at most half the entries)
    DevArray<int> sell_ptr, sell_ci, sell_src;  // sell_src: upper value index or -1
    DevArray<double> sell_vals;
    std::shared_ptr<MpGraph> mp;          // learned-preconditioner graph data (built on first use)
    void build(int n, const int *rowptr, const int *colidx);
    void set_values_host(const double *values);
    void set_values_dev(const double *d_values, hipStream_t st);
};

struct DevicePlan;
struct Numeric;

std::shared_ptr<DevicePlan> build_plan(const Symbolic &S);
// Builds and uploads S.plan once (thread-safe); a no-op when it exists.
void ensure_plan(Symbolic &S);
int plan_level_count(const DevicePlan &P);

// Factorization: values (upper CSR order, float64) already on the device.
void factorize_device(Symbolic &S, const double *d_values, int precision, const mg_options &opts,
                      hipStream_t st, Numeric *&num);
// Schur tile autotuner (per front-size bucket, per precision, stored in the plan): leaves
// num factorized from d_values; tiles receives 64/128 per bucket, ms the tuned time.
void autotune_device(Symbolic &S, const double *d_values, const mg_options &opts, hipStream_t st, Numeric *&num,
                     int *tiles, double *ms);
void set_schur_tiles(Numeric *num, const int *tiles);
void numeric_release(Numeric *num);
void numeric_info(const Numeric *num, mg_numeric_info &info, double *diag_host);
void solve_device(Numeric *num, int nrhs, const double *d_b, int ldb, double *d_x, int ldx, double tol,
                  int max_refine, hipStream_t st, mg_solve_report *rep);
// The same solve with its workspace in the library's static device scratch (no allocation).
void solve_device_noalloc(Numeric *num, int nrhs, const double *d_b, int ldb, double *d_x, int ldx, double tol,
                          int max_refine, hipStream_t st, mg_solve_report *rep);

// SpMV / PCG
void spmv_device(const Matrix &A, int nrhs, const double *d_x, int ldx, double *d_y, int ldy, hipStream_t st);
void pcg_device(const Matrix &A, const double *d_b, const double *d_x0, double tol, int max_it, int use_jacobi,
                double *d_x, hipStream_t st, mg_pcg_result &res);

// General square/rectangular CSR matrix on the device (CSRMatrix.cs), ILU(0) of it and
// BiCGSTAB (BiCgStabSolver.cs, Ilu0Preconditioner.cs): bicgstab.hip
struct GMatrix {
    int rows = 0, cols = 0, nnz = 0;
    std::vector<int> h_rowptr, h_colidx;
    DevArray<int> rowptr, colidx;
    DevArray<double> vals;
};
struct Ilu0 {
    const GMatrix *A = nullptr;
    int n = 0, nlev_l = 0, nlev_u = 0;
    double pivot_tol = 0.0;
    DevArray<int> diag, lptr, lrows, uptr, urows, bad;  // diagonal positions, row levels of L and U
    DevArray<double> f, w;                              // factors in A's pattern, forward workspace
};
void gmatrix_build(GMatrix &G, int rows, int cols, const int *rowptr, const int *colidx, const double *values);
void gspmv_device(const GMatrix &G, const double *d_x, double *d_y, hipStream_t st);
void ilu0_build(Ilu0 &P, const GMatrix &A, double pivot_tol, hipStream_t st);
void ilu0_apply_device(const Ilu0 &P, const double *d_r, double *d_z, hipStream_t st, const int *done = nullptr);
void bicgstab_device(const GMatrix &A, const Ilu0 *pre, int use_pre, const double *d_b, const double *d_x0,
                     double tol, int max_it, double btol, double *d_x, hipStream_t st, mg_pcg_result &res,
                     std::unique_ptr<Ilu0> &owned);

constexpr int kRedBlocks = 512;  // fixed grid of the deterministic vector reductions

// Device scalars of a (flexible) PCG run; kernels of an iteration turn into no-ops once done.
struct PcgScalars {
    double rz, pap, alpha, beta, rnorm, thr, bnorm, best;
    double rz2[2];  // r'z of iteration it in rz2[it & 1] (flexible PCG: two-kernel iteration)
    double al2[2];  // single-reduction PCG: alpha of iteration it in al2[it & 1] (gamma in rz2)
    int done, iters, fail, fallback, restart, best_it, pad0, pad1;
};

// State of the fused flexible PCG between its kernels, one slot per iteration parity: the
// residual-norm test of iteration it (best norm, stall fallback) is written into slot it & 1
// (fallback1 / restart1), the next iteration's first kernel completes it with the direction
// update (rz2, beta, fallback2 / restart2) in the same slot; kernels read the other slot.
struct FcgState {
    double rnorm, best, rz2, beta;
    int best_it, fallback1, restart1, fallback2, restart2, pad;
};

// ---- learned preconditioner (message-passing network on the matrix graph) ----
constexpr int kMpW = 32;                                      // hidden width
constexpr int kMpOffW = 3 * kMpW;                             // W1_0 (then W2_0, W1_1, ...)
constexpr int kMpOffT = 3 * kMpW + 6 * kMpW * kMpW;           // w_t, then (u1, v1, w1), (u2, v2, w2)
constexpr int kMpP = kMpOffT + 7 * kMpW;                      // parameter count
constexpr int kMpBf16Layer = 2 * 4 * 64 * 4;                  // packed B operands per layer (bf16)
constexpr int kMpBf16Vec = 3 * kMpBf16Layer;                  // then W_in (96) and the readouts (224)
constexpr int kMpBlocks = 512;                                // fixed grid of the gradient kernels

// Per-matrix graph data: e1 = a_ij / sqrt(a_ii a_jj) on the full CSR (0 on the diagonal), node
// features [1, log(a_ii / mean a_ii), sum_j |e1_ij|], a_ii, and the pattern of the learned
// factor G (row i: columns j <= i within two hops, ascending, diagonal last) with the weights
// e1_ij and e2_ij = sum_k e1_ik e1_kj of every entry, plus the transposed pattern.
struct MpGraph {
    int n = 0;
    int64_t ng = 0;
    DevArray<float> ew, x, ge1, ge2;
    DevArray<double> diag;
    DevArray<int> gptr, gcol, gtptr, gtrow, gtidx;
    // sliced ELL copies of G and G' for the solves (64-row slices, lane = row; built when
    // the padding stays under half the entries): columns / rows and the index of each
    // entry's value in the factor values (-1: padding)
    int nslices = 0;
    DevArray<int> gs_ptr, gs_col, gs_src, gts_ptr, gts_row, gts_src;
    // when every sliced entry's column (row, for G') is within 32767 of its slot's row: those
    // differences as 16 bits, so that the fused PCG streams one 32-bit word per entry (BF16
    // value in the high half, the difference in the low half)
    bool packed = false;
    DevArray<uint16_t> gs_dl, gts_dl;
};

struct Mpnn {
    DevArray<float> w;              // FP32 master parameters
    DevArray<uint16_t> wb;          // BF16 copy: packed MFMA B operands (3 layers), W_in, readouts
    DevArray<float> grad, am, av;   // gradient and Adam moments
    int adam_t = 0;
    int ws_n = 0;                   // workspaces below are sized for this n (and ng)
    int64_t ws_ng = 0;
    DevArray<float> H, Mb, Pre, dH, dHp, dM, part, gvf, dG, gaE;
    DevArray<uint16_t> hb0, hb1;
    DevArray<float> gv, gvt;        // factor values for the solves (from the BF16 network), also transposed order
    DevArray<float> gvs, gvts;      // the same in the sliced ELL orders of G and G' (when built)
    DevArray<uint32_t> gps, gpts;   // packed sliced entries (BF16 value | 16-bit index difference), when packed
    const MpGraph *gv_graph = nullptr;  // graph the factor values belong to
    DevArray<double> y, z, az, u, dpart, scal;
};

void mp_graph_prepare(Matrix &A, hipStream_t st);
void mpnn_init(Mpnn &net, uint64_t seed, hipStream_t st);
void mpnn_pack_bf16(Mpnn &net, hipStream_t st);
// z = G'G r with the FP32 network (training precision) or the BF16 MFMA network
void mpnn_apply(Mpnn &net, Matrix &A, const double *d_r, double *d_z, int precision, hipStream_t st);
// factor values from the BF16 network (once per matrix and network), then per call
// z = G'G r with per-block partial sums of r.z and z.ap (mpnn_partial_blocks(n) each); the kernels do
// nothing once S->done or S->fallback
void mpnn_factor_bf16(Mpnn &net, Matrix &A, hipStream_t st);
void mpnn_apply_factor(Mpnn &net, Matrix &A, const double *d_r, double *d_z, const PcgScalars *S, const double *d_ap,
                       double *rz_part, double *zap_part, hipStream_t st);
// The fused flexible PCG's preconditioner step of iteration it (both products on the sliced
// ELL copies of G and G', with the residual-norm test folded in); false when the graph has
// no sliced copies.  rz / zap partials: (nslices + 3) / 4 of each.
bool mpnn_fcg_apply(Mpnn &net, Matrix &A, const double *d_r, double *d_z, const double *d_ap, const double *prr,
                    int nstep, int it, int stall, const FcgState *fs_prev, FcgState *fs_cur, PcgScalars *S,
                    double *rz_part, double *zap_part, hipStream_t st);
int mpnn_fcg_partials(const Matrix &A);
int mpnn_partial_blocks(int n);  // partial sums written by mpnn_apply_factor
// mean loss over ns pairs (columns of R, E with leading dimension ld) plus kappa times the log
// Kaporin number of G A G'; gradient into net.grad
double mpnn_loss_grad(Mpnn &net, Matrix &A, int ns, const double *d_R, const double *d_E, int64_t ld, double kappa,
                      hipStream_t st);
void mpnn_adam_step(Mpnn &net, float lr, hipStream_t st);
void mpnn_train(Mpnn &net, Matrix &A, int ns, const double *d_R, const double *d_E, int64_t ld, int epochs, int batch,
                float lr, double kappa, uint64_t seed, double *losses, hipStream_t st);
void fcg_learned_device(Matrix &A, Mpnn &net, const double *d_b, const double *d_x0, double tol, int max_it,
                        int stall_window, double *d_x, hipStream_t st, mg_pcg_result &res, int &fell_back);

}  // namespace mg

struct mg_matrix {
    mg::Matrix m;
};

struct mg_mpnn {
    mg::Mpnn net;
};
