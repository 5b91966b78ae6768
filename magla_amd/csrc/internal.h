// internal.h -- host-side types shared by the symbolic phase, the device plan
// and the C ABI of libmagla.so.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/magla.h"

namespace mg {

// Error carrying an mg_status; the C ABI converts it to a return code plus
// the thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct DevicePlan;  // numeric.hip

// Result of the symbolic phase, restating UtduSymbolicFactorization and the
// structures it owns (PermutedSymmetricPattern, SupernodalStructure).
struct Symbolic {
    int n = 0;
    int matrix_nnz = 0;
    mg_options opts{};
    std::vector<int> a_rowptr, a_colidx;  // structure of the analysed matrix (upper CSR)

    std::vector<int> fill_order;  // AMD (or identity) before the postorder
    std::vector<int> perm, iperm;
    std::vector<int> etree, colcount;

    // PermutedSymmetricPattern of the final permutation
    std::vector<int> pptr, prows, lower_start, lvptr, lvidx, dvidx;

    // SupernodalStructure
    int nsuper = 0;
    std::vector<int> first, super_of, sparent, child_ptr, children, pat_ptr, pat_rows;
    int64_t factor_entries = 0, strict_nonzeros = 0, factor_ops = 0, peak_contribution = 0;
    int max_front = 0, max_width = 0;

    // Batched systems merged into one block-diagonal system: first row (original and
    // elimination order alike) of each system, count+1 entries; empty for a single system.
    std::vector<int> sys_begin;

    // device index maps and level schedule: built by the analysis when a device is visible,
    // else by the first factorization (under plan_mu, so concurrent factorizations from one
    // symbolic handle build it once)
    std::shared_ptr<DevicePlan> plan;
    std::shared_ptr<std::mutex> plan_mu = std::make_shared<std::mutex>();

    int front(int s) const { return pat_ptr[s + 1] - pat_ptr[s]; }
    int width(int s) const { return first[s + 1] - first[s]; }
};

// Symbolic phase entry (symbolic.cpp).
void analyze(int n, const int *rowptr, const int *colidx, const mg_options &opts, Symbolic &out);
// Independent systems: analyse each (in parallel over host threads, each result the same as
// analyze() of that system alone), then relabel them into one block-diagonal Symbolic whose
// supernodal forest is the union of the systems' forests.
void analyze_batch(int count, const int *sizes, const int *const *rowptrs, const int *const *colidxs,
                   const mg_options &opts, std::vector<Symbolic> &each);
void merge_symbolic(const std::vector<Symbolic> &each, Symbolic &out);
// ApproximateMinimumDegreeOrdering.Compute over the pattern of an upper CSR.
void amd_order(int n, const int *rowptr, const int *colidx, std::vector<int> &perm);

}  // namespace mg
