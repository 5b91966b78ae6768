// symbolic.cpp -- host symbolic phase of the UCHOL solver: approximate minimum
// degree ordering, permuted pattern, elimination tree, column counts,
// supernode partition and front row structures.  Integer results must equal
// MAGLA's C# symbolic phase exactly, so every rule (tie breaks, bucket order,
// hash, compaction points, amalgamation thresholds) is kept as the reference
// states it; the data layout and code organisation are this library's own.
// Built with -fwrapv: the reference's int arithmetic wraps on overflow.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <numeric>
#include <thread>

#include "internal.h"

namespace mg {
namespace {

constexpr int kNone = -1;
inline int flipped(int i) { return -i - 2; }

// ---------------------------------------------------------------------------
// Approximate minimum degree (AMD/ApproximateMinimumDegreeOrdering.cs,
// AMD/FactorGraph.cs, AMD/DegreeBuckets.cs, AMD/HashBuckets.cs,
// AMD/MemberLists.cs, AMD/SymmetricPatternGraph.cs).
// ---------------------------------------------------------------------------
class MinimumDegree {
  public:
    MinimumDegree(int n, const int *rowptr, const int *colidx) : n_(n) {
        // adjacency in both directions, degree pass then placement pass
        // (SymmetricPatternGraph.FromMatrix :71-121)
        std::vector<int> start(n + 1, 0);
        for (int r = 0; r < n; ++r)
            for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
                if (colidx[k] != r) { ++start[r + 1]; ++start[colidx[k] + 1]; }
        for (int i = 0; i < n; ++i) start[i + 1] += start[i];
        const int edges = start[n];
        // FactorGraph ctor (:171-195): head room of a fifth plus n+16
        capacity_ = edges + edges / 5 + n + 16;
        adj_.assign(capacity_, 0);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
                const int c = colidx[k];
                if (c == r) continue;
                adj_[fill[r]++] = c;
                adj_[fill[c]++] = r;
            }
        tail_ = edges;
        head_.resize(n);
        len_.resize(n);
        elen_.assign(n, 0);
        nv_.assign(n, 1);
        mark_.assign(n, 1);
        deg_.resize(n);
        for (int i = 0; i < n; ++i) {
            head_[i] = start[i];
            len_[i] = start[i + 1] - start[i];
            deg_[i] = len_[i];
        }
        bucket_first_.assign(n + 1, kNone);
        bucket_next_.assign(n, kNone);
        bucket_prev_.assign(n, kNone);
        hash_first_.assign(n, kNone);
        hash_next_.assign(n, kNone);
        hash_prev_.assign(n, kNone);
        hash_used_.assign(n, 0);
        group_last_.resize(n);
        group_next_.assign(n, kNone);
        std::iota(group_last_.begin(), group_last_.end(), 0);
        mark_limit_ = 2147483647 - n - 1;
    }

    void run(std::vector<int> &order) {
        order_.assign(n_, 0);
        cursor_ = 0;
        int done = seed_buckets();
        flag_ = 2;
        min_bucket_ = 0;
        int widest = 0;
        while (done < n_) {
            reserve_marks(2 * n_ + 2);
            const int piv = pop_lowest();
            if (piv == kNone) throw Error(MG_E_INTERNAL, "AMD: no vertex was found for elimination in a non-empty graph.");
            const int piv_elems = elen_[piv];
            int piv_weight = nv_[piv];
            done += piv_weight;
            nv_[piv] = -piv_weight;
            int degree_sum, lo, hi;
            if (piv_elems == 0)
                gather_variables(piv, degree_sum, lo, hi);
            else
                gather_elements(piv, piv_elems, degree_sum, lo, hi);
            head_[piv] = lo;
            len_[piv] = hi - lo + 1;
            outside_counts(lo, hi);
            new_degrees(piv, lo, hi, degree_sum, piv_weight, done);
            deg_[piv] = degree_sum;
            elen_[piv] = flipped(piv_weight + degree_sum);
            widest = std::max(widest, degree_sum);
            merge_twins();
            reinsert(piv, piv_elems, degree_sum, piv_weight, done, lo, hi);
            output(piv);
            flag_ += widest + 1;
        }
        for (int t = 0; t < (int)dense_.size(); ++t) output(dense_[t]);
        if (cursor_ != n_)
            throw Error(MG_E_INTERNAL, "AMD: an incomplete permutation was built. This is an internal ordering algorithm error.");
        order.swap(order_);
    }

  private:
    // -- degree buckets (DegreeBuckets.cs:56-114) --
    void bucket_push(int v, int d) {
        const int nxt = bucket_first_[d];
        if (nxt != kNone) bucket_prev_[nxt] = v;
        bucket_next_[v] = nxt;
        bucket_prev_[v] = kNone;
        bucket_first_[d] = v;
        if (d < min_bucket_) min_bucket_ = d;
    }
    void bucket_drop(int v) {
        const int before = bucket_prev_[v], after = bucket_next_[v];
        if (after != kNone) bucket_prev_[after] = before;
        if (before != kNone) bucket_next_[before] = after;
        else bucket_first_[deg_[v]] = after;
    }
    int pop_lowest() {
        for (int d = min_bucket_; d <= n_; ++d) {
            const int v = bucket_first_[d];
            if (v == kNone) continue;
            min_bucket_ = d;
            const int after = bucket_next_[v];
            if (after != kNone) bucket_prev_[after] = kNone;
            bucket_first_[d] = after;
            return v;
        }
        return kNone;
    }
    // -- supervariable hash chains (HashBuckets.cs:38-102) --
    void hash_insert(int64_t key, int v) {
        const int b = (int)(key % n_);
        const int old = hash_first_[b];
        if (old == kNone) hash_used_[hash_count_++] = b;
        else hash_prev_[old] = v;
        hash_next_[v] = old;
        hash_prev_[v] = kNone;
        hash_first_[b] = v;
    }
    void hash_remove(int v) {
        const int before = hash_prev_[v], after = hash_next_[v];
        hash_next_[before] = after;
        if (after != kNone) hash_prev_[after] = before;
    }
    // -- member groups (MemberLists.cs:42-47) --
    void absorb_group(int into, int from) {
        group_next_[group_last_[into]] = from;
        group_last_[into] = group_last_[from];
    }
    void output(int rep) {
        for (int v = rep; v != kNone; v = group_next_[v]) order_[cursor_++] = v;
    }

    // ApproximateMinimumDegreeOrdering.cs:647-659
    void reserve_marks(int need) {
        if (flag_ + need < mark_limit_) return;
        for (int x = 0; x < n_; ++x)
            if (mark_[x] != 0) mark_[x] = 1;
        flag_ = 2;
    }

    // ApproximateMinimumDegreeOrdering.cs:231-269
    int seed_buckets() {
        int threshold = (int)(10.0 * std::sqrt((double)n_));
        threshold = std::min(n_, std::max(16, threshold));
        int done = 0;
        for (int v = 0; v < n_; ++v) {
            const int d = deg_[v];
            if (d == 0) {
                head_[v] = kNone; mark_[v] = 0; nv_[v] = 0; elen_[v] = kNone;
                ++done;
                output(v);
            } else if (d > threshold) {
                head_[v] = kNone; nv_[v] = 0; elen_[v] = kNone;
                ++done;
                dense_.push_back(v);
            } else {
                bucket_push(v, d);
            }
        }
        return done;
    }

    // FactorGraph.Compact (FactorGraph.cs:95-143)
    void squeeze(int &lo) {
        std::vector<std::pair<int, int>> live;
        live.reserve(n_);
        for (int v = 0; v < n_; ++v)
            if (head_[v] >= 0 && len_[v] > 0) live.emplace_back(head_[v], v);
        std::sort(live.begin(), live.end());
        int dst = 0;
        for (const auto &pv : live) {
            const int v = pv.second, src = head_[v], cnt = len_[v];
            if (dst != src) std::memmove(&adj_[dst], &adj_[src], sizeof(int) * (size_t)cnt);
            head_[v] = dst;
            dst += cnt;
        }
        const int partial = tail_ - lo;
        if (partial > 0 && dst != lo) std::memmove(&adj_[dst], &adj_[lo], sizeof(int) * (size_t)partial);
        lo = dst;
        tail_ = dst + partial;
        const int need = tail_ + n_ + 16;
        if (need > capacity_) {
            capacity_ = std::max(need, capacity_ + capacity_ / 2);
            adj_.resize(capacity_);
        }
    }

    // ApproximateMinimumDegreeOrdering.cs:285-304
    void gather_variables(int piv, int &degree_sum, int &lo, int &hi) {
        degree_sum = 0;
        lo = head_[piv];
        hi = lo - 1;
        const int end = lo + len_[piv];
        for (int p = lo; p < end; ++p) {
            const int v = adj_[p], wt = nv_[v];
            if (wt <= 0) continue;
            degree_sum += wt;
            nv_[v] = -wt;
            adj_[++hi] = v;
            bucket_drop(v);
        }
    }

    // ApproximateMinimumDegreeOrdering.cs:313-393
    void gather_elements(int piv, int piv_elems, int &degree_sum, int &lo, int &hi) {
        degree_sum = 0;
        int p = head_[piv];
        lo = tail_;
        const int own_vars = len_[piv] - piv_elems;
        for (int pass = 1; pass <= piv_elems + 1; ++pass) {
            const bool own = pass > piv_elems;
            int e, at, cnt;
            if (own) { e = piv; at = p; cnt = own_vars; }
            else { e = adj_[p++]; at = head_[e]; cnt = len_[e]; }
            for (int k = 1; k <= cnt; ++k) {
                const int v = adj_[at++], wt = nv_[v];
                if (wt <= 0) continue;
                if (tail_ >= capacity_) {
                    if (!own) {
                        head_[piv] = p;
                        len_[piv] = (piv_elems - pass) + own_vars;
                        if (len_[piv] == 0) head_[piv] = kNone;
                    }
                    head_[e] = at;
                    len_[e] = cnt - k;
                    if (len_[e] == 0) head_[e] = kNone;
                    squeeze(lo);
                    at = head_[e];
                    p = own ? at : head_[piv];
                }
                degree_sum += wt;
                nv_[v] = -wt;
                adj_[tail_++] = v;
                bucket_drop(v);
            }
            if (e != piv) { head_[e] = flipped(piv); mark_[e] = 0; }
        }
        hi = tail_ - 1;
    }

    // ApproximateMinimumDegreeOrdering.cs:402-428: |Le \ Lme| held in mark_[e] - flag_
    void outside_counts(int lo, int hi) {
        for (int q = lo; q <= hi; ++q) {
            const int v = adj_[q], ne = elen_[v];
            if (ne <= 0) continue;
            const int wt = -nv_[v], base = flag_ - wt;
            const int end = head_[v] + ne;
            for (int p = head_[v]; p < end; ++p) {
                const int e = adj_[p];
                int m = mark_[e];
                if (m >= flag_) m -= wt;
                else if (m != 0) m = deg_[e] + base;
                mark_[e] = m;
            }
        }
    }

    // ApproximateMinimumDegreeOrdering.cs:436-522
    void new_degrees(int piv, int lo, int hi, int &degree_sum, int &piv_weight, int &done) {
        hash_count_ = 0;
        for (int q = lo; q <= hi; ++q) {
            const int v = adj_[q];
            const int beg = head_[v], elem_end = beg + elen_[v];
            int out = beg, d = 0;
            int64_t key = 0;
            for (int p = beg; p < elem_end; ++p) {
                const int e = adj_[p];
                if (mark_[e] == 0) continue;
                const int ext = mark_[e] - flag_;
                if (ext > 0) { d += ext; adj_[out++] = e; key += e; }
                else if (ext == 0) { head_[e] = flipped(piv); mark_[e] = 0; }
            }
            elen_[v] = out - beg + 1;
            const int var_beg = out, list_end = beg + len_[v];
            for (int p = elem_end; p < list_end; ++p) {
                const int u = adj_[p], wu = nv_[u];
                if (wu <= 0) continue;
                d += wu;
                adj_[out++] = u;
                key += u;
            }
            if (elen_[v] == 1 && var_beg == out) {
                head_[v] = flipped(piv);
                const int wt = -nv_[v];
                degree_sum -= wt;
                piv_weight += wt;
                done += wt;
                nv_[v] = 0;
                elen_[v] = kNone;
                absorb_group(piv, v);
            } else {
                deg_[v] = std::min(deg_[v], d);
                adj_[out] = adj_[var_beg];
                adj_[var_beg] = adj_[beg];
                adj_[beg] = piv;
                len_[v] = out - beg + 1;
                hash_insert(key, v);
            }
        }
    }

    // ApproximateMinimumDegreeOrdering.cs:531-592
    void merge_twins() {
        for (int b = 0; b < hash_count_; ++b) {
            const int bucket = hash_used_[b];
            for (int v = hash_first_[bucket]; v != kNone; v = hash_next_[v]) {
                if (nv_[v] >= 0) continue;
                const int l = len_[v], ne = elen_[v];
                for (int p = head_[v] + 1; p < head_[v] + l; ++p) mark_[adj_[p]] = flag_;
                int u = hash_next_[v];
                while (u != kNone) {
                    bool same = nv_[u] < 0 && len_[u] == l && elen_[u] == ne;
                    if (same)
                        for (int p = head_[u] + 1; p < head_[u] + l; ++p)
                            if (mark_[adj_[p]] != flag_) { same = false; break; }
                    const int nxt = hash_next_[u];
                    if (same) {
                        head_[u] = flipped(v);
                        nv_[v] += nv_[u];
                        nv_[u] = 0;
                        elen_[u] = kNone;
                        absorb_group(v, u);
                        hash_remove(u);
                    }
                    u = nxt;
                }
                ++flag_;
            }
            hash_first_[bucket] = kNone;
        }
    }

    // ApproximateMinimumDegreeOrdering.cs:599-640
    void reinsert(int piv, int piv_elems, int degree_sum, int piv_weight, int done, int lo, int hi) {
        int out = lo;
        const int left = n_ - done;
        for (int q = lo; q <= hi; ++q) {
            const int v = adj_[q], wt = -nv_[v];
            if (wt <= 0) continue;
            nv_[v] = wt;
            int d = std::min(deg_[v] + degree_sum - wt, left - wt);
            if (d < 0) d = 0;
            deg_[v] = d;
            bucket_push(v, d);
            adj_[out++] = v;
        }
        nv_[piv] = piv_weight;
        len_[piv] = out - lo;
        if (len_[piv] == 0) { head_[piv] = kNone; mark_[piv] = 0; }
        if (piv_elems != 0) tail_ = out;
    }

    int n_;
    std::vector<int> adj_;
    int capacity_ = 0, tail_ = 0;
    std::vector<int> head_, len_, elen_, nv_, deg_, mark_;
    int flag_ = 2, mark_limit_ = 0;
    std::vector<int> bucket_first_, bucket_next_, bucket_prev_;
    int min_bucket_ = 0;
    std::vector<int> hash_first_, hash_next_, hash_prev_, hash_used_;
    int hash_count_ = 0;
    std::vector<int> group_last_, group_next_;
    std::vector<int> order_, dense_;
    int cursor_ = 0;
};

// ---------------------------------------------------------------------------
// PermutedSymmetricPattern.Create (PermutedSymmetricPattern.cs:81-185)
// ---------------------------------------------------------------------------
struct Pattern {
    std::vector<int> ptr, rows, lower_start, lvptr, lvidx, dvidx;
};

void build_pattern(int n, const int *rowptr, const int *colidx, const std::vector<int> &iperm, Pattern &pt) {
    pt.ptr.assign(n + 1, 0);
    pt.lvptr.assign(n + 1, 0);
    pt.dvidx.assign(n, kNone);
    for (int r = 0; r < n; ++r) {
        const int pr = iperm[r];
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int c = colidx[k];
            if (c == r) { pt.dvidx[pr] = k; continue; }
            const int pc = iperm[c];
            ++pt.ptr[pr + 1];
            ++pt.ptr[pc + 1];
            ++pt.lvptr[std::min(pr, pc) + 1];
        }
    }
    for (int j = 0; j < n; ++j) { pt.ptr[j + 1] += pt.ptr[j]; pt.lvptr[j + 1] += pt.lvptr[j]; }
    pt.rows.assign(pt.ptr[n], 0);
    const int nl = pt.lvptr[n];
    std::vector<std::pair<int, int>> lower(nl);  // (row, value index)
    std::vector<int> at(pt.ptr.begin(), pt.ptr.end() - 1), lat(pt.lvptr.begin(), pt.lvptr.end() - 1);
    for (int r = 0; r < n; ++r) {
        const int pr = iperm[r];
        for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int c = colidx[k];
            if (c == r) continue;
            const int pc = iperm[c];
            pt.rows[at[pr]++] = pc;
            pt.rows[at[pc]++] = pr;
            const int lo = std::min(pr, pc);
            lower[lat[lo]++] = {std::max(pr, pc), k};
        }
    }
    for (int j = 0; j < n; ++j) {
        std::sort(pt.rows.begin() + pt.ptr[j], pt.rows.begin() + pt.ptr[j + 1]);
        std::sort(lower.begin() + pt.lvptr[j], lower.begin() + pt.lvptr[j + 1]);
    }
    pt.lvidx.resize(nl);
    for (int q = 0; q < nl; ++q) pt.lvidx[q] = lower[q].second;
    pt.lower_start.resize(n);
    for (int j = 0; j < n; ++j) pt.lower_start[j] = pt.ptr[j + 1] - (pt.lvptr[j + 1] - pt.lvptr[j]);
}

// EliminationTree.Build (EliminationTree.cs:39-71), path compression via ancestors
std::vector<int> elimination_tree(int n, const std::vector<int> &ptr, const std::vector<int> &rows) {
    std::vector<int> parent(n, kNone), up(n, kNone);
    for (int k = 0; k < n; ++k)
        for (int p = ptr[k]; p < ptr[k + 1]; ++p) {
            int i = rows[p];
            if (i >= k) break;
            while (i != kNone && i < k) {
                const int nxt = up[i];
                up[i] = k;
                if (nxt == kNone) parent[i] = k;
                i = nxt;
            }
        }
    return parent;
}

// EliminationTree.Postorder (EliminationTree.cs:80-139): children ascending
std::vector<int> postorder(int n, const std::vector<int> &parent) {
    std::vector<int> kid(n, kNone), sib(n, kNone), stack(n), order(n);
    for (int j = n - 1; j >= 0; --j)
        if (parent[j] != kNone) { sib[j] = kid[parent[j]]; kid[parent[j]] = j; }
    int k = 0;
    for (int root = 0; root < n; ++root) {
        if (parent[root] != kNone) continue;
        int top = 0;
        stack[0] = root;
        while (top >= 0) {
            const int v = stack[top], c = kid[v];
            if (c == kNone) { --top; order[k++] = v; }
            else { kid[v] = sib[c]; stack[++top] = c; }
        }
    }
    return order;
}

// EliminationTree.ColumnCounts / FindLeaf (EliminationTree.cs:152-249)
std::vector<int> column_counts(int n, const std::vector<int> &ptr, const std::vector<int> &rows,
                               const std::vector<int> &parent, const std::vector<int> &post) {
    std::vector<int> cnt(n), anc(n), maxfirst(n, kNone), prevleaf(n, kNone), firstdesc(n, kNone);
    for (int k = 0; k < n; ++k) {
        int j = post[k];
        cnt[j] = firstdesc[j] == kNone ? 1 : 0;
        for (; j != kNone && firstdesc[j] == kNone; j = parent[j]) firstdesc[j] = k;
    }
    std::iota(anc.begin(), anc.end(), 0);
    for (int k = 0; k < n; ++k) {
        const int j = post[k];
        if (parent[j] != kNone) --cnt[parent[j]];
        for (int p = ptr[j]; p < ptr[j + 1]; ++p) {
            const int i = rows[p];
            if (i <= j || firstdesc[j] <= maxfirst[i]) continue;
            maxfirst[i] = firstdesc[j];
            const int last_leaf = prevleaf[i];
            prevleaf[i] = j;
            ++cnt[j];
            if (last_leaf == kNone) continue;
            int root = last_leaf;
            while (root != anc[root]) root = anc[root];
            for (int s = last_leaf; s != root;) { const int nx = anc[s]; anc[s] = root; s = nx; }
            --cnt[root];
        }
        if (parent[j] != kNone) anc[j] = parent[j];
    }
    for (int j = 0; j < n; ++j)
        if (parent[j] != kNone) cnt[parent[j]] += cnt[j];
    return cnt;
}

// SupernodalStructure.IsMergeWorthwhile (SupernodalStructure.cs:271-286)
bool worth_merging(int cols, int64_t zeros, int64_t stored) {
    if (zeros <= 0) return true;
    const double frac = (double)zeros / (double)stored;
    if (cols <= 4) return frac <= 0.8;
    if (cols <= 16) return frac <= 0.1;
    if (cols <= 48) return frac <= 0.05;
    return false;
}

// SupernodalStructure.Build (:102-166), FindFundamentalSupernodes (:175-193),
// Amalgamate (:203-262), BuildPatterns (:296-355), ComputeCosts (:370-418)
void build_supernodes(Symbolic &S, int limit) {
    const int n = S.n;
    const auto &par = S.etree;
    const auto &cc = S.colcount;
    std::vector<int> nkids(n, 0);
    for (int j = 0; j < n; ++j)
        if (par[j] >= 0) ++nkids[par[j]];
    std::vector<int> fund;
    for (int j = 0; j < n; ++j)
        if (j == 0 || par[j - 1] != j || nkids[j] > 1 || cc[j - 1] != cc[j] + 1) fund.push_back(j);
    fund.push_back(n);
    const int nf = (int)fund.size() - 1;
    std::vector<int64_t> pre(n + 1, 0);
    for (int j = 0; j < n; ++j) pre[j + 1] = pre[j] + cc[j];

    std::vector<int> bounds, fsize;
    for (int idx = 0; idx < nf;) {
        const int lo = fund[idx];
        int hi = fund[idx + 1] - 1;
        int cols = hi - lo + 1;
        int front = cc[lo];
        int64_t nz = pre[hi + 1] - pre[lo];
        while (idx + 1 < nf && par[hi] == hi + 1) {
            const int nlo = fund[idx + 1], nhi = fund[idx + 2] - 1;
            const int mcols = cols + (nhi - nlo + 1);
            if (mcols > limit) break;
            const int64_t mfront = (int64_t)cols + cc[nlo];
            const int64_t stored = mfront * mcols - (int64_t)mcols * (mcols - 1) / 2;
            const int64_t mnz = nz + (pre[nhi + 1] - pre[nlo]);
            if (!worth_merging(mcols, stored - mnz, stored)) break;
            hi = nhi;
            cols = mcols;
            front = (int)mfront;
            nz = mnz;
            ++idx;
        }
        bounds.push_back(lo);
        fsize.push_back(front);
        ++idx;
    }
    bounds.push_back(n);
    const int ns = (int)fsize.size();
    S.nsuper = ns;
    S.first = bounds;
    S.super_of.assign(n, 0);
    for (int s = 0; s < ns; ++s)
        for (int j = S.first[s]; j < S.first[s + 1]; ++j) S.super_of[j] = s;
    S.sparent.assign(ns, kNone);
    S.child_ptr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        const int pc = par[S.first[s + 1] - 1];
        S.sparent[s] = pc < 0 ? kNone : S.super_of[pc];
        if (S.sparent[s] >= 0) ++S.child_ptr[S.sparent[s] + 1];
    }
    for (int s = 0; s < ns; ++s) S.child_ptr[s + 1] += S.child_ptr[s];
    S.children.assign(S.child_ptr[ns], 0);
    {
        std::vector<int> at(S.child_ptr.begin(), S.child_ptr.end() - 1);
        for (int s = 0; s < ns; ++s)
            if (S.sparent[s] >= 0) S.children[at[S.sparent[s]]++] = s;
    }
    S.pat_ptr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) S.pat_ptr[s + 1] = S.pat_ptr[s] + fsize[s];
    S.pat_rows.assign(S.pat_ptr[ns], 0);

    std::vector<int> seen(n, kNone);
    for (int s = 0; s < ns; ++s) {
        int out = S.pat_ptr[s];
        const int lim = S.pat_ptr[s + 1];
        auto add = [&](int row) {
            if (seen[row] == s) return;
            seen[row] = s;
            if (out >= lim) throw Error(MG_E_INTERNAL, "Symbolic phase: predicted front size does not match actual.");
            S.pat_rows[out++] = row;
        };
        for (int j = S.first[s]; j < S.first[s + 1]; ++j) add(j);
        for (int j = S.first[s]; j < S.first[s + 1]; ++j)
            for (int p = S.lower_start[j]; p < S.pptr[j + 1]; ++p) add(S.prows[p]);
        for (int c = S.child_ptr[s]; c < S.child_ptr[s + 1]; ++c) {
            const int ch = S.children[c];
            for (int q = S.pat_ptr[ch] + S.width(ch); q < S.pat_ptr[ch + 1]; ++q) add(S.pat_rows[q]);
        }
        if (out != lim) throw Error(MG_E_INTERNAL, "Symbolic phase: predicted front size does not match actual.");
        std::sort(S.pat_rows.begin() + S.pat_ptr[s], S.pat_rows.begin() + lim);
    }

    int64_t entries = 0, ops = 0, live = 0, peak = 0;
    int maxf = 0, maxw = 0;
    std::vector<int64_t> held(ns, 0);
    for (int s = 0; s < ns; ++s) {
        const int64_t f = S.front(s), w = S.width(s);
        const int64_t blk = f * w - w * (w - 1) / 2;
        if (blk > 2147483647LL)
            throw Error(MG_E_TOO_DENSE, "Factor block of a supernode cannot be represented by a single array. The problem is too dense for the direct solver.");
        entries += blk;
        for (int64_t k = 0; k < w; ++k) ops += (f - k) * (f - k);
        maxf = std::max(maxf, (int)f);
        maxw = std::max(maxw, (int)w);
        peak = std::max(peak, live + f * (f + 1) / 2);
        for (int c = S.child_ptr[s]; c < S.child_ptr[s + 1]; ++c) live -= held[S.children[c]];
        const int64_t t = f - w;
        held[s] = t * (t + 1) / 2;
        live += held[s];
        peak = std::max(peak, live);
    }
    S.factor_entries = entries;
    S.factor_ops = ops;
    S.max_front = maxf;
    S.max_width = maxw;
    S.peak_contribution = peak;
}

}  // namespace

void amd_order(int n, const int *rowptr, const int *colidx, std::vector<int> &perm) {
    perm.clear();
    if (n == 0) return;
    MinimumDegree md(n, rowptr, colidx);
    md.run(perm);
}

// UtduSymbolicFactorization.Analyze (UtduSymbolicFactorization.cs:240-295)
void analyze(int n, const int *rowptr, const int *colidx, const mg_options &opts, Symbolic &S) {
    S.n = n;
    S.matrix_nnz = rowptr[n];
    S.opts = opts;
    S.a_rowptr.assign(rowptr, rowptr + n + 1);
    S.a_colidx.assign(colidx, colidx + rowptr[n]);
    if (opts.ordering == MG_ORDER_AMD) {
        amd_order(n, rowptr, colidx, S.fill_order);
    } else {
        S.fill_order.resize(n);
        std::iota(S.fill_order.begin(), S.fill_order.end(), 0);
    }
    std::vector<int> inv(n);
    for (int k = 0; k < n; ++k) inv[S.fill_order[k]] = k;
    {
        Pattern first_pass;
        build_pattern(n, rowptr, colidx, inv, first_pass);
        const auto parent = elimination_tree(n, first_pass.ptr, first_pass.rows);
        const auto post = postorder(n, parent);
        S.perm.resize(n);
        for (int k = 0; k < n; ++k) S.perm[k] = S.fill_order[post[k]];
    }
    S.iperm.resize(n);
    for (int k = 0; k < n; ++k) S.iperm[S.perm[k]] = k;
    Pattern pt;
    build_pattern(n, rowptr, colidx, S.iperm, pt);
    S.pptr = std::move(pt.ptr);
    S.prows = std::move(pt.rows);
    S.lower_start = std::move(pt.lower_start);
    S.lvptr = std::move(pt.lvptr);
    S.lvidx = std::move(pt.lvidx);
    S.dvidx = std::move(pt.dvidx);
    S.etree = elimination_tree(n, S.pptr, S.prows);
    const auto post2 = postorder(n, S.etree);
    S.colcount = column_counts(n, S.pptr, S.prows, S.etree, post2);
    build_supernodes(S, std::max(1, opts.supernode_amalgamation_limit));
    int64_t strict = 0;
    for (int j = 0; j < n; ++j) strict += S.colcount[j];
    S.strict_nonzeros = strict;
}

}  // namespace mg

// ---- batched systems ---------------------------------------------------------------

namespace mg {

void analyze_batch(int count, const int *sizes, const int *const *rowptrs, const int *const *colidxs,
                   const mg_options &opts, std::vector<Symbolic> &each) {
    each.clear();
    each.resize(count);
    int threads = opts.max_degree_of_parallelism > 0 ? opts.max_degree_of_parallelism
                                                      : (int)std::thread::hardware_concurrency();
    threads = std::max(1, std::min(threads, count));
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> errors(threads);
    auto work = [&](int tid) {
        try {
            for (int t = next++; t < count; t = next++) analyze(sizes[t], rowptrs[t], colidxs[t], opts, each[t]);
        } catch (...) {
            errors[tid] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(work, k);
    work(0);
    for (auto &th : pool) th.join();
    for (auto &e : errors)
        if (e) std::rethrow_exception(e);
}

void merge_symbolic(const std::vector<Symbolic> &each, Symbolic &S) {
    S = Symbolic();
    const int count = (int)each.size();
    int64_t n = 0, nnz = 0, ns = 0;
    for (const auto &t : each) {
        n += t.n;
        nnz += t.matrix_nnz;
        ns += t.nsuper;
    }
    if (n > INT32_MAX || nnz > INT32_MAX) throw Error(MG_E_TOO_DENSE, "The batch exceeds 2^31 rows or entries.");
    S.n = (int)n;
    S.matrix_nnz = (int)nnz;
    S.nsuper = (int)ns;
    if (count) S.opts = each[0].opts;
    S.sys_begin.assign(count + 1, 0);
    S.a_rowptr.assign(1, 0);
    S.pptr.assign(1, 0);
    S.lvptr.assign(1, 0);
    S.first.assign(1, 0);
    S.child_ptr.assign(1, 0);
    S.pat_ptr.assign(1, 0);
    const auto shift = [](std::vector<int> &dst, const std::vector<int> &src, size_t from, int64_t by, bool keep_neg) {
        for (size_t k = from; k < src.size(); ++k) dst.push_back(keep_neg && src[k] < 0 ? src[k] : (int)(src[k] + by));
    };
    int64_t r0 = 0, v0 = 0, s0 = 0;
    for (int t = 0; t < count; ++t) {
        const Symbolic &T = each[t];
        S.sys_begin[t] = (int)r0;
        const int64_t p0 = S.pptr.back(), l0 = S.lvptr.back(), c0 = S.child_ptr.back(), q0 = S.pat_ptr.back();
        shift(S.a_rowptr, T.a_rowptr, 1, v0, false);
        shift(S.a_colidx, T.a_colidx, 0, r0, false);
        shift(S.fill_order, T.fill_order, 0, r0, false);
        shift(S.perm, T.perm, 0, r0, false);
        shift(S.iperm, T.iperm, 0, r0, false);
        shift(S.etree, T.etree, 0, r0, true);
        S.colcount.insert(S.colcount.end(), T.colcount.begin(), T.colcount.end());
        shift(S.pptr, T.pptr, 1, p0, false);
        shift(S.prows, T.prows, 0, r0, false);
        shift(S.lower_start, T.lower_start, 0, p0, false);
        shift(S.lvptr, T.lvptr, 1, l0, false);
        shift(S.lvidx, T.lvidx, 0, v0, false);
        shift(S.dvidx, T.dvidx, 0, v0, true);
        shift(S.first, T.first, 1, r0, false);
        shift(S.super_of, T.super_of, 0, s0, false);
        shift(S.sparent, T.sparent, 0, s0, true);
        shift(S.child_ptr, T.child_ptr, 1, c0, false);
        shift(S.children, T.children, 0, s0, false);
        shift(S.pat_ptr, T.pat_ptr, 1, q0, false);
        shift(S.pat_rows, T.pat_rows, 0, r0, false);
        S.factor_entries += T.factor_entries;
        S.strict_nonzeros += T.strict_nonzeros;
        S.factor_ops += T.factor_ops;
        S.peak_contribution = std::max(S.peak_contribution, T.peak_contribution);
        S.max_front = std::max(S.max_front, T.max_front);
        S.max_width = std::max(S.max_width, T.max_width);
        r0 += T.n;
        v0 += T.matrix_nnz;
        s0 += T.nsuper;
    }
    S.sys_begin[count] = (int)r0;
}

}  // namespace mg
