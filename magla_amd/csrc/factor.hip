// factor.hip -- level-scheduled multifrontal U^T D U factorization and the
// supernodal solves on MI355X (gfx950).
//
// Device layout
//   * Every supernode s owns a dense f_s x f_s column-major front (ld = f_s)
//     in one arena (sum f^2 elements of the factor precision).  After the
//     partial factorization, columns [0, w_s) hold the factor block (unit
//     diagonal, L below it: MAGLA's factor column k = [1, L[k+1.., k]],
//     UtduNumericFactorization.cs:517-519 / 591) and the trailing
//     t_s x t_s block is the contribution read by the parent (:521-530).
//     Nothing is copied between fronts except the extend-add itself.
//   * The supernodal tree is cut into levels by height (leaves = 0); all
//     fronts of a level are independent and go through one launch per step.
//   * Extend-add is a deterministic PULL: each parent entry sums the entries
//     of its children in MAGLA's Children order (:470-509) through
//     per-child inverse row maps, so no atomics and no order dependence.
//
// Kernels per level: assemble (pull + zero-fill), scatter of the original
// entries, then per 64-column panel of the supernode: panel LDL^T
// (diagonal block in LDS + row-block triangular solve), and the Schur update
// F22 -= L21 D L21^T on the lower triangle (MFMA tiles).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device.h"
#include "mfma.h"

namespace mg {

constexpr int kTile = 64;   // tile edge of assembly / Schur tiles
constexpr int kPanel = 64;  // panel width: columns factored by one panel_diag / panel_rows pass
constexpr int kBlock = 256; // Schur blocking: panels of one block update the rest of the front at once
constexpr int kMaxFrontWidth = 4096;  // widest fused front (LDS-resident solve vectors)

// A device front is a chain of consecutive supernodes s, s+1, ..., each the
// elimination-tree parent of the previous one (the last column of s has the
// first column of s+1 as etree parent), fused when the rows nest: the fused
// front has the chain's columns followed by the top supernode's footer, so
// the lower members' columns carry at most a few structural zeros.  The
// elimination order, the supernode partition and every factor entry are
// those of the reference; only the numeric schedule differs (no extend-add
// between chain members, Schur updates of up to kBlock columns).
struct DevicePlan {
    int n = 0, nsuper = 0, nfront = 0, nlevels = 0, max_w = 0, max_f = 0;
    int64_t arena = 0;  // sum f^2 over fronts
    int64_t usum = 0;   // sum (f - w) over fronts
    std::vector<int> h_f, h_w, h_first;
    std::vector<int64_t> h_off;
    std::vector<int> front_of;   // supernode -> front
    std::vector<int> h_pat_ptr, h_pat_rows;
    std::vector<int> lvl_node_ptr, lvl_pull_ptr;
    std::vector<int64_t> lvl_asm_ptr;
    std::vector<int> lvl_npanels;
    std::vector<std::vector<int>> lvl_panel_ptr, lvl_upd_ptr, lvl_diag_ptr;
    std::vector<int> lvl_init_ptr;  // solve: (front, 256-row chunk) tasks of the forward gather
    int max_panel_tasks = 0;        // largest panel-task launch (solve partial-sum buffer)
    int64_t vsum = 0;               // sum f over fronts (solve front vectors)
    // algorithmic work per launch
    std::vector<double> lvl_pull_elems, lvl_pull_index;
    std::vector<std::vector<double>> lvl_upd_flops, lvl_upd_elems;

    DevArray<int> f, w, first, child_ptr, children, pat_ptr, pat_rows, level_nodes, inv_map, perm, diag_nodes, diag_tb;
    DevArray<int64_t> off, inv_off, u_off, asm_pos, vbase;
    DevArray<int2> init_tasks;
    DevArray<int> asm_vidx;
    DevArray<int4> pull_tasks, upd_tasks;
    DevArray<int2> panel_tasks;
};

int plan_level_count(const DevicePlan &P) { return P.nlevels; }

// Schur work after panel q of a front of width w and size f: "inner" updates the
// remaining columns of the current kBlock block with this panel alone, "outer"
// (after the block's last panel) updates the trailing [be, f)^2 with the whole block.
struct SchurStep {
    bool any, inner;
    int ps, pe, bs, be, base, kb, ke, cend, row_tiles, col_tiles;
};
static inline SchurStep schur_step(int f, int w, int q) {
    SchurStep st{};
    st.ps = q * kPanel;
    st.pe = std::min(st.ps + kPanel, w);
    st.bs = (st.ps / kBlock) * kBlock;
    st.be = std::min(st.bs + kBlock, w);
    st.inner = st.pe < st.be;
    if (st.inner) {
        st.base = st.pe; st.kb = st.ps; st.ke = st.pe; st.cend = st.be;
    } else {
        st.base = st.be; st.kb = st.bs; st.ke = st.be; st.cend = f;
    }
    st.any = st.base < f;
    st.row_tiles = (f - st.base + kTile - 1) / kTile;
    st.col_tiles = (st.cend - st.base + kTile - 1) / kTile;
    return st;
}

static std::shared_ptr<DevicePlan> build_plan_host(const Symbolic &S, bool upload) {
    auto P = std::make_shared<DevicePlan>();
    const int n = S.n, ns = S.nsuper;
    P->n = n;
    P->nsuper = ns;

    // ---- fuse chains into fronts ----
    std::vector<int> fbeg;  // first member supernode of each front
    P->front_of.assign(ns, -1);
    for (int s = 0; s < ns;) {
        int top = s, width = S.width(s);
        while (top + 1 < ns && S.sparent[top] == top + 1) {
            const int p = top + 1;
            const int last = S.first[top + 1] - 1;
            if (S.etree[last] != S.first[p]) break;
            if (width + S.width(p) > kMaxFrontWidth) break;
            // rows of the fused front: columns so far + pattern(p); the members' true
            // rows are a subset.  Accept when the structural zeros stay below 10%.
            const int fused = (S.first[p] - S.first[s]) + S.front(p);
            const int own = S.front(top) + (S.first[top] - S.first[s]);
            if (fused > own + own / 10) break;
            top = p;
            width += S.width(p);
        }
        const int g = (int)fbeg.size();
        for (int m = s; m <= top; ++m) P->front_of[m] = g;
        fbeg.push_back(s);
        s = top + 1;
    }
    const int nf = (int)fbeg.size();
    fbeg.push_back(ns);
    P->nfront = nf;
    P->h_f.resize(nf);
    P->h_w.resize(nf);
    P->h_first.resize(nf + 1);
    P->h_pat_ptr.assign(nf + 1, 0);
    for (int g = 0; g < nf; ++g) {
        const int s0 = fbeg[g], top = fbeg[g + 1] - 1;
        const int w = S.first[top + 1] - S.first[s0];
        const int t = S.front(top) - S.width(top);
        P->h_w[g] = w;
        P->h_f[g] = w + t;
        P->h_first[g] = S.first[s0];
        P->h_pat_ptr[g + 1] = P->h_pat_ptr[g] + w + t;
    }
    P->h_first[nf] = n;
    P->h_pat_rows.resize(P->h_pat_ptr[nf]);
    for (int g = 0; g < nf; ++g) {
        const int top = fbeg[g + 1] - 1;
        int *rows = P->h_pat_rows.data() + P->h_pat_ptr[g];
        for (int k = 0; k < P->h_w[g]; ++k) rows[k] = P->h_first[g] + k;
        const int wt = S.width(top), ft = S.front(top);
        for (int a = wt; a < ft; ++a) rows[P->h_w[g] + a - wt] = S.pat_rows[S.pat_ptr[top] + a];
    }
    // front tree: parent front = front of the top member's parent supernode; children in
    // ascending order of the members' children lists (MAGLA's Children order per member)
    std::vector<int> fparent(nf, -1), fchild_ptr(nf + 1, 0), fchildren;
    for (int g = 0; g < nf; ++g) {
        const int top = fbeg[g + 1] - 1;
        fparent[g] = S.sparent[top] >= 0 ? P->front_of[S.sparent[top]] : -1;
    }
    for (int g = 0; g < nf; ++g) {
        for (int m = fbeg[g]; m < fbeg[g + 1]; ++m)
            for (int c = S.child_ptr[m]; c < S.child_ptr[m + 1]; ++c) {
                const int ch = S.children[c];
                if (P->front_of[ch] == g) continue;  // chain member
                fchildren.push_back(P->front_of[ch]);
            }
        fchild_ptr[g + 1] = (int)fchildren.size();
    }
    P->h_off.assign(nf + 1, 0);
    std::vector<int64_t> uoff(nf + 1, 0);
    for (int g = 0; g < nf; ++g) {
        const int f = P->h_f[g], w = P->h_w[g];
        P->max_w = std::max(P->max_w, w);
        P->max_f = std::max(P->max_f, f);
        P->h_off[g + 1] = P->h_off[g] + (int64_t)f * f;
        uoff[g + 1] = uoff[g] + (f - w);
    }
    P->arena = P->h_off[nf];
    P->usum = uoff[nf];

    // levels by height in the front tree
    std::vector<int> level(nf, 0);
    int nl = 0;
    for (int g = 0; g < nf; ++g) {
        int h = 0;
        for (int c = fchild_ptr[g]; c < fchild_ptr[g + 1]; ++c) h = std::max(h, level[fchildren[c]] + 1);
        level[g] = h;
        nl = std::max(nl, h + 1);
    }
    P->nlevels = nl;
    std::vector<std::vector<int>> bylevel(nl);
    for (int g = 0; g < nf; ++g) bylevel[level[g]].push_back(g);

    // inverse row maps: for child front x of front g, inv_x[i] = index of g's row i in x's footer
    std::vector<int64_t> inv_off(nf + 1, 0);
    for (int g = 0; g < nf; ++g) inv_off[g + 1] = inv_off[g] + (fparent[g] >= 0 ? P->h_f[fparent[g]] : 0);
    std::vector<int> inv_map(inv_off[nf], -1);
    std::vector<int> rowmap(std::max(n, 1), -1);
    for (int g = 0; g < nf; ++g) {
        const int fg = P->h_f[g];
        for (int t = 0; t < fg; ++t) rowmap[P->h_pat_rows[P->h_pat_ptr[g] + t]] = t;
        for (int c = fchild_ptr[g]; c < fchild_ptr[g + 1]; ++c) {
            const int x = fchildren[c];
            const int wx = P->h_w[x], tx = P->h_f[x] - wx;
            int *inv = inv_map.data() + inv_off[x];
            for (int a = 0; a < tx; ++a) inv[rowmap[P->h_pat_rows[P->h_pat_ptr[x] + wx + a]]] = a;
        }
    }

    std::vector<int> level_nodes, diag_nodes, diag_tb;
    std::vector<int4> pull, upd;
    std::vector<int2> panel, init;
    std::vector<int64_t> vbase(nf + 1, 0);
    for (int g = 0; g < nf; ++g) vbase[g + 1] = vbase[g] + P->h_f[g];
    P->vsum = vbase[nf];
    P->lvl_init_ptr.push_back(0);
    std::vector<int64_t> asm_pos;
    std::vector<int> asm_vidx;
    P->lvl_node_ptr.push_back(0);
    P->lvl_pull_ptr.push_back(0);
    P->lvl_asm_ptr.push_back(0);
    for (int L = 0; L < nl; ++L) {
        int npan = 0;
        double pe_elems = 0, pe_index = 0;
        for (int g : bylevel[L]) {
            level_nodes.push_back(g);
            const int f = P->h_f[g], w = P->h_w[g];
            npan = std::max(npan, (w + kPanel - 1) / kPanel);
            for (int c = 0; c < (f + 255) / 256; ++c) init.push_back(make_int2(g, c));
            const int T = (f + kTile - 1) / kTile;
            for (int tj = 0; tj < T; ++tj)
                for (int ti = tj; ti < T; ++ti) pull.push_back(make_int4(g, ti, tj, 0));
            pe_elems += (double)f * (f + 1) / 2;
            for (int c = fchild_ptr[g]; c < fchild_ptr[g + 1]; ++c) {
                const double tc = P->h_f[fchildren[c]] - P->h_w[fchildren[c]];
                pe_elems += tc * (tc + 1) / 2;
                pe_index += f;
            }
            // original entries of the own columns (UtduNumericFactorization.cs:450-468)
            for (int t = 0; t < f; ++t) rowmap[P->h_pat_rows[P->h_pat_ptr[g] + t]] = t;
            const int64_t base = P->h_off[g];
            for (int jl = 0; jl < w; ++jl) {
                const int j = P->h_first[g] + jl;
                if (S.dvidx[j] >= 0) {
                    asm_pos.push_back(base + (int64_t)jl * f + jl);
                    asm_vidx.push_back(S.dvidx[j]);
                }
                for (int p = S.lower_start[j]; p < S.pptr[j + 1]; ++p) {
                    asm_pos.push_back(base + (int64_t)jl * f + rowmap[S.prows[p]]);
                    asm_vidx.push_back(S.lvidx[S.lvptr[j] + (p - S.lower_start[j])]);
                }
            }
        }
        P->lvl_node_ptr.push_back((int)level_nodes.size());
        P->lvl_init_ptr.push_back((int)init.size());
        P->lvl_pull_ptr.push_back((int)pull.size());
        P->lvl_asm_ptr.push_back((int64_t)asm_pos.size());
        P->lvl_npanels.push_back(npan);
        P->lvl_pull_elems.push_back(pe_elems);
        P->lvl_pull_index.push_back(pe_index);
        std::vector<int> pp{(int)panel.size()}, up{(int)upd.size()}, dp{(int)diag_nodes.size()};
        std::vector<double> ufl(npan, 0.0), uel(npan, 0.0);
        for (int q = 0; q < npan; ++q) {
            const int pbase = (int)panel.size();
            for (int g : bylevel[L]) {
                const int f = P->h_f[g], w = P->h_w[g];
                if (w <= q * kPanel) continue;
                diag_nodes.push_back(g);
                diag_tb.push_back((int)panel.size() - pbase);
                const int pe = std::min(q * kPanel + kPanel, w);
                const int nrb = (f - pe + kTile - 1) / kTile;
                for (int rb = 0; rb < nrb; ++rb) panel.push_back(make_int2(g, rb));
            }
            P->max_panel_tasks = std::max(P->max_panel_tasks, (int)panel.size() - pbase);
            dp.push_back((int)diag_nodes.size());
            pp.push_back((int)panel.size());
            for (int g : bylevel[L]) {
                const int f = P->h_f[g], w = P->h_w[g];
                if (w <= q * kPanel) continue;
                const SchurStep st = schur_step(f, w, q);
                if (!st.any) continue;
                for (int tj = 0; tj < st.col_tiles; ++tj)
                    for (int ti = tj; ti < st.row_tiles; ++ti) upd.push_back(make_int4(g, ti, tj, 0));
                double elems = 0;
                for (int j = st.base; j < st.cend; ++j) elems += f - j;
                ufl[q] += 2.0 * (st.ke - st.kb) * elems;
                uel[q] += 2.0 * elems + (double)(st.ke - st.kb) * (f - st.base);
            }
            up.push_back((int)upd.size());
        }
        P->lvl_panel_ptr.push_back(pp);
        P->lvl_upd_ptr.push_back(up);
        P->lvl_diag_ptr.push_back(dp);
        P->lvl_upd_flops.push_back(ufl);
        P->lvl_upd_elems.push_back(uel);
    }

    if (!upload) return P;
    P->f.upload(P->h_f);
    P->w.upload(P->h_w);
    P->first.upload(P->h_first);
    P->child_ptr.upload(fchild_ptr);
    P->children.upload(fchildren);
    P->pat_ptr.upload(P->h_pat_ptr);
    P->pat_rows.upload(P->h_pat_rows);
    P->level_nodes.upload(level_nodes);
    P->diag_nodes.upload(diag_nodes);
    P->diag_tb.upload(diag_tb);
    P->init_tasks.upload(init);
    P->vbase.upload(vbase);
    P->inv_map.upload(inv_map);
    P->perm.upload(S.perm);
    P->off.upload(P->h_off);
    P->inv_off.upload(inv_off);
    P->u_off.upload(uoff);
    P->asm_pos.upload(asm_pos);
    P->asm_vidx.upload(asm_vidx);
    P->pull_tasks.upload(pull);
    P->upd_tasks.upload(upd);
    P->panel_tasks.upload(panel);
    return P;
}

std::shared_ptr<DevicePlan> build_plan(const Symbolic &S) { return build_plan_host(S, true); }

// Host-only summary of the device schedule: fronts, levels, arena elements, largest front,
// widest front, kernel launches per factorization.
void plan_stats(const Symbolic &S, int64_t *out) {
    auto P = build_plan_host(S, false);
    int64_t launches = 0;
    for (int L = 0; L < P->nlevels; ++L) {
        launches += 2;
        for (int q = 0; q < P->lvl_npanels[L]; ++q) launches += 3;
    }
    out[0] = P->nfront;
    out[1] = P->nlevels;
    out[2] = P->arena;
    out[3] = P->max_f;
    out[4] = P->max_w;
    out[5] = launches;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Front zero-fill + extend-add as a pull over 64x64 lower tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_assemble_pull(const int4 *__restrict__ tasks, T *__restrict__ arena,
                                                       const int *__restrict__ F, const int *__restrict__ W,
                                                       const int64_t *__restrict__ off,
                                                       const int *__restrict__ child_ptr,
                                                       const int *__restrict__ children,
                                                       const int64_t *__restrict__ inv_off,
                                                       const int *__restrict__ inv_map) {
    const int4 tk = tasks[blockIdx.x];
    const int s = tk.x;
    const int f = F[s];
    const int i = tk.y * kTile + (threadIdx.x & 63);
    const int jbase = tk.z * kTile + (threadIdx.x >> 6);
    T *P = arena + off[s];
    const int c0 = child_ptr[s], c1 = child_ptr[s + 1];
    if (i >= f) return;
    for (int q = 0; q < kTile / 4; ++q) {
        const int j = jbase + 4 * q;
        if (j > i) break;
        T acc = T(0);
        for (int c = c0; c < c1; ++c) {
            const int ch = children[c];
            const int *inv = inv_map + inv_off[ch];
            const int a = inv[j];
            if (a < 0) continue;
            const int b = inv[i];
            if (b < 0) continue;
            const int fc = F[ch], wc = W[ch];
            acc += arena[off[ch] + (int64_t)(wc + a) * fc + (wc + b)];
        }
        P[(int64_t)j * f + i] = acc;
    }
}

// Original matrix entries of the own columns (positions are unique).
template <typename T>
__global__ __launch_bounds__(256) void k_assemble_values(const int64_t *__restrict__ pos, const int *__restrict__ vidx,
                                                         int64_t count, const double *__restrict__ values,
                                                         T *__restrict__ arena) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    arena[pos[e]] += (T)values[vidx[e]];
}

// Panel LDL^T, step 1: one workgroup per front factors the (<=64)^2 diagonal
// block of the panel in LDS and writes it back (unit diagonal, L below) with
// D.  Pivot rule of UtduNumericFactorization.cs:553-577: accept only
// pivot > tol*scale (NaN fails), otherwise regularize to reg*scale and count
// (negative counted separately; bit 2 marks a refused pivot when
// regularization is off); the comparison is made in float64 for both
// precisions.  Column scaling multiplies by the reciprocal as the reference
// does (:577).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_diag(const int *__restrict__ nodes, T *__restrict__ arena,
                                                    const int *__restrict__ F, const int *__restrict__ W,
                                                    const int64_t *__restrict__ off, const int *__restrict__ firstcol,
                                                    int ps, const double *__restrict__ scale, double tol, double regf,
                                                    int allow_reg, double *__restrict__ D,
                                                    unsigned char *__restrict__ flags) {
    __shared__ T L11[kPanel][kPanel + 1];
    __shared__ T dv[kPanel], dinv[kPanel];
    const int s = nodes[blockIdx.x];
    const int f = F[s], w = W[s];
    if (w <= ps) return;
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    T *M = arena + off[s];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r >= c) L11[r][c] = M[(int64_t)(ps + c) * f + ps + r];
    }
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        if (tid == 0) {
            const double thr = tol * scale[0], regv = regf * scale[0];
            double piv = (double)L11[k][k];
            unsigned char fl = 0;
            if (!(piv > thr)) {
                if (!allow_reg) fl |= 4;
                if (piv < 0.0) fl |= 2;
                fl |= 1;
                piv = regv;
            }
            const T pv = (T)piv;
            dv[k] = pv;
            dinv[k] = T(1) / pv;
            D[firstcol[s] + ps + k] = (double)pv;
            flags[firstcol[s] + ps + k] = fl;
        }
        __syncthreads();
        for (int r = k + 1 + tid; r < nb; r += 256) L11[r][k] *= dinv[k];
        __syncthreads();
        const int m = nb - k - 1;
        for (int idx = tid; idx < m * m; idx += 256) {
            const int r = k + 1 + idx % m, c = k + 1 + idx / m;
            if (r >= c) L11[r][c] -= L11[r][k] * (dv[k] * L11[c][k]);
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r > c) M[(int64_t)(ps + c) * f + ps + r] = L11[r][c];
        else if (r == c) M[(int64_t)(ps + c) * f + ps + r] = T(1);
    }
}

// Panel LDL^T, step 2: each workgroup solves one 64-row block below the
// factored diagonal block, X := X L11^{-T} D^{-1} (right-looking over the
// panel columns, as the reference's in-panel loop :580-588).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_rows(const int2 *__restrict__ tasks, T *__restrict__ arena,
                                                    const int *__restrict__ F, const int *__restrict__ W,
                                                    const int64_t *__restrict__ off, const int *__restrict__ firstcol,
                                                    int ps, const double *__restrict__ D) {
    __shared__ T L11[kPanel][kPanel + 1];
    __shared__ T X[kTile][kPanel + 1];
    __shared__ T dv[kPanel], dinv[kPanel];
    const int2 tk = tasks[blockIdx.x];
    const int s = tk.x, rb = tk.y;
    const int f = F[s], w = W[s];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const int r0 = pe + rb * kTile;
    const int nr = min(kTile, f - r0);
    if (nr <= 0) return;
    const T *M0 = arena + off[s];
    T *M = arena + off[s];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r > c) L11[r][c] = M0[(int64_t)(ps + c) * f + ps + r];
    }
    for (int k = tid; k < nb; k += 256) {
        const T pv = (T)D[firstcol[s] + ps + k];
        dv[k] = pv;
        dinv[k] = T(1) / pv;
    }
    for (int idx = tid; idx < nr * nb; idx += 256) {
        const int r = idx % nr, c = idx / nr;
        X[r][c] = M0[(int64_t)(ps + c) * f + r0 + r];
    }
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        for (int r = tid; r < nr; r += 256) X[r][k] *= dinv[k];
        __syncthreads();
        const int m = nb - k - 1;
        for (int idx = tid; idx < nr * m; idx += 256) {
            const int r = idx % nr, c = k + 1 + idx / nr;
            X[r][c] -= X[r][k] * (dv[k] * L11[c][k]);
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nr * nb; idx += 256) {
        const int r = idx % nr, c = idx / nr;
        M[(int64_t)(ps + c) * f + r0 + r] = X[r][c];
    }
}

// Schur update after panel q (cf. ApplyPanelUpdate, UtduNumericFactorization.cs:624-852):
// F[i,j] -= sum_k (L[i,k] d_k) L[j,k] over the step's K columns, on the lower part
// of the step's target region (schur_step).  One 64x64 output tile per workgroup;
// K streamed through LDS in chunks of 64, products on MFMA (mfma.h).
template <typename T>
__global__ __launch_bounds__(256) void k_schur(const int4 *__restrict__ tasks, T *__restrict__ arena,
                                               const int *__restrict__ F, const int *__restrict__ W,
                                               const int64_t *__restrict__ off, const int *__restrict__ firstcol,
                                               int q, const double *__restrict__ D) {
    __shared__ T As[kPanel][kTile + 1];  // As[k][r] = L[i0+r, kc+k] * d_k
    __shared__ T Bs[kPanel][kTile + 1];  // Bs[k][r] = L[j0+r, kc+k]
    const int4 tk = tasks[blockIdx.x];
    const int s = tk.x;
    const int f = F[s], w = W[s];
    const int ps = q * kPanel;
    const int pe = min(ps + kPanel, w), bs = (ps / kBlock) * kBlock, be = min(bs + kBlock, w);
    const bool inner = pe < be;
    const int base = inner ? pe : be, kb = inner ? ps : bs, ke = inner ? pe : be, cend = inner ? be : f;
    const int i0 = base + tk.y * kTile, j0 = base + tk.z * kTile;
    T *M = arena + off[s];
    const int tid = threadIdx.x;
    const double *Dp = D + firstcol[s];
    typename MfmaAcc<T>::type acc[2][2];
    schur_acc_zero<T>(acc);
    for (int kc = kb; kc < ke; kc += kPanel) {
        const int nk = min(kPanel, ke - kc);
        for (int idx = tid; idx < kPanel * kTile; idx += 256) {
            const int r = idx % kTile, k = idx / kTile;
            T av = T(0), bv = T(0);
            if (k < nk) {
                const T *col = M + (int64_t)(kc + k) * f;
                if (i0 + r < f) av = col[i0 + r] * (T)Dp[kc + k];
                if (j0 + r < f) bv = col[j0 + r];
            }
            As[k][r] = av;
            Bs[k][r] = bv;
        }
        __syncthreads();
        schur_acc_tile<T>(As, Bs, (nk + 3) & ~3, acc);
        __syncthreads();
    }
    schur_store_tile<T>(acc, M, f, i0, j0, cend);
}

// Deterministic statistics over D and the pivot flags (one workgroup).
__global__ __launch_bounds__(1024) void k_pivot_stats(const double *__restrict__ D, const unsigned char *__restrict__ flags,
                                                      int n, int *__restrict__ counts, double *__restrict__ mags) {
    __shared__ int sc[3][1024];
    __shared__ double smin[1024], smax[1024];
    int reg = 0, neg = 0, bad = 0;
    double lo = 1.7976931348623157e308, hi = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const unsigned char fl = flags[i];
        reg += fl & 1;
        neg += (fl >> 1) & 1;
        bad += (fl >> 2) & 1;
        const double m = fabs(D[i]);
        lo = fmin(lo, m);
        hi = fmax(hi, m);
    }
    sc[0][threadIdx.x] = reg;
    sc[1][threadIdx.x] = neg;
    sc[2][threadIdx.x] = bad;
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            for (int q = 0; q < 3; ++q) sc[q][threadIdx.x] += sc[q][threadIdx.x + h];
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + h]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[0] = sc[0][0];
        counts[1] = sc[1][0];
        counts[2] = sc[2][0];
        mags[0] = smin[0];
        mags[1] = smax[0];
    }
}

// ---- solves ----------------------------------------------------------------

// Y[i] = (T) B[perm[i]] (elimination order), nrhs columns
template <typename T>
__global__ void k_permute_in(const double *__restrict__ B, int ldb, const int *__restrict__ perm, int n, int nrhs,
                             T *__restrict__ Y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    const int r = (int)(e / n), i = (int)(e % n);
    Y[e] = (T)B[(int64_t)r * ldb + perm[i]];
}

// X[perm[i]] (+)= Xe[i]
template <typename T>
__global__ void k_permute_out(const T *__restrict__ Xe, const int *__restrict__ perm, int n, int nrhs,
                              double *__restrict__ X, int ldx, int accumulate) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    const int r = (int)(e / n), i = (int)(e % n);
    double *dst = X + (int64_t)r * ldx + perm[i];
    if (accumulate) *dst += (double)Xe[e];
    else *dst = (double)Xe[e];
}

// Supernodal solves, level by level over the fronts, in 64-column panels
// (cf. ForwardSubstitution / BackwardSubstitution, UtduNumericFactorization.cs:139-207).
// Every front owns a vector V (f rows x nrhs) in a workspace: the forward pass
// gathers it (own rows from the permuted right-hand side, plus the children's
// footer vectors through the inverse row maps), then per panel solves the unit
// lower triangle (k_fwd_diag) and subtracts the panel's product from the rows
// below (k_fwd_update); the footer left in V is the update the parent pulls.
// The backward pass walks the panels from the last one down: partial products
// of the rows below with the final x (k_bwd_partial, one 64-row block each) are
// summed in a fixed order by k_bwd_diag, which divides by D and solves the
// transposed unit triangle.  Y holds y after the forward pass and x after the
// backward pass (elimination order, column-major n x nrhs).
constexpr int kRhsChunk = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_init(const int2 *__restrict__ tasks, const int *__restrict__ F,
                                                  const int *__restrict__ W, const int *__restrict__ firstcol,
                                                  const int *__restrict__ child_ptr, const int *__restrict__ children,
                                                  const int64_t *__restrict__ inv_off, const int *__restrict__ inv_map,
                                                  const int64_t *__restrict__ vbase, T *__restrict__ V,
                                                  const T *__restrict__ Y, int n, int nrhs) {
    const int2 tk = tasks[blockIdx.x];
    const int g = tk.x, f = F[g], w = W[g];
    const int i = tk.y * 256 + threadIdx.x;
    if (i >= f) return;
    const int c0 = child_ptr[g], c1 = child_ptr[g + 1];
    T *Vg = V + vbase[g] * nrhs;
    for (int r = 0; r < nrhs; ++r) {
        T acc = i < w ? Y[(int64_t)r * n + firstcol[g] + i] : T(0);
        for (int c = c0; c < c1; ++c) {
            const int x = children[c];
            const int a = inv_map[inv_off[x] + i];
            if (a >= 0) acc += V[vbase[x] * nrhs + (int64_t)r * F[x] + W[x] + a];
        }
        Vg[(int64_t)r * f + i] = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_diag(const int *__restrict__ nodes, const T *__restrict__ arena,
                                                  const int *__restrict__ F, const int *__restrict__ W,
                                                  const int64_t *__restrict__ off, const int *__restrict__ firstcol,
                                                  const int64_t *__restrict__ vbase, T *__restrict__ V,
                                                  T *__restrict__ Y, int ps, int n, int nrhs) {
    __shared__ T L[kPanel][kPanel + 1];
    __shared__ T y[kPanel][kRhsChunk + 1];
    const int g = nodes[blockIdx.x];
    const int f = F[g], w = W[g];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const T *M = arena + off[g];
    T *Vg = V + vbase[g] * nrhs;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r > c) L[r][c] = M[(int64_t)(ps + c) * f + ps + r];
    }
    for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nrhs - r0);
        for (int idx = tid; idx < nb * nr; idx += 256) {
            const int k = idx % nb, rr = idx / nb;
            y[k][rr] = Vg[(int64_t)(r0 + rr) * f + ps + k];
        }
        __syncthreads();
        for (int k = 0; k < nb - 1; ++k) {
            const int m = nb - k - 1;
            for (int idx = tid; idx < m * nr; idx += 256) {
                const int q = k + 1 + idx % m, rr = idx / m;
                const T yk = y[k][rr];
                if (yk != T(0)) y[q][rr] -= L[q][k] * yk;
            }
            __syncthreads();
        }
        for (int idx = tid; idx < nb * nr; idx += 256) {
            const int k = idx % nb, rr = idx / nb;
            Vg[(int64_t)(r0 + rr) * f + ps + k] = y[k][rr];
            Y[(int64_t)(r0 + rr) * n + firstcol[g] + ps + k] = y[k][rr];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_update(const int2 *__restrict__ tasks, const T *__restrict__ arena,
                                                    const int *__restrict__ F, const int *__restrict__ W,
                                                    const int64_t *__restrict__ off, const int64_t *__restrict__ vbase,
                                                    T *__restrict__ V, int ps, int nrhs) {
    __shared__ T y[kPanel][kRhsChunk + 1];
    const int2 tk = tasks[blockIdx.x];
    const int g = tk.x, f = F[g], w = W[g];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const int r0 = pe + tk.y * kTile, nrow = min(kTile, f - r0);
    const T *M = arena + off[g];
    T *Vg = V + vbase[g] * nrhs;
    const int tid = threadIdx.x, row = tid & 63;
    for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nrhs - c0);
        for (int idx = tid; idx < nb * nr; idx += 256) {
            const int k = idx % nb, rr = idx / nb;
            y[k][rr] = Vg[(int64_t)(c0 + rr) * f + ps + k];
        }
        __syncthreads();
        if (row < nrow)
            for (int rr = tid >> 6; rr < nr; rr += 4) {
                T acc = Vg[(int64_t)(c0 + rr) * f + r0 + row];
                for (int k = 0; k < nb; ++k) {
                    const T yk = y[k][rr];
                    if (yk != T(0)) acc -= M[(int64_t)(ps + k) * f + r0 + row] * yk;
                }
                Vg[(int64_t)(c0 + rr) * f + r0 + row] = acc;
            }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_partial(const int2 *__restrict__ tasks, const T *__restrict__ arena,
                                                     const int *__restrict__ F, const int *__restrict__ W,
                                                     const int64_t *__restrict__ off, const int *__restrict__ pat_ptr,
                                                     const int *__restrict__ pat_rows, const T *__restrict__ Y,
                                                     T *__restrict__ part, int ps, int n, int nrhs) {
    __shared__ T Lb[kTile][kPanel + 1];
    __shared__ T xs[kTile][kRhsChunk + 1];
    const int2 tk = tasks[blockIdx.x];
    const int g = tk.x, f = F[g], w = W[g];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const int r0 = pe + tk.y * kTile, nrow = min(kTile, f - r0);
    const T *M = arena + off[g];
    const int *rows = pat_rows + pat_ptr[g] + r0;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < kTile * nb; idx += 256) {
        const int i = idx % kTile, k = idx / kTile;
        Lb[i][k] = i < nrow ? M[(int64_t)(ps + k) * f + r0 + i] : T(0);
    }
    for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nrhs - c0);
        for (int idx = tid; idx < kTile * nr; idx += 256) {
            const int i = idx % kTile, rr = idx / kTile;
            xs[i][rr] = i < nrow ? Y[(int64_t)(c0 + rr) * n + rows[i]] : T(0);
        }
        __syncthreads();
        for (int idx = tid; idx < nb * nr; idx += 256) {
            const int k = idx % nb, rr = idx / nb;
            T acc = T(0);
            for (int i = 0; i < nrow; ++i) acc += Lb[i][k] * xs[i][rr];
            part[((int64_t)blockIdx.x * kPanel + k) * nrhs + c0 + rr] = acc;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_diag(const int *__restrict__ nodes, const int *__restrict__ task_begin,
                                                  const T *__restrict__ arena, const int *__restrict__ F,
                                                  const int *__restrict__ W, const int64_t *__restrict__ off,
                                                  const int *__restrict__ firstcol, const double *__restrict__ D,
                                                  T *__restrict__ Y, const T *__restrict__ part, int ps, int n,
                                                  int nrhs) {
    __shared__ T L[kPanel][kPanel + 1];
    __shared__ T z[kPanel][kRhsChunk + 1];
    const int g = nodes[blockIdx.x];
    const int f = F[g], w = W[g];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const int nrb = (f - pe + kTile - 1) / kTile, tb = task_begin[blockIdx.x];
    const int fc = firstcol[g] + ps;
    const T *M = arena + off[g];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nb * nb; idx += 256) {
        const int r = idx % nb, c = idx / nb;
        if (r > c) L[r][c] = M[(int64_t)(ps + c) * f + ps + r];
    }
    for (int c0 = 0; c0 < nrhs; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nrhs - c0);
        for (int idx = tid; idx < nb * nr; idx += 256) {
            const int k = idx % nb, rr = idx / nb;
            T s = T(0);
            for (int t = 0; t < nrb; ++t) s += part[((int64_t)(tb + t) * kPanel + k) * nrhs + c0 + rr];
            z[k][rr] = Y[(int64_t)(c0 + rr) * n + fc + k] / (T)D[fc + k] - s;
        }
        __syncthreads();
        for (int k = nb - 1; k > 0; --k) {
            for (int idx = tid; idx < k * nr; idx += 256) {
                const int m = idx % k, rr = idx / k;
                z[m][rr] -= L[k][m] * z[k][rr];
            }
            __syncthreads();
        }
        for (int idx = tid; idx < nb * nr; idx += 256) {
            const int k = idx % nb, rr = idx / nb;
            Y[(int64_t)(c0 + rr) * n + fc + k] = z[k][rr];
        }
        __syncthreads();
    }
}

// ---- refinement helpers ------------------------------------------------------

// R = B - A X over full CSR (one row per group of lanes), and per-column
// partial sums of R^2 and B^2 for the relative residual.
__global__ void k_residual(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v, int n,
                           int nrhs, const double *__restrict__ B, int ldb, const double *__restrict__ X, int ldx,
                           double *__restrict__ R, double *__restrict__ part) {
    // one thread per row; part[(col * gridDim.x + block) * 2 + {0,1}]
    __shared__ double sr[256], sb[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int c = 0; c < nrhs; ++c) {
        double rr = 0.0, bb = 0.0;
        if (i < n) {
            const double *xc = X + (int64_t)c * ldx;
            double acc = 0.0;
            for (int k = rp[i]; k < rp[i + 1]; ++k) acc += v[k] * xc[ci[k]];
            const double b = B[(int64_t)c * ldb + i];
            const double res = b - acc;
            R[(int64_t)c * n + i] = res;
            rr = res * res;
            bb = b * b;
        }
        sr[threadIdx.x] = rr;
        sb[threadIdx.x] = bb;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) {
                sr[threadIdx.x] += sr[threadIdx.x + h];
                sb[threadIdx.x] += sb[threadIdx.x + h];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[((int64_t)c * gridDim.x + blockIdx.x) * 2 + 0] = sr[0];
            part[((int64_t)c * gridDim.x + blockIdx.x) * 2 + 1] = sb[0];
        }
        __syncthreads();
    }
}

__global__ void k_residual_final(const double *__restrict__ part, int nblocks, int nrhs, double *__restrict__ relres) {
    __shared__ double sr[256], sb[256];
    for (int c = 0; c < nrhs; ++c) {
        double rr = 0.0, bb = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) {
            rr += part[((int64_t)c * nblocks + b) * 2 + 0];
            bb += part[((int64_t)c * nblocks + b) * 2 + 1];
        }
        sr[threadIdx.x] = rr;
        sb[threadIdx.x] = bb;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) {
                sr[threadIdx.x] += sr[threadIdx.x + h];
                sb[threadIdx.x] += sb[threadIdx.x + h];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double bn = sqrt(sb[0]);
            relres[c] = sqrt(sr[0]) / (bn > 0.0 ? bn : 1.0);
        }
        __syncthreads();
    }
}

__global__ void k_zero_cols(double *__restrict__ X, int ldx, int n, int nrhs) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    X[(e / n) * ldx + e % n] = 0.0;
}

__global__ void k_gather_values(const int *__restrict__ src, int64_t nnz, const double *__restrict__ upper,
                                double *__restrict__ full) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz) full[e] = upper[src[e]];
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

// Max |A_ii| (ComputeDiagonalScale, UtduNumericFactorization.cs:411-418; Math.Max
// propagates NaN; 1 when the maximum is not positive).
__global__ __launch_bounds__(1024) void k_diag_scale(const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                                     const double *__restrict__ vals, int n, double *__restrict__ out) {
    __shared__ double sm[1024];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        double d = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
            if (colidx[k] == i) d = vals[k];
        const double a = fabs(d);
        m = (isnan(a) || isnan(m)) ? NAN : fmax(m, a);
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const double a = sm[threadIdx.x], b = sm[threadIdx.x + h];
            sm[threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : fmax(a, b);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0] > 0.0 ? sm[0] : 1.0;
}

struct Numeric {
    std::shared_ptr<DevicePlan> plan;
    const Symbolic *sym = nullptr;
    int precision = MG_FP64;
    int n = 0;
    DevArray<unsigned char> arena;  // f^2 fronts, element type by precision
    DevArray<double> diag;
    DevArray<unsigned char> flags;
    DevArray<int> counts;
    DevArray<double> mags, scale;
    DevArray<double> values;  // upper CSR values (float64), kept for refinement
    Matrix A;                 // full CSR for residuals
    mg_numeric_info info{};
    // solve workspaces
    DevArray<unsigned char> wsY, wsU, wsP;
    DevArray<double> wsR, wsX, wsPart, wsRel;
    int ws_nrhs = 0;
};

static inline unsigned blocks_for(int64_t count, int bs) { return (unsigned)((count + bs - 1) / bs); }


// ---- launch timing (kernel-level roofline inside the library) ---------------

struct KernelTally {
    const char *name;
    int64_t launches = 0;
    double ms = 0, flops = 0, bytes = 0;
};
struct Profiler {
    bool on = false;
    std::vector<KernelTally> k;
    struct Pending {
        hipEvent_t a, b;
        int kind;
        double flops, bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    Profiler() {
        for (const char *nm : {"assemble_pull", "assemble_values", "panel_diag", "panel_rows", "schur", "forward", "backward"}) {
            KernelTally t;
            t.name = nm;
            k.push_back(t);
        }
    }
    hipEvent_t ev() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        MG_HIP(hipEventCreate(&e));
        return e;
    }
};
static Profiler g_prof;
enum { KP_PULL = 0, KP_VALUES, KP_PDIAG, KP_PROWS, KP_SCHUR, KP_FWD, KP_BWD };

struct ProfScope {
    hipStream_t st;
    int kind;
    double flops, bytes;
    hipEvent_t a = nullptr;
    ProfScope(hipStream_t s, int kd, double fl, double by) : st(s), kind(kd), flops(fl), bytes(by) {
        if (!g_prof.on) return;
        a = g_prof.ev();
        MG_HIP(hipEventRecord(a, st));
    }
    ~ProfScope() {
        if (!a) return;
        hipEvent_t b = g_prof.ev();
        (void)hipEventRecord(b, st);
        g_prof.pending.push_back({a, b, kind, flops, bytes});
    }
};

void profile_enable(int on) {
    g_prof.on = on != 0;
    for (auto &t : g_prof.k) t.launches = 0, t.ms = 0, t.flops = 0, t.bytes = 0;
}

int profile_read(int kind, const char **name, int64_t *launches, double *ms, double *flops, double *bytes) {
    if (kind < 0 || kind >= (int)g_prof.k.size()) return 0;
    if (!g_prof.pending.empty()) {
        MG_HIP(hipDeviceSynchronize());
        for (auto &p : g_prof.pending) {
            float el = 0;
            MG_HIP(hipEventElapsedTime(&el, p.a, p.b));
            auto &t = g_prof.k[p.kind];
            t.launches += 1;
            t.ms += el;
            t.flops += p.flops;
            t.bytes += p.bytes;
            g_prof.pool.push_back(p.a);
            g_prof.pool.push_back(p.b);
        }
        g_prof.pending.clear();
    }
    const auto &t = g_prof.k[kind];
    *name = t.name;
    *launches = t.launches;
    *ms = t.ms;
    *flops = t.flops;
    *bytes = t.bytes;
    return 1;
}

template <typename T>
static void run_factor(Numeric &N, const mg_options &opts, hipStream_t st) {
    DevicePlan &P = *N.plan;
    T *arena = reinterpret_cast<T *>(N.arena.p);
    const double e = sizeof(T);
    for (int L = 0; L < P.nlevels; ++L) {
        const int np = P.lvl_pull_ptr[L + 1] - P.lvl_pull_ptr[L];
        if (np > 0) {
            ProfScope ps_(st, KP_PULL, 0.0, e * P.lvl_pull_elems[L] + 4.0 * P.lvl_pull_index[L]);
            hipLaunchKernelGGL(k_assemble_pull<T>, dim3(np), dim3(256), 0, st, P.pull_tasks.p + P.lvl_pull_ptr[L], arena,
                               P.f.p, P.w.p, P.off.p, P.child_ptr.p, P.children.p, P.inv_off.p, P.inv_map.p);
        }
        const int64_t na = P.lvl_asm_ptr[L + 1] - P.lvl_asm_ptr[L];
        if (na > 0) {
            ProfScope ps_(st, KP_VALUES, 0.0, (double)na * (8.0 + 4.0 + 8.0 + 2 * e));
            hipLaunchKernelGGL(k_assemble_values<T>, dim3(blocks_for(na, 256)), dim3(256), 0, st,
                               P.asm_pos.p + P.lvl_asm_ptr[L], P.asm_vidx.p + P.lvl_asm_ptr[L], na, N.values.p, arena);
        }
        for (int q = 0; q < P.lvl_npanels[L]; ++q) {
            const int ps = q * kPanel;
            const int b0 = P.lvl_panel_ptr[L][q], b1 = P.lvl_panel_ptr[L][q + 1];
            const int d0 = P.lvl_diag_ptr[L][q], d1 = P.lvl_diag_ptr[L][q + 1];
            if (d1 > d0) {
                ProfScope ps_(st, KP_PDIAG, 0.0, 0.0);
                hipLaunchKernelGGL(k_panel_diag<T>, dim3(d1 - d0), dim3(256), 0, st, P.diag_nodes.p + d0, arena, P.f.p,
                                   P.w.p, P.off.p, P.first.p, ps, N.scale.p, opts.pivot_tolerance,
                                   opts.diagonal_regularization, opts.allow_diagonal_regularization, N.diag.p,
                                   N.flags.p);
            }
            if (b1 > b0) {
                ProfScope ps_(st, KP_PROWS, 0.0, 0.0);
                hipLaunchKernelGGL(k_panel_rows<T>, dim3(b1 - b0), dim3(256), 0, st, P.panel_tasks.p + b0, arena, P.f.p,
                                   P.w.p, P.off.p, P.first.p, ps, N.diag.p);
            }
            const int u0 = P.lvl_upd_ptr[L][q], u1 = P.lvl_upd_ptr[L][q + 1];
            if (u1 > u0) {
                ProfScope ps_(st, KP_SCHUR, P.lvl_upd_flops[L][q], e * P.lvl_upd_elems[L][q]);
                hipLaunchKernelGGL(k_schur<T>, dim3(u1 - u0), dim3(256), 0, st, P.upd_tasks.p + u0, arena, P.f.p, P.w.p,
                                   P.off.p, P.first.p, q, N.diag.p);
            }
        }
    }
    MG_HIP(hipGetLastError());
}

void factorize_device(Symbolic &S, const double *d_values, int precision, const mg_options &opts, hipStream_t st,
                      Numeric *&num) {
    if (precision != MG_FP64 && precision != MG_FP32) throw Error(MG_E_ARG, "Unknown factor precision.");
    if (!S.plan) S.plan = build_plan(S);
    if (S.plan->max_w > 8192)
        throw Error(MG_E_TOO_DENSE, "A supernode wider than 8192 columns is not supported on the device.");
    const size_t esz = precision == MG_FP32 ? sizeof(float) : sizeof(double);
    Numeric *N = num;
    if (N && (N->sym != &S || N->precision != precision))
        throw Error(MG_E_ARG, "The numeric handle belongs to another symbolic factorization or precision.");
    bool created = false;
    if (!N) {
        N = new Numeric();
        created = true;
    }
    try {
        if (created) {
            N->plan = S.plan;
            N->sym = &S;
            N->precision = precision;
            N->n = S.n;
            N->arena.alloc((size_t)S.plan->arena * esz);
            N->diag.alloc(std::max(S.n, 1));
            N->flags.alloc(std::max(S.n, 1));
            N->counts.alloc(4);
            N->mags.alloc(4);
            N->scale.alloc(1);
            N->values.alloc(std::max(S.matrix_nnz, 1));
            N->A.build(S.n, S.a_rowptr.data(), S.a_colidx.data());
        }
        N->info = mg_numeric_info{};
        N->info.precision = precision;
        if (S.n > 0) {
            if (S.matrix_nnz > 0)
                MG_HIP(hipMemcpyAsync(N->values.p, d_values, sizeof(double) * S.matrix_nnz, hipMemcpyDeviceToDevice, st));
            N->A.set_values_dev(N->values.p, st);
            hipLaunchKernelGGL(k_diag_scale, dim3(1), dim3(1024), 0, st, N->A.rowptr.p, N->A.colidx.p, N->A.vals.p, S.n,
                               N->scale.p);
            if (precision == MG_FP32) run_factor<float>(*N, opts, st);
            else run_factor<double>(*N, opts, st);
            hipLaunchKernelGGL(k_pivot_stats, dim3(1), dim3(1024), 0, st, N->diag.p, N->flags.p, S.n, N->counts.p,
                               N->mags.p);
            MG_HIP(hipGetLastError());
            int counts[3];
            double mags[2];
            MG_HIP(hipMemcpyAsync(counts, N->counts.p, sizeof(counts), hipMemcpyDeviceToHost, st));
            MG_HIP(hipMemcpyAsync(mags, N->mags.p, sizeof(mags), hipMemcpyDeviceToHost, st));
            MG_HIP(hipStreamSynchronize(st));
            if (counts[2] > 0)
                throw Error(MG_E_NOT_PD,
                            "Numerical factorization has not completed: a pivot is not positive (the matrix is not "
                            "positive definite). Enable diagonal regularization if factorization must be completed.");
            N->info.regularized_pivot_count = counts[0];
            N->info.negative_pivot_count = counts[1];
            N->info.smallest_pivot_magnitude = mags[0];
            N->info.largest_pivot_magnitude = mags[1];
        }
        N->info.is_positive_definite = N->info.regularized_pivot_count == 0 && N->info.negative_pivot_count == 0;
    } catch (...) {
        if (created) delete N;
        throw;
    }
    num = N;
}

void numeric_release(Numeric *num) { delete num; }

void numeric_info(const Numeric *N, mg_numeric_info &info, double *diag_host) {
    info = N->info;
    if (diag_host && N->n > 0) MG_HIP(hipMemcpy(diag_host, N->diag.p, sizeof(double) * N->n, hipMemcpyDeviceToHost));
}

template <typename T>
static void run_solve_once(Numeric &N, int nrhs, const double *d_rhs, int ldr, double *d_x, int ldx, int accumulate,
                           hipStream_t st) {
    DevicePlan &P = *N.plan;
    const int n = N.n;
    T *Y = reinterpret_cast<T *>(N.wsY.p);
    T *V = reinterpret_cast<T *>(N.wsU.p);
    T *part = reinterpret_cast<T *>(N.wsP.p);
    const T *arena = reinterpret_cast<const T *>(N.arena.p);
    const int64_t tot = (int64_t)n * nrhs;
    hipLaunchKernelGGL(k_permute_in<T>, dim3(blocks_for(tot, 256)), dim3(256), 0, st, d_rhs, ldr, P.perm.p, n, nrhs, Y);
    for (int L = 0; L < P.nlevels; ++L) {
        const int i0 = P.lvl_init_ptr[L], i1 = P.lvl_init_ptr[L + 1];
        hipLaunchKernelGGL(k_fwd_init<T>, dim3(i1 - i0), dim3(256), 0, st, P.init_tasks.p + i0, P.f.p, P.w.p, P.first.p,
                           P.child_ptr.p, P.children.p, P.inv_off.p, P.inv_map.p, P.vbase.p, V, Y, n, nrhs);
        for (int q = 0; q < P.lvl_npanels[L]; ++q) {
            const int ps = q * kPanel;
            const int d0 = P.lvl_diag_ptr[L][q], d1 = P.lvl_diag_ptr[L][q + 1];
            const int b0 = P.lvl_panel_ptr[L][q], b1 = P.lvl_panel_ptr[L][q + 1];
            hipLaunchKernelGGL(k_fwd_diag<T>, dim3(d1 - d0), dim3(256), 0, st, P.diag_nodes.p + d0, arena, P.f.p, P.w.p,
                               P.off.p, P.first.p, P.vbase.p, V, Y, ps, n, nrhs);
            if (b1 > b0)
                hipLaunchKernelGGL(k_fwd_update<T>, dim3(b1 - b0), dim3(256), 0, st, P.panel_tasks.p + b0, arena, P.f.p,
                                   P.w.p, P.off.p, P.vbase.p, V, ps, nrhs);
        }
    }
    for (int L = P.nlevels - 1; L >= 0; --L) {
        for (int q = P.lvl_npanels[L] - 1; q >= 0; --q) {
            const int ps = q * kPanel;
            const int d0 = P.lvl_diag_ptr[L][q], d1 = P.lvl_diag_ptr[L][q + 1];
            const int b0 = P.lvl_panel_ptr[L][q], b1 = P.lvl_panel_ptr[L][q + 1];
            if (b1 > b0)
                hipLaunchKernelGGL(k_bwd_partial<T>, dim3(b1 - b0), dim3(256), 0, st, P.panel_tasks.p + b0, arena, P.f.p,
                                   P.w.p, P.off.p, P.pat_ptr.p, P.pat_rows.p, Y, part, ps, n, nrhs);
            hipLaunchKernelGGL(k_bwd_diag<T>, dim3(d1 - d0), dim3(256), 0, st, P.diag_nodes.p + d0, P.diag_tb.p + d0,
                               arena, P.f.p, P.w.p, P.off.p, P.first.p, N.diag.p, Y, part, ps, n, nrhs);
        }
    }
    hipLaunchKernelGGL(k_permute_out<T>, dim3(blocks_for(tot, 256)), dim3(256), 0, st, Y, P.perm.p, n, nrhs, d_x, ldx,
                       accumulate);
    MG_HIP(hipGetLastError());
}

static void ensure_ws(Numeric &N, int nrhs) {
    if (N.ws_nrhs >= nrhs) return;
    const size_t esz = N.precision == MG_FP32 ? sizeof(float) : sizeof(double);
    const int n = std::max(N.n, 1);
    N.wsY.alloc((size_t)n * nrhs * esz);
    N.wsU.alloc((size_t)std::max<int64_t>(N.plan->vsum, 1) * nrhs * esz);
    N.wsP.alloc((size_t)std::max(N.plan->max_panel_tasks, 1) * kPanel * nrhs * esz);
    N.wsR.alloc((size_t)n * nrhs);
    const int nb = (int)blocks_for(n, 256);
    N.wsPart.alloc((size_t)nb * nrhs * 2);
    N.wsRel.alloc(nrhs);
    N.ws_nrhs = nrhs;
}

// relres[c] = ||B - A X||/||B|| per column; R receives the residuals.
static void residual(Numeric &N, int nrhs, const double *d_b, int ldb, const double *d_x, int ldx, hipStream_t st,
                     std::vector<double> &relres) {
    const int n = N.n;
    const int nb = (int)blocks_for(n, 256);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, N.A.rowptr.p, N.A.colidx.p, N.A.vals.p, n, nrhs, d_b, ldb,
                       d_x, ldx, N.wsR.p, N.wsPart.p);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, st, N.wsPart.p, nb, nrhs, N.wsRel.p);
    relres.resize(nrhs);
    MG_HIP(hipMemcpyAsync(relres.data(), N.wsRel.p, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
}

// UtduNumericFactorization.Solve (:90-112) for nrhs columns, followed by FP64
// iterative refinement: x += solve(b - A x) until max relres <= tol.
void solve_device(Numeric *Np, int nrhs, const double *d_b, int ldb, double *d_x, int ldx, double tol, int max_refine,
                  hipStream_t st, mg_solve_report *rep) {
    Numeric &N = *Np;
    if (nrhs < 0) throw Error(MG_E_ARG, "The number of right-hand sides cannot be negative.");
    if (rep) *rep = mg_solve_report{0, 1, 0.0};
    if (N.n == 0 || nrhs == 0) return;
    if (ldb < N.n || ldx < N.n) throw Error(MG_E_ARG, "The right-hand side length does not match the matrix size.");
    ensure_ws(N, nrhs);
    auto once = [&](const double *rhs, int ldr, int acc) {
        if (N.precision == MG_FP32) run_solve_once<float>(N, nrhs, rhs, ldr, d_x, ldx, acc, st);
        else run_solve_once<double>(N, nrhs, rhs, ldr, d_x, ldx, acc, st);
    };
    once(d_b, ldb, 0);
    if (!rep && max_refine <= 0) return;
    std::vector<double> rel;
    residual(N, nrhs, d_b, ldb, d_x, ldx, st, rel);
    double worst = *std::max_element(rel.begin(), rel.end());
    int steps = 0;
    while (steps < max_refine && !(worst <= tol)) {
        once(N.wsR.p, N.n, 1);
        ++steps;
        residual(N, nrhs, d_b, ldb, d_x, ldx, st, rel);
        worst = *std::max_element(rel.begin(), rel.end());
    }
    if (rep) {
        rep->refinement_steps = steps;
        rep->relative_residual = worst;
        rep->converged = worst <= tol;
    }
}

// ---- test hooks -----------------------------------------------------------

// Factor block of supernode s as MAGLA stores it (UtduNumericFactorization.cs:517-519, 591):
// f_s x w_s column-major over the supernode's own pattern rows, unit diagonal, L below;
// read out of the fused device front that holds it, as float64.
void numeric_front(const Numeric *N, int s, double *out) {
    const DevicePlan &P = *N->plan;
    const Symbolic &S = *N->sym;
    if (s < 0 || s >= P.nsuper) throw Error(MG_E_RANGE, "supernode index is out of range.");
    const int g = P.front_of[s];
    const int fg = P.h_f[g];
    const int64_t cnt = (int64_t)fg * fg;
    std::vector<double> front(cnt);
    if (N->precision == MG_FP32) {
        std::vector<float> tmp(cnt);
        MG_HIP(hipMemcpy(tmp.data(), reinterpret_cast<const float *>(N->arena.p) + P.h_off[g], cnt * sizeof(float),
                         hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < cnt; ++e) front[e] = tmp[e];
    } else {
        MG_HIP(hipMemcpy(front.data(), reinterpret_cast<const double *>(N->arena.p) + P.h_off[g],
                         cnt * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::vector<int> rowmap(S.n, -1);
    for (int t = 0; t < fg; ++t) rowmap[P.h_pat_rows[P.h_pat_ptr[g] + t]] = t;
    const int fs = S.front(s), ws = S.width(s);
    const int c0 = S.first[s] - P.h_first[g];
    for (int k = 0; k < ws; ++k)
        for (int t = 0; t < fs; ++t)
            out[(int64_t)k * fs + t] = front[(int64_t)(c0 + k) * fg + rowmap[S.pat_rows[S.pat_ptr[s] + t]]];
}

// One 16x16x4 MFMA with A[i][k] = (k==0 ? i : k==1), B[k][j] = (k==0 ? 1 : k==1 ? 16 j : 0), so that
// D[i][j] = i + 16 j; operands use the lane mapping of the Schur kernel and each accumulator
// register is stored at the (row, column) MfmaAcc<T>::row assigns to it: out[i*16 + j] must equal
// i + 16 j.
template <typename T>
__global__ void k_mfma_probe(double *out) {
    const int lane = threadIdx.x;
    const int i = lane & 15, k = lane >> 4;
    const T a = k == 0 ? T(i) : (k == 1 ? T(1) : T(0));
    const T b = k == 0 ? T(1) : (k == 1 ? T(16 * (lane & 15)) : T(0));
    typename MfmaAcc<T>::type acc;
    for (int q = 0; q < 4; ++q) acc[q] = T(0);
    acc = MfmaAcc<T>::step(a, b, acc);
    for (int q = 0; q < 4; ++q) out[MfmaAcc<T>::row(k, q) * 16 + (lane & 15)] = (double)acc[q];
}

void selftest_mfma(int precision, double *host_out) {
    DevArray<double> d(256);
    if (precision == MG_FP32) hipLaunchKernelGGL(k_mfma_probe<float>, dim3(1), dim3(64), 0, 0, d.p);
    else hipLaunchKernelGGL(k_mfma_probe<double>, dim3(1), dim3(64), 0, 0, d.p);
    MG_HIP(hipGetLastError());
    MG_HIP(hipMemcpy(host_out, d.p, 256 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace mg
