// factor.hip -- level-scheduled multifrontal U^T D U factorization and the
// supernodal solves on MI355X (gfx950).
//
// Device layout
//   * A front of f rows and w pivot columns lives in two places of one arena
//     (elements of the factor precision, layout in front.h).  Its pivot
//     columns [0, w) are an f x w column-major block in the factor store,
//     kept for the solves: unit diagonal, L below it (MAGLA's factor column
//     k = [1, L[k+1.., k]], UtduNumericFactorization.cs:517-519 / 591).  Its
//     trailing t x t block (t = f - w) is the contribution read by the
//     parent (:521-530); it lives in a contribution pool only from the
//     front's level to its parent's, and blocks whose lifetimes do not
//     overlap share memory, as the reference's peak-contribution estimate
//     assumes (SupernodalStructure.cs:368-418).
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
#include <atomic>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <climits>
#include <map>
#include <mutex>
#include <set>
#include <type_traits>

#include "device.h"
#include "front.h"
#include "mfma.h"

namespace mg {

constexpr int kTile = 64;   // tile edge of assembly / Schur tiles
constexpr int kPanel = 64;  // panel width: columns factored by one panel_diag / panel_rows pass
constexpr int kBlock = 256; // Schur blocking: panels of one block update the rest of its outer block at once
constexpr int kOuter = 1024; // outer blocking: the trailing update after every kOuter columns
constexpr int kSBlock = 1024; // columns of one block step of the large-front solves (diagonal blocks inverted)
static_assert(kOuter % kBlock == 0 && kBlock % kPanel == 0, "nested blocking");
constexpr int kMaxFrontWidth = 4096;  // widest chain fusion
// Schur tile shape of the outer steps per front-size bucket (f < 128, < 256, < 512, < 1024,
// < 2048, >= 2048): 64x64 (k_schur) or 128x128 (k_schur128) output tiles, chosen by
// autotune_device; 128 for f >= 1024 until tuned.  Inner and middle steps (narrow targets)
// use 64x64.
constexpr int kSchurBuckets = MG_SCHUR_BUCKETS;
constexpr int kUb64Stride = 2 * (kSchurBuckets + 1) + 1;  // 64x64 task-list offsets per (level, panel)
constexpr int kUb128Stride = 2 * kSchurBuckets + 1;       // 128x128 task-list offsets per (level, panel)
static inline int schur_bucket(int f) {
    int b = 0;
    for (int lim = 128; b < kSchurBuckets - 1 && f >= lim; lim *= 2) ++b;
    return b;
}
constexpr int kSmallFront = 64;       // fronts this small are solved by one wave each (all panels, all RHS)
constexpr int kSmallFactor = 32;      // fronts this small are factored whole by a lane group (k_factor_small)
constexpr int kArenaPad = 64;         // tail elements after the factor store and after the pool (16-byte vector over-reads)
constexpr int kPoolAlign = 64;        // contribution blocks start at multiples of this many elements

typedef unsigned int mg_u4u __attribute__((ext_vector_type(4), aligned(4)));  // unaligned 16-byte load
constexpr int kFootChunk = 256;       // footer rows per partial product in the backward solve
constexpr int kRhsChunk = 16;         // right-hand sides per pass of the solve kernels
constexpr int kPanelFusedRows = 4;    // fronts with at most this many row blocks below a panel: one launch

// A device front is a chain of consecutive supernodes s, s+1, ..., each the
// elimination-tree parent of the previous one (the last column of s has the
// first column of s+1 as etree parent), fused when the rows nest: the fused
// front has the chain's columns followed by the top supernode's footer, so
// the lower members' columns carry at most a few structural zeros.  The
// elimination order, the supernode partition and every factor entry are
// those of the reference; only the numeric schedule differs (no extend-add
// between chain members, Schur updates of up to kBlock columns).
// A Schur tile task with the front's size, width, first column and arena offset (one
// dependent load before the panel loads instead of two).
struct __align__(8) SchurTask {
    int g, i0, j0, f, w, fc;  // front, tile's first row and column, size, width, first column
    int64_t off, coff;        // front's factor-store and contribution-pool offsets
};

// A workgroup's task in the panel solves, carrying the front's sizes, offsets and system so
// the kernel's data loads depend on one load instead of two.
struct __align__(8) PanelTask {
    int g, blk, f, w;     // front, row block (forward) or column block (backward), size, width
    int64_t off, vbase;   // front's arena offset and vector base
    int fc, sys;          // first column, system
};

// A front of the small-front solve lists with everything the solve kernels read of it, so
// that the lane groups' data loads depend on this one load.
struct __align__(16) SmallRec {
    int f, w, fc, sys;    // size, width, first column, system
    int64_t off, vbase;   // factor-store offset, vector base
    int c0, c1, pat, pad; // children-array range, first pattern row
    int64_t inv0;         // inverse row map of its first child (the children's maps follow, f entries each)
    int64_t coff;         // contribution-pool offset
};
// A children-array slot: the child's vector base, factor-store and pool offsets, size and width.
struct __align__(16) SmallChild {
    int64_t vb, off, coff;
    int f, w;
};

struct DevicePlan {
    int n = 0, nsuper = 0, nfront = 0, nlevels = 0, max_w = 0, max_f = 0;
    int64_t arena = 0;  // factor store, pad, contribution pool, pad (elements)
    int64_t fstore = 0; // factor store: sum f w over fronts
    int64_t pool = 0;   // contribution pool (peak of the live blocks, with fragmentation)
    int64_t pool_live = 0;  // largest sum of contribution blocks live at once
    int64_t usum = 0;   // sum (f - w) over fronts
    std::vector<int> h_f, h_w, h_first;
    std::vector<int64_t> h_off, h_coff;  // per front: factor-store offset, contribution-pool offset
    std::vector<int> h_level, h_parent;  // per front: level (height in the front tree), parent front (-1: root)
    std::vector<int> front_of;   // supernode -> front
    std::vector<int> h_pat_ptr, h_pat_rows;
    std::vector<int> lvl_node_ptr, lvl_pull_ptr;
    std::vector<int64_t> lvl_asm_ptr;
    std::vector<int> lvl_npanels;
    std::vector<std::vector<int>> lvl_panel_ptr, lvl_ptask_ptr;  // per panel: row-block tasks, panel tasks
    // Schur tasks per level and panel q, in two parts p (0: before the next panel, 1: may overlap
    // it).  upd_tasks (64x64 tiles), from u = &ub64[q * kUb64Stride + p * (kB + 1)]: inner and
    // middle steps [u[0], u[1]), outer steps of bucket b [u[1 + b], u[2 + b]).  upd128_tasks
    // (128x128 tiles, outer steps only), from u = &ub128[q * kUb128Stride + p * kB]: bucket b
    // [u[b], u[b + 1]).  The outer steps of a bucket run in the shape schur128 selects.
    std::vector<std::vector<int>> lvl_ub64, lvl_ub128;
    // per panel: 1 when some front's step writes columns past its 256-column block (those
    // steps wait for the previous step's overlapped part)
    std::vector<std::vector<char>> lvl_beyond;
    // [precision is FP32][bucket]: 1 = 128x128 tiles for the outer steps (fronts of 1024 rows
    // and more until tuned)
    int schur128[2][kSchurBuckets] = {{0, 0, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1}};
    int64_t bucket_tasks[kSchurBuckets] = {};
    std::vector<int> lvl_init_ptr;  // solve: (front, 256-row chunk) tasks of the forward gather
    std::vector<int> lvl_small_ptr; // solve: fronts of at most kSmallFront rows (one wave each),
                                    // per level four classes: f <= 8, <= 16, <= 32, <= 64
    DevArray<int> small_nodes;
    DevArray<SmallRec> small_recs;     // small_nodes' records, same order
    DevArray<SmallChild> child_recs;   // per children-array slot
    std::vector<int> lvl_foot_ptr;  // solve: (front, 64-column block, row chunk) footer partial products
    std::vector<int> lvl_footr_ptr; // solve: (front, 64-column block) footer reductions

    // large-front solves per kSBlock-column block J of a level: forward y of the block's
    // sub-blocks (fa) and the rows below it (fb), backward x of the sub-blocks (ba) and the
    // 64-column blocks to its left (bb); task ranges [ptr[J], ptr[J + 1]) of the task arrays
    std::vector<std::vector<int>> lvl_fa_ptr, lvl_fb_ptr, lvl_ba_ptr, lvl_bb_ptr;
    // solve-block (kSBlock columns) diagonal inverses (end of the factorization): (front, block,
    // a, b) tasks of distance d = a - b in [inv_ptr[d - 1], inv_ptr[d])
    int inv_ptr[kSBlock / kPanel] = {};
    DevArray<int4> inv_tasks;
    int64_t vsum = 0;               // sum f over fronts (solve front vectors)
    // algorithmic work per launch
    std::vector<double> lvl_pull_elems, lvl_pull_index;
    // solve work per level: factor entries the sweep reads (L block of each front with the
    // stored inverse triangle, f w - w (w - 1) / 2), sum of f and sum of w over the fronts
    std::vector<double> lvl_solve_fac, lvl_solve_f, lvl_solve_w;
    std::vector<std::vector<double>> lvl_upd_flops, lvl_upd_elems;
    std::vector<std::vector<double>> lvl_panel_elems;  // sum over fronts of (f - pe) * nb (L below the panel)

    int nsys = 1;                   // independent systems merged into this plan (batched solves)
    std::vector<int> h_sys_begin;   // first row of each system, nsys+1
    DevArray<int> front_sys, sys_begin;
    DevArray<int> row_sys;  // system of every row (elimination order = input order per system)
    DevArray<int> f, w, first, child_ptr, children, pat_ptr, pat_rows, level_nodes, inv_map, perm, iperm;
    // per inverse-map entry a (>= 0): the child's contribution column a as an offset from its
    // pool offset, contrib_col(0, t, a) (below 2^32: t (t + 1) / 2 <= INT32_MAX); ~0u where a < 0
    DevArray<uint32_t> inv_col;
    DevArray<int64_t> off, coff, inv_off, u_off, asm_pos, vbase;
    DevArray<int2> init_tasks;
    DevArray<PanelTask> fa_tasks, fb_tasks, ba_tasks, bb_tasks;
    DevArray<int4> footr_tasks;    // (front, column block, first partial task, partial count)
    DevArray<int4> foot_tasks;     // (front, column block, row chunk, index of the chunk in the block)

    int max_foot_tasks = 0;
    DevArray<int> asm_vidx;
    DevArray<int4> pull_tasks;
    DevArray<SchurTask> upd_tasks, upd128_tasks;
    DevArray<int2> panel_tasks;  // k_panel_rows: (front, row block) of fronts with many row blocks
    DevArray<int2> ptasks;       // k_panel: (front, row block), row block -1 = diagonal block only
};

int plan_level_count(const DevicePlan &P) { return P.nlevels; }

// Schur work after panel q of a front of width w and size f, blocked at three widths
// (64-column panels in 256-column blocks in kOuter-column outer blocks):
//  * inner (not the block's last panel): the next panel's columns [pe, pe + 64), rows [pe, f),
//    with every panel of the block so far (K = pe - bs <= 192; left-looking inside the block:
//    each column of the block is updated once, just before its panel, with the K of all the
//    block's panels before it);
//  * middle (the block's last panel, not the outer block's last): the outer block's remaining
//    columns [be, oe), rows [be, f), with the block (K <= 256);
//  * outer (the outer block's last panel): the trailing [oe, f)^2 with the whole outer block
//    (K <= kOuter), which keeps the matrix cores fed longest per tile.
// Host and device decode a step the same way.
struct SchurStep {
    bool any, inner, middle;
    int ps, pe, bs, be, os, oe, base, kb, ke, cend;
};
__host__ __device__ inline SchurStep schur_step(int f, int w, int q) {
    SchurStep st{};
    st.ps = q * kPanel;
    st.pe = st.ps + kPanel < w ? st.ps + kPanel : w;
    st.bs = (st.ps / kBlock) * kBlock;
    st.be = st.bs + kBlock < w ? st.bs + kBlock : w;
    st.os = (st.ps / kOuter) * kOuter;
    st.oe = st.os + kOuter < w ? st.os + kOuter : w;
    st.inner = st.pe < st.be;
    st.middle = !st.inner && st.be < st.oe;
    if (st.inner) {
        st.base = st.pe; st.kb = st.bs; st.ke = st.pe; st.cend = st.pe + kPanel < st.be ? st.pe + kPanel : st.be;
    } else if (st.middle) {
        st.base = st.be; st.kb = st.bs; st.ke = st.be; st.cend = st.oe;
    } else {
        st.base = st.oe; st.kb = st.os; st.ke = st.oe; st.cend = f;
    }
    st.any = st.base < f;
    return st;
}

// Contribution pool.  Front g's block (contrib_size(f - w) elements) is written by the assembly
// of g's level and read by the assembly of its parent's level, so it is live over the levels
// [level(g), level(parent)], both included; levels run one after another, so blocks whose level
// ranges do not meet may share memory.  Before the fronts of level L are placed, the blocks last
// read at level L - 1 are released (coalesced with free neighbours); the level's blocks then go,
// largest first, to the smallest free range that holds them (lowest offset among equals), or at
// the end of the pool.  O(log) per block: batches have millions of fronts.  Sets h_coff (arena
// offsets: the pool follows the factor store and its padding), pool, pool_live and arena.
static void assign_contributions(DevicePlan &P, const std::vector<int> &level, const std::vector<int> &parent,
                                 const std::vector<std::vector<int>> &bylevel) {
    const int nf = P.nfront, nl = P.nlevels;
    std::vector<int64_t> size(nf, 0);
    for (int g = 0; g < nf; ++g)
        size[g] = (contrib_size(P.h_f[g] - P.h_w[g]) + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
    std::vector<int64_t> at(nf, 0);
    std::vector<std::vector<int>> release(nl + 1);        // blocks released before level L
    std::map<int64_t, int64_t> by_off;                     // free ranges: offset -> size, none adjacent
    std::set<std::pair<int64_t, int64_t>> by_size;         // the same ranges as (size, offset)
    auto take = [&](std::map<int64_t, int64_t>::iterator it) {
        by_size.erase({it->second, it->first});
        return by_off.erase(it);
    };
    int64_t top = 0, live = 0;  // the pool's free tail [top, ...) is not in the lists
    P.pool_live = 0;
    for (int L = 0; L < nl; ++L) {
        for (int g : release[L]) {
            int64_t o = at[g], sz = size[g];
            live -= sz;
            auto nx = by_off.lower_bound(o);
            if (nx != by_off.end() && nx->first == o + sz) {
                sz += nx->second;
                nx = take(nx);
            }
            if (nx != by_off.begin()) {
                auto pv = std::prev(nx);
                if (pv->first + pv->second == o) {
                    o = pv->first;
                    sz += pv->second;
                    take(pv);
                }
            }
            if (o + sz == top) {
                top = o;
            } else {
                by_off[o] = sz;
                by_size.insert({sz, o});
            }
        }
        std::vector<int> blocks;
        for (int g : bylevel[L])
            if (size[g] > 0) blocks.push_back(g);
        std::stable_sort(blocks.begin(), blocks.end(), [&](int a, int b) { return size[a] > size[b]; });
        for (int g : blocks) {
            const int64_t sz = size[g];
            int64_t o;
            auto fit = by_size.lower_bound({sz, INT64_MIN});
            if (fit != by_size.end()) {
                o = fit->second;
                const int64_t rest = fit->first - sz;
                take(by_off.find(o));
                if (rest > 0) {
                    by_off[o + sz] = rest;
                    by_size.insert({rest, o + sz});
                }
            } else {
                o = top;
                top += sz;
            }
            at[g] = o;
            live += sz;
            P.pool_live = std::max(P.pool_live, live);
            release[(parent[g] >= 0 ? level[parent[g]] : L) + 1].push_back(g);
        }
        P.pool = std::max(P.pool, top);
    }
    const int64_t base = P.fstore + kArenaPad;
    P.h_coff.assign(nf, base);
    for (int g = 0; g < nf; ++g)
        if (size[g] > 0) P.h_coff[g] = base + at[g];
    P.arena = base + P.pool + kArenaPad;
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
        P->h_off[g + 1] = P->h_off[g] + (int64_t)f * w;
        uoff[g + 1] = uoff[g] + (f - w);
    }
    P->fstore = P->h_off[nf];
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
    assign_contributions(*P, level, fparent, bylevel);
    P->h_level = level;
    P->h_parent = fparent;

    // inverse row maps: for child front x of front g, inv_x[i] = index of g's row i in x's footer;
    // laid out in children-array order, so the maps of g's children are consecutive (f_g each)
    std::vector<int64_t> inv_off(nf + 1, 0);
    int64_t inv_total = 0;
    for (int g = 0; g < nf; ++g)
        for (int c = fchild_ptr[g]; c < fchild_ptr[g + 1]; ++c) {
            inv_off[fchildren[c]] = inv_total;
            inv_total += P->h_f[g];
        }
    for (int g = 0; g < nf; ++g)
        if (fparent[g] < 0) inv_off[g] = inv_total;  // roots: no map
    inv_off[nf] = inv_total;
    std::vector<int> inv_map(inv_total, -1);
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

    std::vector<int> level_nodes;
    std::vector<int4> pull, upd, upd128;
    std::vector<int2> panel, init, ptask;
    std::vector<int64_t> vbase(nf + 1, 0);
    for (int g = 0; g < nf; ++g) vbase[g + 1] = vbase[g] + P->h_f[g];
    P->vsum = vbase[nf];
    P->lvl_init_ptr.push_back(0);
    P->lvl_small_ptr.push_back(0);
    std::vector<int> small;
    P->lvl_foot_ptr.push_back(0);
    P->lvl_footr_ptr.push_back(0);
    std::vector<int2> fa, fb, ba, bb;
    std::vector<int4> foot, footr;
    std::vector<int4> inv_d[kSBlock / kPanel - 1];

    std::vector<int64_t> asm_pos;
    std::vector<int> asm_vidx;
    P->lvl_node_ptr.push_back(0);
    P->lvl_pull_ptr.push_back(0);
    P->lvl_asm_ptr.push_back(0);
    for (int L = 0; L < nl; ++L) {
        int npan = 0;
        double pe_elems = 0, pe_index = 0;
        std::vector<int> small_cls[4];
        for (int g : bylevel[L]) {
            level_nodes.push_back(g);
            const int f = P->h_f[g], w = P->h_w[g];
            npan = std::max(npan, (w + kPanel - 1) / kPanel);
            if (f <= kSmallFront) small_cls[f <= 8 ? 0 : (f <= 16 ? 1 : (f <= 32 ? 2 : 3))].push_back(g);
            else
                for (int c = 0; c < (f + 255) / 256; ++c) init.push_back(make_int2(g, c));
            // 64x64 tiles of the lower triangle, (front, first row, first column): the pivot
            // columns tiled from column 0, the contribution's columns from column w, so that a
            // tile's columns lie in one region (small fronts: k_assemble_small)
            if (f > kSmallFactor)
                for (int fam = 0; fam < 2; ++fam)
                    for (int j0 = fam ? w : 0; j0 < (fam ? f : w); j0 += kTile)
                        for (int i0 = j0; i0 < f; i0 += kTile) pull.push_back(make_int4(g, i0, j0, 0));
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
        // solve tasks
        const int foot_base = (int)foot.size();
        for (int g : bylevel[L]) {
            const int f = P->h_f[g], w = P->h_w[g];
            if (f <= kSmallFront) continue;
            const int nchunk = (f - w + kFootChunk - 1) / kFootChunk;
            for (int cb = 0; cb < (w + kPanel - 1) / kPanel; ++cb) {
                footr.push_back(make_int4(g, cb, (int)foot.size() - foot_base, nchunk));
                for (int ch = 0; ch < nchunk; ++ch) foot.push_back(make_int4(g, cb, ch, 0));
            }
        }
        P->max_foot_tasks = std::max(P->max_foot_tasks, (int)foot.size() - foot_base);
        // per block J: tasks [ptr[2J], ptr[2J + 1]) are those of blocks of at most 64 columns
        // (one tile in registers), [ptr[2J + 1], ptr[2J + 2]) those of wider blocks
        // (kernels with NS = 1 and NS = kSSub)
        std::vector<int> pfa{(int)fa.size()}, pfb{(int)fb.size()}, pba{(int)ba.size()}, pbb{(int)bb.size()};
        const int nblk = (npan + kSBlock / kPanel - 1) / (kSBlock / kPanel);
        for (int J = 0; J < nblk; ++J) {
            std::vector<int2> t[4][2];  // fa, fb, ba, bb x (narrow, wide)
            for (int g : bylevel[L]) {
                const int f = P->h_f[g], w = P->h_w[g];
                if (w <= J * kSBlock || f <= kSmallFront) continue;
                const int bs = J * kSBlock, be = std::min(bs + kSBlock, w);
                const int nsub = (be - bs + kPanel - 1) / kPanel;
                const int wide = nsub > 1;
                for (int a = 0; a < nsub; ++a) {
                    t[0][wide].push_back(make_int2(g, a));
                    t[2][wide].push_back(make_int2(g, a));
                }
                for (int rb = 0; rb < (f - be + kTile - 1) / kTile; ++rb) t[1][wide].push_back(make_int2(g, rb));
                for (int cb = 0; cb < bs / kPanel; ++cb) t[3][wide].push_back(make_int2(g, cb));
                for (int a = 1; a < nsub; ++a)
                    for (int b = 0; b < a; ++b) inv_d[a - b - 1].push_back(make_int4(g, J, a, b));
            }
            std::vector<int2> *dst[4] = {&fa, &fb, &ba, &bb};
            std::vector<int> *ptr[4] = {&pfa, &pfb, &pba, &pbb};
            for (int k = 0; k < 4; ++k)
                for (int c = 0; c < 2; ++c) {
                    dst[k]->insert(dst[k]->end(), t[k][c].begin(), t[k][c].end());
                    ptr[k]->push_back((int)dst[k]->size());
                }
        }
        P->lvl_fa_ptr.push_back(pfa);
        P->lvl_fb_ptr.push_back(pfb);
        P->lvl_ba_ptr.push_back(pba);
        P->lvl_bb_ptr.push_back(pbb);
        P->lvl_foot_ptr.push_back((int)foot.size());
        P->lvl_footr_ptr.push_back((int)footr.size());
        P->lvl_node_ptr.push_back((int)level_nodes.size());
        P->lvl_init_ptr.push_back((int)init.size());
        for (int c = 0; c < 4; ++c) {
            small.insert(small.end(), small_cls[c].begin(), small_cls[c].end());
            P->lvl_small_ptr.push_back((int)small.size());
        }
        P->lvl_pull_ptr.push_back((int)pull.size());
        P->lvl_asm_ptr.push_back((int64_t)asm_pos.size());
        P->lvl_npanels.push_back(npan);
        P->lvl_pull_elems.push_back(pe_elems);
        {
            double fac = 0, sf = 0, sw = 0;
            for (int g : bylevel[L]) {
                const double f = P->h_f[g], w = P->h_w[g];
                fac += f * w - w * (w - 1) / 2;
                sf += f;
                sw += w;
            }
            P->lvl_solve_fac.push_back(fac);
            P->lvl_solve_f.push_back(sf);
            P->lvl_solve_w.push_back(sw);
        }
        P->lvl_pull_index.push_back(pe_index);
        std::vector<int> pp{(int)panel.size()}, tp{(int)ptask.size()}, ub64, ub128;
        std::vector<char> bey;
        std::vector<double> ufl(npan, 0.0), uel(npan, 0.0), pel(npan, 0.0);
        for (int q = 0; q < npan; ++q) {
            for (int g : bylevel[L]) {
                const int f = P->h_f[g], w = P->h_w[g];
                if (w <= q * kPanel || f <= kSmallFactor) continue;
                const int pe = std::min(q * kPanel + kPanel, w);
                pel[q] += (double)(f - pe) * (pe - q * kPanel);
                const int nrb = (f - pe + kTile - 1) / kTile;
                // few row blocks: each repeats the diagonal factorization in one launch; many:
                // the diagonal block alone, then the row blocks in a k_panel_rows launch
                if (nrb == 0 || nrb > kPanelFusedRows) ptask.push_back(make_int2(g, -1));
                for (int rb = 0; rb < nrb; ++rb)
                    (nrb > kPanelFusedRows ? panel : ptask).push_back(make_int2(g, rb));
            }
            pp.push_back((int)panel.size());
            tp.push_back((int)ptask.size());
            // Schur tasks of this panel, grouped by front-size bucket, in both tile shapes, in two
            // parts: part 0 must precede the next panel of its front (inner steps, and the next
            // 256-column block of a middle or outer step); part 1 is the rest of a middle or outer
            // step, which the next block's panels do not touch (run_factor may overlap it)
            char beyond = 0;  // some step writes columns past its own 256-column block
            for (int g : bylevel[L]) {
                const int f = P->h_f[g], w = P->h_w[g];
                if (w <= q * kPanel || f <= kSmallFactor) continue;
                const SchurStep st = schur_step(f, w, q);
                if (st.any && !st.inner) beyond = 1;
            }
            bey.push_back(beyond);
            // one front's tiles of one part in one shape, (front, first row, first column):
            // columns in ascending order, rows down from the diagonal.  Target columns before w
            // are tiled from the step's base, those from w on (the contribution, outer steps)
            // from w, so that every wave's columns lie in one region and one 64-column block
            // of the contribution (FrontCols::run)
            auto push_tiles = [&](int g, const SchurStep &st, int part, bool big) {
                const int f = P->h_f[g], w = P->h_w[g];
                const int tile = big ? 128 : kTile, next_tiles = kBlock / tile;
                const bool has_next = !st.inner && st.base < w;
                int count = 0;
                for (int fam = 0; fam < 2; ++fam) {
                    const int c0 = fam ? std::max(st.base, w) : st.base, c1 = fam ? st.cend : std::min(st.cend, w);
                    for (int j0 = c0, tj = 0; j0 < c1; j0 += tile, ++tj) {
                        const bool first = st.inner || (fam == 0 && has_next && tj < next_tiles);
                        if (first != (part == 0)) continue;
                        for (int i0 = j0; i0 < f; i0 += tile) {
                            (big ? upd128 : upd).push_back(make_int4(g, i0, j0, big));
                            ++count;
                        }
                    }
                }
                return count;
            };
            auto steps_of = [&](auto fn) {  // fn(g, step) for the fronts with Schur work after panel q
                for (int g : bylevel[L]) {
                    const int f = P->h_f[g], w = P->h_w[g];
                    if (w <= q * kPanel || f <= kSmallFactor) continue;
                    const SchurStep st = schur_step(f, w, q);
                    if (st.any) fn(g, st);
                }
            };
            for (int part = 0; part < 2; ++part) {
                // 64x64: inner and middle steps, then the outer steps bucket by bucket
                ub64.push_back((int)upd.size());
                steps_of([&](int g, const SchurStep &st) {
                    if (st.inner || st.middle) push_tiles(g, st, part, false);
                });
                for (int b = 0; b < kSchurBuckets; ++b) {
                    ub64.push_back((int)upd.size());
                    steps_of([&](int g, const SchurStep &st) {
                        if (!st.inner && !st.middle && schur_bucket(P->h_f[g]) == b)
                            P->bucket_tasks[b] += push_tiles(g, st, part, false);
                    });
                }
                // 128x128: the outer steps bucket by bucket
                for (int b = 0; b < kSchurBuckets; ++b) {
                    ub128.push_back((int)upd128.size());
                    steps_of([&](int g, const SchurStep &st) {
                        if (!st.inner && !st.middle && schur_bucket(P->h_f[g]) == b) push_tiles(g, st, part, true);
                    });
                }
            }
            ub64.push_back((int)upd.size());
            ub128.push_back((int)upd128.size());
            steps_of([&](int g, const SchurStep &st) {
                const int f = P->h_f[g];
                double elems = 0;
                for (int j = st.base; j < st.cend; ++j) elems += f - j;
                ufl[q] += 2.0 * (st.ke - st.kb) * elems;
                uel[q] += 2.0 * elems + (double)(st.ke - st.kb) * (f - st.base);
            });
        }
        P->lvl_panel_ptr.push_back(pp);
        P->lvl_ub64.push_back(ub64);
        P->lvl_ub128.push_back(ub128);
        P->lvl_beyond.push_back(bey);
        P->lvl_ptask_ptr.push_back(tp);
        P->lvl_upd_flops.push_back(ufl);
        P->lvl_upd_elems.push_back(uel);
        P->lvl_panel_elems.push_back(pel);
    }

    // system of every front (pivot scale and residual norms are per system)
    P->h_sys_begin = S.sys_begin.empty() ? std::vector<int>{0, n} : S.sys_begin;
    P->nsys = (int)P->h_sys_begin.size() - 1;
    std::vector<int> fsys(nf, 0);
    for (int g = 0; g < nf; ++g)
        fsys[g] = (int)(std::upper_bound(P->h_sys_begin.begin(), P->h_sys_begin.end(), P->h_first[g]) -
                        P->h_sys_begin.begin()) - 1;

    if (!upload) return P;
    P->front_sys.upload(fsys);
    P->sys_begin.upload(P->h_sys_begin);
    {
        std::vector<int> rsys(std::max(n, 1), 0);
        for (int t = 0; t < P->nsys; ++t)
            for (int i = P->h_sys_begin[t]; i < P->h_sys_begin[t + 1]; ++i) rsys[i] = t;
        P->row_sys.upload(rsys);
    }
    P->f.upload(P->h_f);
    P->w.upload(P->h_w);
    P->first.upload(P->h_first);
    P->child_ptr.upload(fchild_ptr);
    P->children.upload(fchildren);
    P->pat_ptr.upload(P->h_pat_ptr);
    P->pat_rows.upload(P->h_pat_rows);
    P->level_nodes.upload(level_nodes);
    P->init_tasks.upload(init);
    P->small_nodes.upload(small);
    {
        std::vector<SmallRec> recs(small.size());
        for (size_t k = 0; k < small.size(); ++k) {
            const int g = small[k], c0 = fchild_ptr[g], c1 = fchild_ptr[g + 1];
            recs[k] = SmallRec{P->h_f[g], P->h_w[g], P->h_first[g], fsys[g], P->h_off[g], vbase[g],
                               c0, c1, P->h_pat_ptr[g], 0, c1 > c0 ? inv_off[fchildren[c0]] : 0, P->h_coff[g]};
        }
        P->small_recs.upload(recs);
        std::vector<SmallChild> crec(std::max<size_t>(fchildren.size(), 1));
        for (size_t c = 0; c < fchildren.size(); ++c) {
            const int x = fchildren[c];
            crec[c] = SmallChild{vbase[x], P->h_off[x], P->h_coff[x], P->h_f[x], P->h_w[x]};
        }
        P->child_recs.upload(crec);
    }
    P->foot_tasks.upload(foot);
    P->footr_tasks.upload(footr);

    auto pack = [&](const std::vector<int2> &v) {
        std::vector<PanelTask> o(v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            const int g = v[i].x;
            o[i] = PanelTask{g, v[i].y, P->h_f[g], P->h_w[g], P->h_off[g], vbase[g], P->h_first[g], fsys[g]};
        }
        return o;
    };
    P->fa_tasks.upload(pack(fa));
    P->fb_tasks.upload(pack(fb));
    P->ba_tasks.upload(pack(ba));
    P->bb_tasks.upload(pack(bb));
    {
        std::vector<int4> all;
        for (int d = 0; d < kSBlock / kPanel - 1; ++d) {
            all.insert(all.end(), inv_d[d].begin(), inv_d[d].end());
            P->inv_ptr[d + 1] = (int)all.size();
        }
        P->inv_tasks.upload(all);
    }
    P->vbase.upload(vbase);
    P->inv_map.upload(inv_map);
    {
        std::vector<uint32_t> inv_col(inv_map.size(), ~0u);
        for (int x = 0; x < nf; ++x) {
            if (fparent[x] < 0) continue;
            const int tx = P->h_f[x] - P->h_w[x], fp = P->h_f[fparent[x]];
            for (int k = 0; k < fp; ++k) {
                const int a = inv_map[inv_off[x] + k];
                if (a >= 0) inv_col[inv_off[x] + k] = (uint32_t)contrib_col(0, tx, a);
            }
        }
        P->inv_col.upload(inv_col);
    }
    P->perm.upload(S.perm);
    P->iperm.upload(S.iperm);
    P->off.upload(P->h_off);
    P->coff.upload(P->h_coff);
    P->inv_off.upload(inv_off);
    P->u_off.upload(uoff);
    P->asm_pos.upload(asm_pos);
    P->asm_vidx.upload(asm_vidx);
    P->pull_tasks.upload(pull);
    auto pack_schur = [&](const std::vector<int4> &v) {
        std::vector<SchurTask> o(v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            const int g = v[i].x;
            o[i] = SchurTask{g, v[i].y, v[i].z, P->h_f[g], P->h_w[g], P->h_first[g], P->h_off[g], P->h_coff[g]};
        }
        return o;
    };
    P->upd_tasks.upload(pack_schur(upd));
    P->upd128_tasks.upload(pack_schur(upd128));
    P->panel_tasks.upload(panel);
    P->ptasks.upload(ptask);
    return P;
}

std::shared_ptr<DevicePlan> build_plan(const Symbolic &S) { return build_plan_host(S, true); }

void ensure_plan(Symbolic &S) {
    std::lock_guard<std::mutex> lk(*S.plan_mu);
    if (!S.plan) S.plan = build_plan(S);
}

// Host-only summary of the device schedule: fronts, levels, arena elements, largest front,
// widest front, kernel launches per factorization, Schur tiles (64x64) of the lookahead part
// that may overlap the next panels, factor-store elements, contribution-pool elements, largest
// sum of contribution blocks live at once.
void plan_stats(const Symbolic &S, int64_t *out) {
    auto P = build_plan_host(S, false);
    int64_t launches = 0, part1 = 0;
    for (int L = 0; L < P->nlevels; ++L) {
        launches += 2;
        for (int q = 0; q < P->lvl_npanels[L]; ++q) {
            launches += 3;
            const int *ub = &P->lvl_ub64[L][q * kUb64Stride];
            part1 += ub[2 * (kSchurBuckets + 1)] - ub[kSchurBuckets + 1];
        }
    }
    out[0] = P->nfront;
    out[1] = P->nlevels;
    out[2] = P->arena;
    out[3] = P->max_f;
    out[4] = P->max_w;
    out[5] = launches;
    out[6] = part1;
    out[7] = P->fstore;
    out[8] = P->pool;
    out[9] = P->pool_live;
}

void plan_layout(const Symbolic &S, int64_t *out) {
    auto P = build_plan_host(S, false);
    for (int g = 0; g < P->nfront; ++g) {
        int64_t *o = out + 6 * (int64_t)g;
        o[0] = P->h_f[g];
        o[1] = P->h_w[g];
        o[2] = P->h_level[g];
        o[3] = P->h_parent[g];
        o[4] = P->h_off[g];
        o[5] = P->h_coff[g];
    }
}

int64_t front_addr(int64_t off, int64_t coff, int f, int w, int i, int j) { return front_col(off, coff, f, w, j) + i; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Front zero-fill + extend-add as a pull over 64x64 lower tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_assemble_pull(const int4 *__restrict__ tasks, T *__restrict__ arena,
                                                       const int *__restrict__ F, const int *__restrict__ W,
                                                       const int64_t *__restrict__ off,
                                                       const int64_t *__restrict__ coff,
                                                       const int *__restrict__ child_ptr,
                                                       const int *__restrict__ children,
                                                       const int64_t *__restrict__ inv_off,
                                                       const int *__restrict__ inv_map,
                                                       const uint32_t *__restrict__ inv_col) {
    const int4 tk = tasks[blockIdx.x];  // (front, first row, first column)
    const int s = tk.x;
    const int f = F[s];
    const int i = tk.y + (threadIdx.x & 63);
    const int jbase = tk.z + (threadIdx.x >> 6);
    const ColRun dst = FrontCols{off[s], coff[s], f, W[s]}.run(tk.z, tk.z, kTile);
    const int c0 = child_ptr[s], c1 = child_ptr[s + 1];
    if (i >= f) return;
    // columns j = jbase + 4q <= i of row i inside the tile's run; each child's row map entry
    // for i and its 16 column entries are read together, then the 16 update values together
    const int nq = i < jbase ? 0 : min(min(kTile / 4, (i - jbase) / 4 + 1), (dst.jend - jbase + 3) / 4);
    if (nq <= 0) return;
    T acc[kTile / 4];
#pragma unroll
    for (int q = 0; q < kTile / 4; ++q) acc[q] = T(0);
    for (int c = c0; c < c1; ++c) {  // children in order: each sum keeps the child order
        const int ch = children[c];
        const int b = inv_map[inv_off[ch] + i];
        if (b < 0) continue;
        // the child's contribution entry (b, a), b >= a, is U[inv_col entry of a + b]
        const T *U = arena + coff[ch];
        const uint32_t *icol = inv_col + inv_off[ch];
        // unconditional loads, masked afterwards: past the thread's last column the
        // indices repeat that column's (cache hits, no extra traffic)
        uint32_t a[kTile / 4];
#pragma unroll
        for (int q = 0; q < kTile / 4; ++q) a[q] = icol[jbase + 4 * min(q, nq - 1)];
        T u[kTile / 4];
#pragma unroll
        for (int q = 0; q < kTile / 4; ++q) u[q] = U[(a[q] == ~0u ? 0u : a[q]) + (uint32_t)b];
#pragma unroll
        for (int q = 0; q < kTile / 4; ++q) acc[q] = (q < nq && a[q] != ~0u) ? acc[q] + u[q] : acc[q];
    }
#pragma unroll
    for (int q = 0; q < kTile / 4; ++q)
        if (q < nq) arena[dst.at(i, jbase + 4 * q)] = acc[q];
}

// Zero-fill + extend-add of small fronts (f <= FR, FR = 8, 16 or 32) by lane groups, as
// k_assemble_pull does for 64x64 tiles (the same sums in the same child order): lane i builds
// row i of the front's lower triangle, 64 / FR fronts per wave.
template <typename T, int FR>
__global__ __launch_bounds__(256) void k_assemble_small(const SmallRec *__restrict__ recs, int count,
                                                        T *__restrict__ arena, const SmallChild *__restrict__ child,
                                                        const int *__restrict__ inv_map,
                                                        const uint32_t *__restrict__ inv_col) {
    constexpr int GW = 64 / FR;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wave_first = (blockIdx.x * 4 + wv) * GW;
    if (wave_first >= count) return;  // whole wave
    const int i = lane % FR;
    const int idx = wave_first + lane / FR;
    if (idx >= count) return;  // no barriers or wave-level syncs below
    const SmallRec rc = recs[idx];
    const int f = rc.f;
    if (i >= f) return;
    T acc[FR];
#pragma unroll
    for (int j = 0; j < FR; ++j) acc[j] = T(0);
    for (int c = rc.c0; c < rc.c1; ++c) {  // children in order: each sum keeps the child order
        // the children's inverse row maps follow each other, f entries each
        const SmallChild ch = child[c];
        const int64_t m0 = rc.inv0 + (int64_t)(c - rc.c0) * f;
        const int b = inv_map[m0 + i];
        if (b < 0) continue;
        // unconditional loads (indices clamped to the lane's row), masked afterwards; the
        // child's contribution entry (b, a), b >= a, is U[inv_col entry of a + b]
        const T *U = arena + ch.coff;
        uint32_t a[FR];
#pragma unroll
        for (int j = 0; j < FR; ++j) a[j] = inv_col[m0 + min(j, i)];
        T u[FR];
#pragma unroll
        for (int j = 0; j < FR; ++j) u[j] = U[(a[j] == ~0u ? 0u : a[j]) + (uint32_t)b];
#pragma unroll
        for (int j = 0; j < FR; ++j) acc[j] = (j <= i && a[j] != ~0u) ? acc[j] + u[j] : acc[j];
    }
    const FrontCols dst{rc.off, rc.coff, f, rc.w};  // f <= 32: one contribution column block
#pragma unroll
    for (int j = 0; j < FR; ++j)
        if (j <= i) arena[dst.col_small(j) + i] = acc[j];
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

// Panel LDL^T, step 1: the (<=64)^2 diagonal block of the panel is factored in
// registers and LDS (k_panel), giving D and the block's inverse Li.  Pivot rule of
// UtduNumericFactorization.cs:553-577: accept only
// pivot > tol*scale (NaN fails), otherwise regularize to reg*scale and count
// (negative counted separately; bit 2 marks a refused pivot when
// regularization is off); the comparison is made in float64 for both
// precisions.  One barrier per pair of pivots: the column is scaled by the reciprocal
// (:577), the trailing update is a_rc -= (l_r d) l_c, and the inverse
// Li = L11^{-1} is built alongside (Li <- (I - l_k e_k^T) Li).  Li's strict lower
// part is stored transposed in the unused upper triangle of the diagonal block
// (front row ps+m, column ps+k holds Li[k][m]) for the row-block kernel and the
// solves.
// 64x64 tiles staged through registers: thread (lane r, group g = tid / 64) holds
// column c = g + 4j, j < 16, of rows r0 + r.  Loads are unconditional (addresses
// clamped into the tile, which is not empty: nr, nc >= 1) and raw; the caller masks the
// values when it stores them to LDS, so every load of a tile is in flight before the first
// one is consumed (a branch around the loads would make the next loads wait for them).
template <typename T>
__device__ inline void tile_load(T (&v)[kPanel / 4], const T *__restrict__ M, int f, int r0, int c0, int nr, int nc) {
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    const T *base = M + (int64_t)c0 * f + r0 + min(r, nr - 1);
#pragma unroll
    for (int j = 0; j < kPanel / 4; ++j) v[j] = base[(int64_t)min(g + 4 * j, nc - 1) * f];
}
template <typename T>
__device__ inline T lane_value(T v, int lane);
template <>
__device__ inline float lane_value<float>(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <>
__device__ inline double lane_value<double>(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Makes a loaded value opaque to the compiler.  Applied after the last load of a batch is
// issued, it keeps every load of the batch unconditional: a select between a loaded value and
// zero is otherwise turned into a branch around its load, with a wait for that one load.
template <typename T>
__device__ __forceinline__ void pin(T &x) {
    asm volatile("" : "+v"(x));
}

// 1 / x for a pivot: hardware reciprocal refined by Newton steps (one for float,
// two for double), within an ulp of the rounded quotient and without the branchy
// IEEE division sequence on the pivot-to-pivot path.
__device__ __forceinline__ float pivot_recip(float x) {
    float y = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, y, 1.0f);
    return __builtin_fmaf(y, e, y);
}
__device__ __forceinline__ double pivot_recip(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// Pivot rule of UtduNumericFactorization.cs:553-577 for a pivot candidate, compared in
// float64: returns the pivot to use and sets the step's flags.
__device__ __forceinline__ double pivot_rule(double piv, double thr, double regv, int allow_reg, unsigned char &fl) {
    fl = 0;
    if (!(piv > thr)) {
        if (!allow_reg) fl |= 4;
        if (piv < 0.0) fl |= 2;
        fl |= 1;
        piv = regv;
    }
    return piv;
}

// Diagonal block of a panel: the right-looking LDL^T of UtduNumericFactorization.cs:560-600 on a
// (<=64)^2 block, taken in sixteen groups of four pivots by one workgroup of four waves.  Wave v
// holds block columns 16v..16v+15 of the block (a) and of its inverse Li = L11^{-1} (li) as
// four 16x16 matrix-core accumulator tiles: register q of tile t in lane (n, g), n = lane mod 16,
// g = lane / 16, holds row 16t + g + 4q, column 16v + n.  For the group of pivots k..k+3:
//  - the wave owning columns k..k+3 copies them to LDS and reads them back one row per lane,
//    takes the four pivots in order (pivot rule; l_r = a_rK / d_K; a_rc -= (l_r d_K) l_c on the
//    group's later columns) and publishes -(l d), l and -(L Q) for all 64 rows, Q the inverse of
//    the group's unit lower 4x4 block P of L;
//  - after one barrier every wave applies the group on the matrix cores: the rank-4 update
//    a -= (L D) L^T on its columns right of the group, and the four row operations of the
//    eliminations on the inverse, Li <- Li - (L Q) Li[k..k+3, :], which leaves Q Li[k..k+3, :]
//    in the group's rows (M = L Q is I - Q there, the stored l being zero on and above the
//    diagonal) and Li[r, :] - sum_i l_ri Li_new[k+i, :] below them.
// Entries above the diagonal and in eliminated columns receive updates that nothing reads.
// Sixteen barriers; per group four serial pivots in one wave and at most eight products per wave.
constexpr int kDgLd = 80;  // row stride of the published columns: the four operand rows g = 0..3 of
                           // a read start 16 banks apart
constexpr int kDgColLd = 68;  // the owner's copy: its 4 x 4 writing lanes hit distinct banks
constexpr int kDgStage = 68;  // row stride of the block's staging copy: the tile reads (16 columns
                              // x 4 rows per instruction) hit distinct banks
template <typename T>
struct DiagLds {
    T wn[2][4][kDgLd];            // -(l_r d_K), K = k + i, row r (two groups in flight)
    T lp[2][4][kDgLd];            // l_r
    T mn[2][4][kDgLd];            // -(L Q)[r][i]
    T pcol[4][kDgColLd];          // the next group's four columns, as its owner wave reads them
    T rd[kPanel];                 // 1 / d_k for the row block's operand
    T cand[kPanel];               // pivot candidates (the pivot rule's input)
    T piv[kPanel];                // pivots
};

// M index m of a product's first operand -> row offset in its tile, so that accumulator
// register q of lane group g holds row g + 4q for both precisions (the FP32 instruction
// writes M rows 4g + q, the FP64 one g + 4q).
template <typename T>
__device__ __forceinline__ int dg_row(int m) {
    if constexpr (sizeof(T) == 4) return (m >> 2) | ((m & 3) << 2);
    else return m;
}

// The owner wave of group G copies block columns 4G..4G+3 (its lanes n = 4 (G mod 4) + i)
// through S.pcol into c, one row per lane: c[i] = A[r][4G + i].
template <typename T>
__device__ __forceinline__ void dg_extract(DiagLds<T> &S, const typename MfmaAcc<T>::type (&a)[4], int G, T (&c)[4]) {
    const int lane = threadIdx.x & 63, n = lane & 15, g = lane >> 4, i = n - 4 * (G & 3);
    if (i >= 0 && i < 4) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) S.pcol[i][16 * t + g + 4 * q] = a[t][q];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = S.pcol[j][lane];
}

// Pivot rule operands of one front: the float64 threshold and regularized value, and the same
// rule in the working precision (FP32: cand > thr_t, thr_t the largest float <= thr, decides
// exactly as (double)cand > thr).
template <typename T>
struct DgRule {
    double thr, regv;
    T thr_t, regv_t;
    int allow_reg;
    __device__ DgRule(double t, double r, int a) : thr(t), regv(r), regv_t((T)r), allow_reg(a) {
        if constexpr (sizeof(T) == 4) {
            float tf = (float)t;
            if ((double)tf > t) tf = nextafterf(tf, -INFINITY);
            thr_t = tf;
        } else {
            thr_t = t;
        }
    }
};

// The owner wave of group G takes pivots k = 4G .. k+3 (lane r: row r of the four columns in
// c) and publishes the group's operands into buffer b.  A pivot past nb (identity padding) has
// candidate 1 and a zero column: whatever the rule makes of it multiplies zeros.
// The group's 4x4 diagonal block is factored first on wave-uniform values (the pivot-to-pivot
// chain); the rows follow, l_r = a_rK / d_K and a_rc -= (l_r d_K) l_c, with the same operations
// in the same order as the uniform values of rows k..k+3.
template <typename T>
__device__ __forceinline__ void dg_factor_group(DiagLds<T> &S, int G, int b, const DgRule<T> &R, T (&c)[4]) {
    const int r = threadIdx.x & 63, k = 4 * G;
    T e[4][4];  // e[i2][i] = A[k + i2][k + i], i2 >= i (uniform)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int i2 = i; i2 < 4; ++i2) e[i2][i] = lane_value<T>(c[i], k + i2);
    T pv[4], pinv[4], p[4][4], cand[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        cand[i] = e[i][i];
        pv[i] = cand[i] > R.thr_t ? cand[i] : R.regv_t;  // padding: candidate 1, column zero
        pinv[i] = pivot_recip(pv[i]);
#pragma unroll
        for (int i2 = i + 1; i2 < 4; ++i2) p[i2][i] = e[i2][i] * pinv[i];
#pragma unroll
        for (int i3 = i + 1; i3 < 4; ++i3) {
            const T w3 = p[i3][i] * pv[i];
#pragma unroll
            for (int i2 = i + 1; i2 <= i3; ++i2) e[i3][i2] -= w3 * p[i2][i];
        }
    }
    T l[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int K = k + i;
        l[i] = r > K ? c[i] * pinv[i] : T(0);
        const T w = l[i] * pv[i];
        S.wn[b][i][r] = -w;
        S.lp[b][i][r] = l[i];
#pragma unroll
        for (int i2 = i + 1; i2 < 4; ++i2) c[i2] -= w * p[i2][i];
    }
    if (r < 4) {  // flags are formed from the candidates at the end (dg_pivots)
        S.cand[k + r] = r == 0 ? cand[0] : (r == 1 ? cand[1] : (r == 2 ? cand[2] : cand[3]));
        S.piv[k + r] = r == 0 ? pv[0] : (r == 1 ? pv[1] : (r == 2 ? pv[2] : pv[3]));
    }
    // Q = P^{-1}: Q[a][e] = -(P[a][e] + sum_{e<m<a} P[a][m] Q[m][e])
    T q[4][4];
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int e2 = a - 1; e2 >= 0; --e2) {
            T s = p[a][e2];
#pragma unroll
            for (int m = e2 + 1; m < a; ++m) s += p[a][m] * q[m][e2];
            q[a][e2] = -s;
        }
    // (L Q)[r][e] = l_e + sum_{a>e} l_a Q[a][e]
#pragma unroll
    for (int e2 = 0; e2 < 4; ++e2) {
        T m = l[e2];
#pragma unroll
        for (int a = e2 + 1; a < 4; ++a) m += l[a] * q[a][e2];
        S.mn[b][e2][r] = -m;
    }
}

// Groups 4 T0 .. 4 T0 + 3 of the diagonal block (the pivots of tile column T0); every wave
// passes the same barriers.  The owner of the next group takes it right after updating its
// tiles and applies its inverse update last.
template <typename T, int T0>
__device__ __forceinline__ void dg_tile_column(DiagLds<T> &S, typename MfmaAcc<T>::type (&a)[4],
                                               typename MfmaAcc<T>::type (&li)[4], int v, int nb, const DgRule<T> &R) {
    const int lane = threadIdx.x & 63, n = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int gp = 0; gp < 4; ++gp) {
        const int G = 4 * T0 + gp, b = G & 1;
        if (4 * G >= nb) break;  // uniform
        if (v > T0 || (v == T0 && gp < 3)) {  // columns right of the group
            const T bl = S.lp[b][g][16 * v + n];
#pragma unroll
            for (int t = T0; t < 4; ++t)
                if (t >= v) a[t] = MfmaAcc<T>::step(S.wn[b][g][16 * t + dg_row<T>(n)], bl, a[t]);
        }
        const int G1 = G + 1;
        if (G1 < kPanel / 4 && 4 * G1 < nb && v == G1 / 4) {
            T c[4];
            dg_extract<T>(S, a, G1, c);
            dg_factor_group<T>(S, G1, b ^ 1, R, c);
        }
        if (v <= T0) {  // inverse columns up to the group's last
            const T bli = gp == 0 ? li[T0][0] : (gp == 1 ? li[T0][1] : (gp == 2 ? li[T0][2] : li[T0][3]));
#pragma unroll
            for (int t = T0; t < 4; ++t) li[t] = MfmaAcc<T>::step(S.mn[b][g][16 * t + dg_row<T>(n)], bli, li[t]);
        }
        __syncthreads();
    }
}

// Factors the panel's diagonal block (nb columns, identity past nb) held in a, forming li.
// Ends with S.cand / S.piv complete.
template <typename T>
__device__ __forceinline__ void dg_factor(DiagLds<T> &S, typename MfmaAcc<T>::type (&a)[4],
                                          typename MfmaAcc<T>::type (&li)[4], int v, int nb, const DgRule<T> &R) {
    if (v == 0) {
        T c[4];
        dg_extract<T>(S, a, 0, c);
        dg_factor_group<T>(S, 0, 0, R, c);
    }
    __syncthreads();
    dg_tile_column<T, 0>(S, a, li, v, nb, R);
    if (nb > 16) dg_tile_column<T, 1>(S, a, li, v, nb, R);
    if (nb > 32) dg_tile_column<T, 2>(S, a, li, v, nb, R);
    if (nb > 48) dg_tile_column<T, 3>(S, a, li, v, nb, R);
}

// Row stride of the scaled inverse Bm in LDS: the MFMA operand reads (4 rows x 16
// consecutive columns per wave) hit distinct banks for FP32.
template <typename T>
constexpr int kBmLd = sizeof(T) == 4 ? kPanel : kPanel + 1;
// FP32: entry (m, k) of Bm sits at column k ^ h(m), h(m) = 16 (m mod 4) + (m / 4) mod 16 (a
// bijection of 0..63): the 64 rows of one column written together hit 64 different banks, and
// the MFMA reads (rows 4 s .. 4 s + 3, 16 consecutive columns) fall on four different 16-bank
// groups.  FP64: plain, stride 65.
template <typename T>
__device__ __forceinline__ int bm_col(int m, int k) {
    if constexpr (sizeof(T) == 4) return k ^ (((m & 3) << 4) | ((m >> 2) & 15));
    else return k;
}

// Row-block product of a panel on the matrix cores: L[r0 + i][ps + k] = sum_m X[i][m] Bm[m][k]
// for i < nr, k < nb, with X the row block's assembled values.  Computed as
// C^T = Bm^T X^T in 16x16x4 tiles: wave v takes rows i = 16 v .. 16 v + 15 (the MFMA columns,
// so loads and stores run down front columns), all 64 k as four 16-row tiles.
// rows_load issues the X operand loads (X[16 v + l16][4 s + lk], clamped) ahead of use.
template <typename T>
__device__ __forceinline__ void rows_load(T (&xb)[kPanel / 4], const T *__restrict__ M, int f, int ps, int nb, int r0,
                                          int nr) {
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4, v = threadIdx.x >> 6;
    const T *src = M + r0 + min(16 * v + l16, max(nr, 1) - 1);
#pragma unroll
    for (int s = 0; s < kPanel / 4; ++s) xb[s] = src[(int64_t)(ps + min(4 * s + lk, nb - 1)) * f];
}
template <typename T>
__device__ __forceinline__ void rows_product(T *__restrict__ M, int f, int ps, int nb, int r0, int nr,
                                             const T (*Bm)[kBmLd<T>], const T (&xb)[kPanel / 4]) {
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4, v = threadIdx.x >> 6;
    if (16 * v >= nr) return;  // wave-uniform: no rows of this wave
    // only the k-steps and output tiles inside nb (uniform): small panels cost little
    const int ns = (nb + 3) / 4, nt = (nb + 15) / 16;
    typename MfmaAcc<T>::type acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[t][q] = T(0);
    if (nb == kPanel) {  // full panel: straight-line, the compiler interleaves reads and MFMAs
#pragma unroll
        for (int s = 0; s < kPanel / 4; ++s)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = MfmaAcc<T>::step(Bm[4 * s + lk][bm_col<T>(4 * s + lk, 16 * t + l16)], xb[s], acc[t]);
    } else {
#pragma unroll
        for (int s = 0; s < kPanel / 4; ++s) {
            if (s >= ns) break;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < nt) acc[t] = MfmaAcc<T>::step(Bm[4 * s + lk][bm_col<T>(4 * s + lk, 16 * t + l16)], xb[s], acc[t]);
        }
    }
    const int i = 16 * v + l16;
    if (i < nr)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = 16 * t + MfmaAcc<T>::row(lk, q);
                if (k < nb) M[(int64_t)(ps + k) * f + r0 + i] = acc[t][q];
            }
}

// Panel LDL^T, step 2: each workgroup turns one 64-row block below the diagonal
// block into L rows: X := X B with B = L11^{-T} D^{-1}, B[m][k] = Li[k][m] / d_k
// (Li from the diagonal block's upper triangle, Li[k][k] = 1).
template <typename T>
__global__ __launch_bounds__(256, 2) void k_panel_rows(const int2 *__restrict__ tasks, T *__restrict__ arena,
                                                    const int *__restrict__ F, const int *__restrict__ W,
                                                    const int64_t *__restrict__ off, const int *__restrict__ firstcol,
                                                    int ps, const double *__restrict__ D) {
    // Bm[m][k]: Li[k][m] / d_k for m < k, 1 / d_k for m = k, else 0 (the MFMA A operand)
    __shared__ T Bm[kPanel][kBmLd<T>];
    const int2 tk = tasks[blockIdx.x];
    const int s = tk.x, rb = tk.y;
    const int f = F[s], w = W[s];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const int r0 = pe + rb * kTile;
    const int nr = min(kTile, f - r0);
    if (nr <= 0) return;
    T *M = arena + off[s];
    const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
    const int fc = firstcol[s] + ps;
    T xb[kPanel / 4];
    rows_load<T>(xb, M, f, ps, nb, r0, nr);
    {
        T ti[kPanel / 4], td[kPanel / 4];
        tile_load<T>(ti, M, f, ps, ps, nb, nb);
#pragma unroll
        for (int j = 0; j < kPanel / 4; ++j) {
            const int k = g + 4 * j;
            td[j] = (T)D[fc + min(k, nb - 1)];
        }
#pragma unroll
        for (int j = 0; j < kPanel / 4; ++j) {
            const int k = g + 4 * j;
            const T rdk = T(1) / td[j];
            Bm[lane][bm_col<T>(lane, k)] = (lane < k && k < nb) ? ti[j] * rdk : (lane == k && k < nb ? rdk : T(0));
        }
    }
    __syncthreads();
    rows_product<T>(M, f, ps, nb, r0, nr, Bm, xb);
}

// Panel LDL^T: every workgroup factors the front's 64x64 diagonal block (dg_factor) and then,
// for a front with at most kPanelFusedRows row blocks below the panel, turns its 64-row block
// into L rows, X := X B with B = L11^{-T} D^{-1}, B[m][k] = Li[k][m] / d_k (as k_panel_rows).
// Those row-block workgroups repeat the diagonal factorization instead of waiting for a launch
// of its own; the row block's loads are in flight meanwhile.  A front with more row blocks gets
// one workgroup here (row block -1) and its row blocks go to k_panel_rows.  The workgroup of row
// block 0 or -1 writes D, the pivot flags and Li.  The diagonal block's lower triangle keeps its
// assembled values: the other workgroups of the launch read them, and nothing later does (the
// row blocks, the Schur updates and the solves use Li and D; the host's factor export rebuilds
// L11 from Li).
template <typename T>
__global__ __launch_bounds__(256, 2) void k_panel(const int2 *__restrict__ tasks, T *__restrict__ arena,
                                               const int *__restrict__ F, const int *__restrict__ W,
                                               const int64_t *__restrict__ off, const int *__restrict__ firstcol,
                                               int ps, const double *__restrict__ scale,
                                               const int *__restrict__ front_sys, double tol, double regf,
                                               int allow_reg, double *__restrict__ D,
                                               unsigned char *__restrict__ flags) {
    __shared__ DiagLds<T> S;
    // The block's staging copy (stg[c][r], wave v's columns 16v..16v+15), later
    // Bm[m][k]: Li[k][m] / d_k for m < k, 1 / d_k for m = k, else 0 (the MFMA A operand)
    __shared__ __align__(16) T Bst[kPanel * kDgStage];
    T (*stg)[kDgStage] = reinterpret_cast<T (*)[kDgStage]>(Bst);
    T (*Bm)[kBmLd<T>] = reinterpret_cast<T (*)[kBmLd<T>]>(Bst);
    const int2 tk = tasks[blockIdx.x];
    const int s = tk.x, rb = tk.y;
    const int f = F[s], w = W[s];
    const int pe = min(ps + kPanel, w), nb = pe - ps;
    const int r0 = pe + max(rb, 0) * kTile, nr = rb < 0 ? 0 : min(kTile, f - r0);
    T *M = arena + off[s];
    const int lane = threadIdx.x & 63, n = lane & 15, g = lane >> 4;
    const int v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T xb[kPanel / 4];
    if (nr > 0) rows_load<T>(xb, M, f, ps, nb, r0, nr);  // in flight during the factorization
    const double sc = scale[front_sys[s]];
    // Block column c = 16v + n, rows 16t + g + 4q: lower triangle, identity past nb, zero above.
    // Wave v loads its columns a column per instruction (lane = row; addresses clamped into the
    // block) and turns them into its tiles through stg; masks applied on the way.
    typename MfmaAcc<T>::type a[4], li[4];
    const int c = 16 * v + n;
    {
        T x[kPanel / 4];
        const T *col0 = M + (int64_t)ps * f + ps + min(lane, nb - 1);
#pragma unroll
        for (int j = 0; j < kPanel / 4; ++j) x[j] = col0[(int64_t)min(16 * v + j, nb - 1) * f];
#pragma unroll
        for (int j = 0; j < kPanel / 4; ++j) stg[16 * v + j][lane] = x[j];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = 16 * t + g + 4 * q;
                const T y = stg[c][rr];
                a[t][q] = (rr < nb && c < nb && rr >= c) ? y : (rr == c ? T(1) : T(0));
                li[t][q] = rr == c ? T(1) : T(0);
            }
    }
    const DgRule<T> R(tol * sc, regf * sc, allow_reg);
    dg_factor<T>(S, a, li, v, nb, R);
    if (rb <= 0) {
        if (threadIdx.x < nb) {
            const int fc = firstcol[s] + ps;
            unsigned char fl;
            (void)pivot_rule((double)S.cand[threadIdx.x], R.thr, R.regv, R.allow_reg, fl);
            D[fc + threadIdx.x] = (double)S.piv[threadIdx.x];
            flags[fc + threadIdx.x] = fl;
        }
        // Li[rr][c] at front row ps + c, column ps + rr
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = 16 * t + g + 4 * q;
                if (c < rr && rr < nb) M[(int64_t)(ps + rr) * f + ps + c] = li[t][q];
            }
    }
    if (nr <= 0) return;  // uniform: the diagonal-block-only workgroup
    if (threadIdx.x < kPanel) S.rd[threadIdx.x] = T(1) / S.piv[min((int)threadIdx.x, nb - 1)];
    __syncthreads();
    // Bm[c][rr] = Li[rr][c] / d_rr
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = 16 * t + g + 4 * q;
            const T rd = S.rd[rr];
            Bm[c][bm_col<T>(c, rr)] = (c < rr && rr < nb) ? li[t][q] * rd : (c == rr && rr < nb ? rd : T(0));
        }
    __syncthreads();
    rows_product<T>(M, f, ps, nb, r0, nr, Bm, xb);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  Block b
// takes task xcd_task(b, n) instead of b, so that every XCD works through one contiguous run
// of the task list: neighbouring tasks (tiles of one front sharing panel rows) then meet in
// the same L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_task(int b, int n) {
    constexpr int kXcd = 8;
    const int x = b % kXcd, r = b / kXcd, q = n / kXcd, rem = n % kXcd;
    return x * q + min(x, rem) + r;
}

// Schur update after panel q (cf. ApplyPanelUpdate, UtduNumericFactorization.cs:624-852):
// F[i,j] -= sum_k (L[i,k] d_k) L[j,k] over the step's K columns, on the lower part
// of the step's target region (schur_step).  One 64x64 output tile per workgroup;
// K streamed through LDS in chunks of 64, products on MFMA (mfma.h).  Operand rows past the
// front's end are not masked: they only reach output rows or columns >= f, which are not
// stored.  Columns past the step's K (a narrow block's last chunk) are zeroed.
template <typename T>
__global__ __launch_bounds__(256, 3) void k_schur(const SchurTask *__restrict__ tasks, T *__restrict__ arena, int q,
                                               const double *__restrict__ D) {
    __shared__ __align__(16) T As[kPanel][kSchurLd<T>];  // As[k][r] = L[i0+r, kc+k] * d_k
    __shared__ __align__(16) T Bs[kPanel][kSchurLd<T>];  // Bs[k][r] = L[j0+r, kc+k]
    __shared__ T Ds[kOuter];                              // d_k of the step's K columns
    const SchurTask tk = tasks[xcd_task(blockIdx.x, gridDim.x)];
    const int f = tk.f;
    const SchurStep sp = schur_step(f, tk.w, q);
    const int kb = sp.kb, ke = sp.ke;
    const int i0 = tk.i0, j0 = tk.j0;
    const T *M = arena + tk.off;  // the step's K columns (pivot columns, factor store)
    const FrontCols fr{tk.off, tk.coff, f, tk.w};  // the target tile's columns
    const double *Dp = D + tk.fc;
    // loader mapping: 16-byte vectors of VEC rows, TPC threads per column, CPI columns per
    // pass (k = cq + CPI j).  The vector start is clamped into the front (reads past a
    // column's end land in the next column or the arena's tail padding).  Byte offsets from
    // a chunk's first column fit 32 bits (f <= 65535), so the loads take a uniform base.
    constexpr int VEC = 16 / sizeof(T), TPC = kTile / VEC, CPI = 256 / TPC, NJ = kPanel / CPI;
    const int qv = threadIdx.x % TPC, cq = threadIdx.x / TPC;
    const uint32_t rowa = min(i0 + VEC * qv, f - 1), rowb = min(j0 + VEC * qv, f - 1);
    for (int t = threadIdx.x; t < ke - kb; t += 256) Ds[t] = (T)Dp[kb + t];
    T ra[NJ][VEC], rb[NJ][VEC];
    // Chunk kc of K into registers; all loads are issued before any is consumed, so the next
    // chunk is in flight while the current one runs on the MFMA units.
    auto fetch = [&](int kc) {
        const char *col = reinterpret_cast<const char *>(M + (int64_t)kc * f);
        const int nk = ke - kc;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const uint32_t oc = (uint32_t)min(cq + CPI * j, nk - 1) * f;
            const mg_u4u va = *reinterpret_cast<const mg_u4u *>(col + (oc + rowa) * sizeof(T));
            const mg_u4u vb = *reinterpret_cast<const mg_u4u *>(col + (oc + rowb) * sizeof(T));
            __builtin_memcpy(ra[j], &va, 16);
            __builtin_memcpy(rb[j], &vb, 16);
        }
    };
    typename MfmaAcc<T>::type acc[2][2];
    schur_acc_zero<T>(acc);
    static_assert(NJ * VEC == 16, "the target tile's 16 values reuse the chunk registers");
    fetch(kb);
    for (int kc = kb; kc < ke; kc += kPanel) {
        const int nk = ke - kc;
        __syncthreads();  // the previous chunk's MFMA reads are done
        typedef T vec_t __attribute__((ext_vector_type(VEC)));
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int k = cq + CPI * j;
            const bool live = nk >= kPanel || k < nk;
            const T d = Ds[kc - kb + k];  // < kOuter; past K only under the mask
            vec_t va, vb;
#pragma unroll
            for (int i = 0; i < VEC; ++i) {
                va[i] = live ? ra[j][i] * d : T(0);
                vb[i] = live ? rb[j][i] : T(0);
            }
            if constexpr (kSchurLd<T> % VEC == 0) {  // aligned rows: one 16-byte store each
                *reinterpret_cast<vec_t *>(&As[k][sch_col<T>(k, VEC * qv)]) = va;
                *reinterpret_cast<vec_t *>(&Bs[k][sch_col<T>(k, VEC * qv)]) = vb;
            } else {
#pragma unroll
                for (int i = 0; i < VEC; ++i) {
                    As[k][sch_col<T>(k, VEC * qv + i)] = va[i];
                    Bs[k][sch_col<T>(k, VEC * qv + i)] = vb[i];
                }
            }
        }
        __syncthreads();
        // last chunk (uniform; kb < ke always): the target tile's values go into the chunk
        // registers, which are free now, and load during the chunk's products
        if (kc + kPanel < ke) fetch(kc + kPanel);
        else schur_load_tile<T>(*reinterpret_cast<T(*)[16]>(&ra[0][0]), arena, fr, i0, j0);
        schur_acc_tile<T>(As, Bs, acc);
    }
    schur_store_tile<T>(acc, *reinterpret_cast<const T(*)[16]>(&ra[0][0]), arena, fr, i0, j0, sp.cend);
}

// The same update with 128x128 output tiles (large fronts).  K goes through two LDS buffers
// in chunks of KC: while chunk c runs on the MFMA units, chunk c+1 (in registers) is staged
// into the other buffer and chunk c+2 is fetched, piece by piece between the k-steps, with
// one barrier per chunk.  Pivots past the step's K are zero and columns past K are clamped
// to its last column, so a partial last chunk adds nothing and no operand is masked.
// Tasks remapped per XCD as in k_schur.
template <typename T>
__global__ __launch_bounds__(256, 2) void k_schur128(const SchurTask *__restrict__ tasks, T *__restrict__ arena, int q,
                                                     const double *__restrict__ D) {
    constexpr int KC = kT128K<T>;
    __shared__ __align__(16) T As[2][KC][kT128Pad];  // As[.][k][r] = L[i0+r, kc+k] * d_k
    __shared__ __align__(16) T Bs[2][KC][kT128Pad];  // Bs[.][k][r] = L[j0+r, kc+k]
    __shared__ T Ds[kOuter];                         // d_k of the step's K columns, zero past K
    const SchurTask tk = tasks[xcd_task(blockIdx.x, gridDim.x)];
    const int f = tk.f;
    const SchurStep sp = schur_step(f, tk.w, q);
    const int kb = sp.kb, ke = sp.ke;
    const int i0 = tk.i0, j0 = tk.j0;
    const T *M = arena + tk.off;  // the step's K columns (pivot columns, factor store)
    const FrontCols fr{tk.off, tk.coff, f, tk.w};  // the target tile's columns
    const double *Dp = D + tk.fc;
    for (int t = threadIdx.x; t < kOuter; t += 256) Ds[t] = t < ke - kb ? (T)Dp[kb + t] : T(0);
    // loader mapping: 16-byte vectors of VEC rows, TPC threads per column, CPI columns per
    // piece (k = cq + CPI j, NJ pieces per chunk); vector starts clamped into the front (rows
    // past f only reach outputs that are not stored).  Byte offsets from a chunk's first
    // column fit 32 bits (f <= 65535).
    constexpr int VEC = 16 / sizeof(T), TPC = 128 / VEC, CPI = 256 / TPC, NJ = KC / CPI;
    constexpr int kSteps = KC / 4;
    static_assert(NJ <= kSteps, "one staged piece per k-step at most");
    const int qv = threadIdx.x % TPC, cq = threadIdx.x / TPC;
    const uint32_t rowa = min(i0 + VEC * qv, f - 1), rowb = min(j0 + VEC * qv, f - 1);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    T ra[NJ][VEC], rb[NJ][VEC];
    auto fetch = [&](int kc, int j) {
        const char *col = reinterpret_cast<const char *>(M + (int64_t)kc * f);
        const uint32_t oc = (uint32_t)min(cq + CPI * j, ke - 1 - kc) * f;
        const mg_u4u va = *reinterpret_cast<const mg_u4u *>(col + (oc + rowa) * sizeof(T));
        const mg_u4u vb = *reinterpret_cast<const mg_u4u *>(col + (oc + rowb) * sizeof(T));
        __builtin_memcpy(ra[j], &va, 16);
        __builtin_memcpy(rb[j], &vb, 16);
    };
    auto stage = [&](int kc, int buf, int j) {  // kc - kb + k < kOuter
        const int k = cq + CPI * j;
        const T d = Ds[kc - kb + k];
        vec_t va, vb;
#pragma unroll
        for (int i = 0; i < VEC; ++i) {
            va[i] = ra[j][i] * d;
            vb[i] = rb[j][i];
        }
        *reinterpret_cast<vec_t *>(&As[buf][k][VEC * qv]) = va;
        *reinterpret_cast<vec_t *>(&Bs[buf][k][VEC * qv]) = vb;
    };
    typename MfmaAcc<T>::type acc[4][4];
    schur128_acc_zero<T>(acc);
    constexpr int NA = sizeof(T) == 4 ? 4 : 1;  // row blocks of old values held at once
    T old[NA * 16];
    // products of chunk kc (buffer buf); STAGE: chunk kc+KC into the other buffer; FETCH:
    // chunk kc+2KC into the registers
    auto chunk = [&](int kc, int buf, auto stage_t, auto fetch_t) {
        constexpr bool STAGE = decltype(stage_t)::value, FETCH = decltype(fetch_t)::value;
        T colv[2][4], rowv[2][4];
        schur128_read_ops<T>(As[buf], Bs[buf], 0, colv[0], rowv[0]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            const int cur = s & 1;
            if (s + 1 < kSteps) schur128_read_ops<T>(As[buf], Bs[buf], 4 * (s + 1), colv[cur ^ 1], rowv[cur ^ 1]);
            schur128_step<T>(colv[cur], rowv[cur], acc);
            if constexpr (STAGE) {
                if (s < NJ) {
                    stage(kc + KC, buf ^ 1, s);
                    if constexpr (FETCH) fetch(kc + 2 * KC, s);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // chunk kb into buffer 0, chunk kb+KC into the registers (or, for one chunk, the old values)
#pragma unroll
    for (int j = 0; j < NJ; ++j) fetch(kb, j);
    __syncthreads();  // Ds complete
#pragma unroll
    for (int j = 0; j < NJ; ++j) stage(kb, 0, j);
    int kc = kb, buf = 0;
    if (kb + KC < ke) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) fetch(kb + KC, j);
    } else {
        schur128_load_old<T, NA, 0>(old, arena, fr, i0, j0);
    }
    __syncthreads();
    for (; kc + 2 * KC < ke; kc += KC, buf ^= 1) {
        chunk(kc, buf, std::true_type{}, std::true_type{});
        __syncthreads();
    }
    if (kc + KC < ke) {
        chunk(kc, buf, std::true_type{}, std::false_type{});
        schur128_load_old<T, NA, 0>(old, arena, fr, i0, j0);
        __syncthreads();
        kc += KC;
        buf ^= 1;
    }
    chunk(kc, buf, std::false_type{}, std::false_type{});
    schur128_store_old<T, NA, 0>(acc, old, arena, fr, i0, j0, sp.cend);
    if constexpr (NA == 1) {
        schur128_load_old<T, NA, 1>(old, arena, fr, i0, j0);
        schur128_store_old<T, NA, 1>(acc, old, arena, fr, i0, j0, sp.cend);
        schur128_load_old<T, NA, 2>(old, arena, fr, i0, j0);
        schur128_store_old<T, NA, 2>(acc, old, arena, fr, i0, j0, sp.cend);
        schur128_load_old<T, NA, 3>(old, arena, fr, i0, j0);
        schur128_store_old<T, NA, 3>(acc, old, arena, fr, i0, j0, sp.cend);
    }
}

// Deterministic statistics over D and the pivot flags: kStatParts workgroups reduce fixed
// strided slices (counts of regularized / negative / failed pivots, min and max |d|), then
// one workgroup combines the partials in a fixed order.
constexpr int kStatParts = 256;

__device__ inline void stat_reduce(double (*v)[256], int t) {
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            for (int q = 0; q < 3; ++q) v[q][t] += v[q][t + h];
            v[3][t] = fmin(v[3][t], v[3][t + h]);
            v[4][t] = fmax(v[4][t], v[4][t + h]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_pivot_stats_part(const double *__restrict__ D,
                                                          const unsigned char *__restrict__ flags, int n,
                                                          double *__restrict__ part) {
    __shared__ double v[5][256];
    const int t = threadIdx.x;
    int reg = 0, neg = 0, bad = 0;
    double lo = 1.7976931348623157e308, hi = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < n; i += (int64_t)kStatParts * 256) {
        const unsigned char fl = flags[i];
        reg += fl & 1;
        neg += (fl >> 1) & 1;
        bad += (fl >> 2) & 1;
        const double m = fabs(D[i]);
        lo = fmin(lo, m);
        hi = fmax(hi, m);
    }
    v[0][t] = reg;
    v[1][t] = neg;
    v[2][t] = bad;
    v[3][t] = lo;
    v[4][t] = hi;
    __syncthreads();
    stat_reduce(v, t);
    if (t < 5) part[(int64_t)t * kStatParts + blockIdx.x] = v[t][0];
}

__global__ __launch_bounds__(256) void k_pivot_stats_final(const double *__restrict__ part, int *__restrict__ counts,
                                                           double *__restrict__ mags) {
    static_assert(kStatParts == 256, "one partial per thread");
    __shared__ double v[5][256];
    const int t = threadIdx.x;
    for (int q = 0; q < 5; ++q) v[q][t] = part[q * kStatParts + t];
    __syncthreads();
    stat_reduce(v, t);
    if (t == 0) {
        counts[0] = (int)v[0][0];
        counts[1] = (int)v[1][0];
        counts[2] = (int)v[2][0];
        mags[0] = v[3][0];
        mags[1] = v[4][0];
    }
}

// ---- solves ----------------------------------------------------------------

// Y[i] = (T) B[perm[i]] (elimination order): a thread per row i, its columns eight at a time
// (the eight gathered loads issued together, the stores coalesced across the wave).  Columns at
// or past the count of the row's system (cols, per system; rows of a system are contiguous and
// permuted among themselves) are left out: no solve kernel reads them.
template <typename T>
__global__ __launch_bounds__(256) void k_permute_in(const double *__restrict__ B, int ldb, const int *__restrict__ perm,
                                                    int n, int nrhs, T *__restrict__ Y, const int *__restrict__ row_sys,
                                                    const int *__restrict__ cols) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int nc = cols ? min(nrhs, cols[row_sys[i]]) : nrhs;
    const double *src = B + perm[i];
    for (int r0 = 0; r0 < nc; r0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[(int64_t)min(r0 + u, nc - 1) * ldb];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (r0 + u < nc) Y[(int64_t)(r0 + u) * n + i] = (T)v[u];
    }
}

// X[j] (+)= Xe[iperm[j]]: a thread per row j in the caller's order, its columns eight at a time
// (the FP64 solution read and written in order, the gather on the elimination-order vector).  A
// column past the row's system count has a zero right-hand side (a first solve: the solution is
// zero) or a zero residual (a correction: nothing to add).
template <typename T>
__global__ __launch_bounds__(256) void k_permute_out(const T *__restrict__ Xe, const int *__restrict__ iperm, int n,
                                                     int nrhs, double *__restrict__ X, int ldx, int accumulate,
                                                     const int *__restrict__ row_sys, const int *__restrict__ cols) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int nc = cols ? min(nrhs, cols[row_sys[j]]) : nrhs;
    const T *src = Xe + iperm[j];
    double *dst = X + j;
    if (!accumulate)
        for (int r = nc; r < nrhs; ++r) dst[(int64_t)r * ldx] = 0.0;
    for (int r0 = 0; r0 < nc; r0 += 8) {
        T v[8];
        double o[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int r = min(r0 + u, nc - 1);
            v[u] = src[(int64_t)r * n];
            o[u] = accumulate ? dst[(int64_t)r * ldx] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (r0 + u < nc) dst[(int64_t)(r0 + u) * ldx] = o[u] + (double)v[u];
    }
}

// Supernodal solves, level by level over the fronts
// (cf. ForwardSubstitution / BackwardSubstitution, UtduNumericFactorization.cs:139-207).
// Every front owns a vector V (f rows x nrhs) in a workspace: the forward pass
// gathers it (own rows from the permuted right-hand side, plus the children's
// footer vectors through the inverse row maps), then walks the own columns in
// kSBlock-column blocks, two launches each (k_fwd_diag: y of the block from its stored
// inverse; k_fwd_upd: the product with the rows below); the footer left in V is the
// update the parent pulls.  The backward pass starts each level with the
// footer product (two-stage: k_bwd_foot_part / k_bwd_foot_red, including the
// division by D) and walks the blocks from the last one down, right-looking
// (k_bwd_diag, k_bwd_upd).  Y holds y after the
// forward pass and x after the backward pass (elimination order, n x nrhs).

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_init(const int2 *__restrict__ tasks, const int *__restrict__ F,
                                                  const int *__restrict__ W, const int *__restrict__ firstcol,
                                                  const int *__restrict__ child_ptr, const int *__restrict__ children,
                                                  const SmallChild *__restrict__ child,
                                                  const int64_t *__restrict__ inv_off, const int *__restrict__ inv_map,
                                                  const int64_t *__restrict__ vbase, T *__restrict__ V,
                                                  const T *__restrict__ Y, int n, int nrhs,
        const int *__restrict__ fsys, const int *__restrict__ cols) {
    const int2 tk = tasks[blockIdx.x];
    const int nc = cols ? min(nrhs, cols[fsys[tk.x]]) : nrhs;  // right-hand sides the system needs
    if (nc == 0) return;
    const int g = tk.x, f = F[g], w = W[g];
    const int i = tk.y * 256 + threadIdx.x;
    if (i >= f) return;
    const int c0 = child_ptr[g], c1 = child_ptr[g + 1];
    // the children's inverse row maps follow each other, f entries each
    const int64_t inv0 = c1 > c0 ? inv_off[children[c0]] : 0;
    T *Vg = V + vbase[g] * nrhs;
    constexpr int R = 8;  // right-hand sides per pass: their loads are issued together
    const T *yrow = Y + firstcol[g] + min(i, max(w - 1, 0));
    for (int r0 = 0; r0 < nc; r0 += R) {
        const int nr = min(R, nc - r0);
        T acc[R];
#pragma unroll
        for (int q = 0; q < R; ++q) acc[q] = yrow[(int64_t)(r0 + min(q, nr - 1)) * n];
#pragma unroll
        for (int q = 0; q < R; ++q) pin(acc[q]);
#pragma unroll
        for (int q = 0; q < R; ++q) acc[q] = i < w ? acc[q] : T(0);
        for (int c = c0; c < c1; ++c) {  // children in order: each sum keeps the child order
            const SmallChild ch = child[c];
            const int a = inv_map[inv0 + (int64_t)(c - c0) * f + i];
            // a child's footer row w_x + max(a, 0) < f_x (children have footer rows)
            const T *src = V + ch.vb * nrhs + ch.w + max(a, 0);
            T u[R];
#pragma unroll
            for (int q = 0; q < R; ++q) u[q] = src[(int64_t)(r0 + min(q, nr - 1)) * ch.f];
#pragma unroll
            for (int q = 0; q < R; ++q) pin(u[q]);
#pragma unroll
            for (int q = 0; q < R; ++q) acc[q] = a >= 0 ? acc[q] + u[q] : acc[q];
        }
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (q < nr) Vg[(int64_t)(r0 + q) * f + i] = acc[q];
    }
}

// Small fronts (f <= 64, hence one panel).  A wave holds 64 / FR fronts of one size class
// (FR = 8, 16, 32 or 64 >= f): lane l works on row i = l % FR of front l / FR, for every
// right-hand side.  Each lane keeps the FR front entries its row needs in registers, loaded
// once per front.  Right-hand sides go in batches of KP: all loads of a batch are issued
// together, and the vector values the rows multiply are broadcast through LDS (one slot per
// lane and right-hand side of the batch, broadcast reads, two wave-level syncs per batch).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum_{j < FR} R[j] * b[j], b 16-byte aligned in LDS
template <typename T, int FR>
__device__ __forceinline__ T lds_dot(const T (&R)[FR], const T *b) {
    constexpr int VW = 16 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(VW)));
    T acc0 = T(0), acc1 = T(0);
#pragma unroll
    for (int j = 0; j < FR; j += VW) {
        const vec q = *reinterpret_cast<const vec *>(b + j);
#pragma unroll
        for (int u = 0; u < VW; ++u) {
            if ((j / VW + u) & 1) acc1 += R[j + u] * q[u];
            else acc0 += R[j + u] * q[u];
        }
    }
    return acc0 + acc1;
}

template <typename T, int FR>
struct SmallCfg {
    static constexpr int GW = 64 / FR;                             // fronts per wave
    static constexpr int KP = (sizeof(T) == 8 && FR == 64) ? 4 : 8;  // right-hand sides per batch
};

// Small fronts (f <= FR, FR = 8, 16 or 32): the front's whole partial factorization in one
// FR-lane group of a wave (64 / FR fronts of one level and size class per wave).  Lane i
// holds row i of the front's lower triangle and row i of Li in registers.  Pivot k
// (UtduNumericFactorization.cs:541-601; the pivot rule in float64, as in k_panel) comes
// from lane k by a shuffle; the scaled column and row k of Li are broadcast through LDS
// (wave-level syncs, no barriers); each lane applies the rank-1 update to its own row.  After
// w pivots a row holds L (columns < w) and the Schur complement (columns >= w), as the panel
// and Schur kernels leave them: stores are L below the diagonal, Li transposed above it,
// the contribution block, D and the pivot flags.
template <typename T, int FR>
__global__ __launch_bounds__(256) void k_factor_small(const int *__restrict__ nodes, int count, T *__restrict__ arena,
                                                      const int *__restrict__ F, const int *__restrict__ W,
                                                      const int64_t *__restrict__ off,
                                                      const int64_t *__restrict__ coff,
                                                      const int *__restrict__ firstcol, const double *__restrict__ scale,
                                                      const int *__restrict__ front_sys, double tol, double regf,
                                                      int allow_reg, double *__restrict__ D,
                                                      unsigned char *__restrict__ flags) {
    constexpr int GW = 64 / FR, VW = 16 / sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    __shared__ __align__(16) T colb[4][64];  // per wave: the scaled pivot column, lane slots
    __shared__ __align__(16) T rowb[4][64];  // per wave: row k of Li of each front
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wave_first = (blockIdx.x * 4 + wv) * GW;
    if (wave_first >= count) return;  // whole wave
    const int i = lane % FR, base = lane - i;
    const int idx = wave_first + lane / FR;
    const bool live = idx < count;
    const int g = nodes[live ? idx : count - 1];
    const int f = live ? F[g] : 0, w = live ? W[g] : 0;
    T *M = arena + off[g];
    const FrontCols fr{off[g], coff[g], f, w};  // f <= 32: one contribution column block
    const double sc = scale[front_sys[g]];
    const double thr = tol * sc, regv = regf * sc;
    // row i of the lower triangle, A[i][j] = arena[fr.col(j) + i] for j <= i (clamped loads,
    // masked)
    const int last = max(f - 1, 0), ri = min(i, last);
    T a[FR], li[FR];
#pragma unroll
    for (int j = 0; j < FR; ++j) a[j] = arena[fr.col_small(min(j, ri)) + ri];
#pragma unroll
    for (int j = 0; j < FR; ++j) pin(a[j]);  // all loads issued before the masks
#pragma unroll
    for (int j = 0; j < FR; ++j) {
        a[j] = (i < f && j <= i) ? a[j] : T(0);
        li[j] = i == j ? T(1) : T(0);
    }
    T my_d = T(1);
    unsigned char my_fl = 0;
    T *cb = &colb[wv][base];
    T *rbf = &rowb[wv][base];
#pragma unroll
    for (int k = 0; k < FR; ++k) {
        if (!__any(k < w)) break;  // wave-uniform: every front of the wave is done
        const T akk = __shfl(a[k], base + k, 64);
        double piv = (double)akk;
        unsigned char fl = 0;
        if (!(piv > thr)) {
            if (!allow_reg) fl |= 4;
            if (piv < 0.0) fl |= 2;
            fl |= 1;
            piv = regv;
        }
        const T pv = (T)piv, pinv = pivot_recip(pv);
        if (i == k) {
            my_d = pv;
            my_fl = fl;
        }
        const bool below = k < w && i > k && i < f;
        const T l = below ? a[k] * pinv : T(0);
        a[k] = below ? l : a[k];
        cb[i] = l;
        if (i == k) {
#pragma unroll
            for (int j = 0; j < FR; j += VW) {
                vec_t v;
#pragma unroll
                for (int e = 0; e < VW; ++e) v[e] = li[j + e];
                *reinterpret_cast<vec_t *>(rbf + j) = v;
            }
        }
        wave_sync();
        const T t = l * pv;
#pragma unroll
        for (int j = k + 1; j < FR; j += 1) a[j] -= t * cb[j];
#pragma unroll
        for (int j = 0; j <= k; ++j) li[j] -= l * rbf[j];
        wave_sync();
    }
    if (live && i < f) {
        const int fc = firstcol[g];
#pragma unroll
        for (int k = 0; k < FR; ++k)
            if (k < w && k < i) M[(int64_t)k * f + i] = a[k];  // L below the diagonal
        if (i < w) {
#pragma unroll
            for (int k = 0; k < FR; ++k)
                if (k < i) M[(int64_t)i * f + k] = li[k];  // Li[i][k] at row k, column i
            D[fc + i] = (double)my_d;
            flags[fc + i] = my_fl;
        }
#pragma unroll
        for (int j = 0; j < FR; ++j)
            if (j >= w && j <= i) arena[fr.col_small(j) + i] = a[j];  // contribution block
    }
}

struct SmallFwdArgs {
    const SmallChild *child;   // per children-array slot
    const int *inv_map;        // a front's children's inverse row maps, consecutive, f entries each
    int n, nrhs;
    const int *cols;           // per system: leading right-hand sides still to solve (NULL: all)
};


// v[p] = own right-hand side (own rows) plus the footer contributions of child slots q < NQ,
// for columns r0 + p: straight-line, so every load of the chunk is in flight at once.
template <typename T, int KP, int NQ>
__device__ __forceinline__ void fwd_gather(T (&v)[KP], const T *__restrict__ Y, const T *__restrict__ V, int r0,
                                           int nc, int i, int f, bool own, int n, int yi, const int64_t (&cb)[4],
                                           const int (&cf)[4], int64_t vself) {
    T yl[KP], cl[KP][NQ > 0 ? NQ : 1];
#pragma unroll
    for (int p = 0; p < KP; ++p) {
        const int r = r0 + p;
        const int64_t rr = (r < nc && i < f) ? r : 0;
        yl[p] = Y[rr * n + yi];
#pragma unroll
        for (int q = 0; q < NQ; ++q) cl[p][q] = V[cb[q] >= 0 ? cb[q] + rr * cf[q] : vself + rr * f];
    }
#pragma unroll
    for (int p = 0; p < KP; ++p) {
        pin(yl[p]);
#pragma unroll
        for (int q = 0; q < NQ; ++q) pin(cl[p][q]);
    }
#pragma unroll
    for (int p = 0; p < KP; ++p) {
        const bool act = r0 + p < nc && i < f;
        T acc = (act && own) ? yl[p] : T(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc += (act && cb[q] >= 0) ? cl[p][q] : T(0);
        v[p] = acc;
    }
}

// Forward: v = own right-hand-side rows plus the children's footer contributions (through
// the inverse row maps); y = Li v on the own rows (Li[i][m], m < i, stored at row m,
// column i; unit diagonal); v -= L y on the footer rows.
template <typename T, int FR>
__global__ __launch_bounds__(256) void k_fwd_small(const SmallRec *__restrict__ recs, int count, SmallFwdArgs a,
                                                   const T *__restrict__ arena, T *__restrict__ V,
                                                   T *__restrict__ Y) {
    constexpr int GW = SmallCfg<T, FR>::GW, KP = SmallCfg<T, FR>::KP;
    __shared__ __align__(16) T sb[4][2][KP][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wave_first = (blockIdx.x * 4 + wv) * GW;
    if (wave_first >= count) return;  // whole wave
    const int i = lane % FR, base = lane - i;
    const int idx = wave_first + lane / FR;
    const SmallRec rc = recs[idx < count ? idx : count - 1];
    // right-hand sides past the system's count are zero (padding, or converged in a
    // refinement step): left out, their vectors stay zero; a front with none is left out
    // like a group past the end of the list
    const int nc = idx < count ? (a.cols ? min(a.nrhs, a.cols[rc.sys]) : a.nrhs) : 0;
    const bool live = nc > 0;
    if (!__any(live)) return;  // whole wave
    const int f = live ? rc.f : 0, w = live ? rc.w : 0, fc = rc.fc;
    const bool own = i < w, foot = i >= w && i < f;
    const T *M = arena + rc.off;
    // own lanes walk column i of the front (stride 1), footer lanes row i (stride f);
    // offsets stay below 64 * 64, so 32-bit per-lane arithmetic
    const int lim = own ? i : (foot ? w : 0);  // entries m < lim are used
    const int start = own ? i * f : (foot ? i : 0), stride = own ? 1 : f, last = max(lim - 1, 0);
    T R[FR];
#pragma unroll
    for (int m = 0; m < FR; ++m) {
        const T val = M[start + min(m, last) * stride];  // clamped: always inside the front
        R[m] = m < lim ? val : T(0);
    }
    // the first kKeep children's footer offsets stay in registers across right-hand sides
    constexpr int kKeep = 4;
    const int nrhs = a.nrhs, n = a.n;
    const int c0 = live ? rc.c0 : 0, c1 = live ? rc.c1 : 0;
    int64_t cb[kKeep];
    int cf[kKeep];
    // child slots some front of the wave uses (wave-uniform): picks a straight-line gather
    int nq = 0;
#pragma unroll
    for (int q = 0; q < kKeep; ++q)
        if (__any(c0 + q < c1)) nq = q + 1;
    // Unconditional loads, from this front's record where there is no such child: under a branch
    // a load lands in a temporary whose copy waits for it, so the children would load one after
    // the other.
    // Every child's loads are issued before any of them is used.
    const SmallRec *rsafe = recs + (idx < count ? idx : count - 1);
    int64_t chv[kKeep];
    int chw[kKeep], chf[kKeep], amv[kKeep];
#pragma unroll
    for (int q = 0; q < kKeep; ++q) {
        const bool has = c0 + q < c1 && i < f;
        const SmallChild *chp = has ? a.child + (c0 + q) : reinterpret_cast<const SmallChild *>(rsafe);
        chv[q] = chp->vb;
        chw[q] = chp->w;
        chf[q] = chp->f;
        amv[q] = *(has ? a.inv_map + (rc.inv0 + (int64_t)q * f + i) : reinterpret_cast<const int *>(rsafe));
    }
#pragma unroll
    for (int q = 0; q < kKeep; ++q) {
        const bool has = c0 + q < c1 && i < f;
        cf[q] = has ? chf[q] : 0;
        cb[q] = has && amv[q] >= 0 ? chv[q] * nrhs + chw[q] + amv[q] : -1;
    }
    T *Vg = V + rc.vbase * nrhs;
    const int yi = own ? fc + i : fc;  // clamped: other lanes read a valid row, then drop it
    // a missing child's load is clamped into this front's own vector block: an address of the
    // lane's neighbourhood (a fixed address would put every wave of the launch on one line)
    const int64_t vself = rc.vbase * nrhs + min(i, max(f - 1, 0));
    T(*b0)[64] = sb[wv][0];
    T(*b1)[64] = sb[wv][1];
    for (int r0 = 0; r0 < nrhs; r0 += KP) {
        if (!__any(r0 < nc)) break;  // wave-uniform
        T v[KP], y[KP];
        switch (nq) {
        case 0: fwd_gather<T, KP, 0>(v, Y, V, r0, nc, i, f, own, n, yi, cb, cf, vself); break;
        case 1: fwd_gather<T, KP, 1>(v, Y, V, r0, nc, i, f, own, n, yi, cb, cf, vself); break;
        case 2: fwd_gather<T, KP, 2>(v, Y, V, r0, nc, i, f, own, n, yi, cb, cf, vself); break;
        case 3: fwd_gather<T, KP, 3>(v, Y, V, r0, nc, i, f, own, n, yi, cb, cf, vself); break;
        default: fwd_gather<T, KP, 4>(v, Y, V, r0, nc, i, f, own, n, yi, cb, cf, vself); break;
        }
        if (c1 - c0 > kKeep && i < f && r0 < nc) {
            // children past the first kKeep, in order: per child its record and map entry, then
            // the chunk's values together (columns clamped to the last one, masked)
            for (int c = c0 + kKeep; c < c1; ++c) {
                const SmallChild ch = a.child[c];
                const int am = a.inv_map[rc.inv0 + (int64_t)(c - c0) * f + i];
                if (am < 0) continue;
                const T *src = V + ch.vb * nrhs + ch.w + am;
                T u[KP];
#pragma unroll
                for (int p = 0; p < KP; ++p) u[p] = src[(int64_t)min(r0 + p, nc - 1) * ch.f];
#pragma unroll
                for (int p = 0; p < KP; ++p) pin(u[p]);
#pragma unroll
                for (int p = 0; p < KP; ++p) v[p] += r0 + p < nc ? u[p] : T(0);
            }
        }
#pragma unroll
        for (int p = 0; p < KP; ++p) b0[p][lane] = own ? v[p] : T(0);
        wave_sync();
#pragma unroll
        for (int p = 0; p < KP; ++p) y[p] = v[p] + lds_dot<T, FR>(R, &b0[p][base]);
#pragma unroll
        for (int p = 0; p < KP; ++p) b1[p][lane] = own ? y[p] : T(0);
        wave_sync();
#pragma unroll
        for (int p = 0; p < KP; ++p) {
            const int r = r0 + p;
            const T vf = v[p] - lds_dot<T, FR>(R, &b1[p][base]);
            if (r < nc && i < f) {
                if (own) Y[(int64_t)r * n + fc + i] = y[p];
                else Vg[(int64_t)r * f + i] = vf;
            }
        }
    }
}

// Backward: z = y / d - L_foot' x_foot, then x = Li' z on the own rows.  Own lane i keeps
// R[j] = Li[j][i] (row i, column j of the front) for i < j < w and R[j] = L[j][i]
// (column i, row j) for w <= j < f: the two index ranges are disjoint, so one register row
// serves both products, once against x_foot (zero below w) and once against z (zero from w).
template <typename T, int FR>
__global__ __launch_bounds__(256) void k_bwd_small(const SmallRec *__restrict__ recs, int count,
                                                   const int *__restrict__ pat_rows, const T *__restrict__ arena,
                                                   const double *__restrict__ D, T *__restrict__ Y, int n, int nrhs,
                                                   const int *__restrict__ cols) {
    constexpr int GW = SmallCfg<T, FR>::GW, KP = SmallCfg<T, FR>::KP;
    __shared__ __align__(16) T sb[4][2][KP][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wave_first = (blockIdx.x * 4 + wv) * GW;
    if (wave_first >= count) return;  // whole wave
    const int i = lane % FR, base = lane - i;
    const int idx = wave_first + lane / FR;
    const SmallRec rc = recs[idx < count ? idx : count - 1];
    // right-hand sides past the system's count are zero and left out (as in k_fwd_small)
    const int nc = idx < count ? (cols ? min(nrhs, cols[rc.sys]) : nrhs) : 0;
    const bool live = nc > 0;
    if (!__any(live)) return;  // whole wave
    const int f = live ? rc.f : 0, w = live ? rc.w : 0, fc = rc.fc;
    const bool own = i < w, foot = i >= w && i < f;
    const T *M = arena + rc.off;
    // j < w: used for i < j (row i, stride f); j >= w: used for j < f (column i, stride 1)
    const int ii = own ? i : 0, flast = max(f - 1, 0);
    T R[FR];
#pragma unroll
    for (int j = 0; j < FR; ++j) R[j] = M[j < w ? ii + j * f : ii * f + min(j, flast)];  // clamped: inside the front
#pragma unroll
    for (int j = 0; j < FR; ++j) pin(R[j]);  // all loads issued before the masks
#pragma unroll
    for (int j = 0; j < FR; ++j) R[j] = (own && (j < w ? j > i : j < f)) ? R[j] : T(0);
    const int yi = own ? fc + i : fc;
    const int row = foot ? pat_rows[rc.pat + i] : yi;  // clamped to the lane's own row
    const T dv = own ? (T)D[fc + i] : T(1);
    T(*b0)[64] = sb[wv][0];
    T(*b1)[64] = sb[wv][1];
    for (int r0 = 0; r0 < nrhs; r0 += KP) {
        if (!__any(r0 < nc)) break;  // wave-uniform
        T xf[KP], yo[KP], z[KP];
#pragma unroll
        for (int p = 0; p < KP; ++p) {
            const int r = r0 + p;
            const bool act = r < nc;
            const int64_t rr = act ? r : 0;
            const T xv = Y[rr * n + row], yv = Y[rr * n + yi];
            xf[p] = (act && foot) ? xv : T(0);
            yo[p] = (act && own) ? yv : T(0);
        }
#pragma unroll
        for (int p = 0; p < KP; ++p) b0[p][lane] = xf[p];
        wave_sync();
#pragma unroll
        for (int p = 0; p < KP; ++p) z[p] = own ? yo[p] / dv - lds_dot<T, FR>(R, &b0[p][base]) : T(0);
#pragma unroll
        for (int p = 0; p < KP; ++p) b1[p][lane] = z[p];
        wave_sync();
#pragma unroll
        for (int p = 0; p < KP; ++p) {
            const int r = r0 + p;
            const T x = z[p] + lds_dot<T, FR>(R, &b1[p][base]);
            if (r < nc && own) Y[(int64_t)r * n + fc + i] = x;
        }
    }
}

// Backward footer, gather: the x values of every front's footer rows (final:
// ancestors are solved first) into the front's vector, V[g][rr][w + a] = x[rows[w + a]].
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_gather(const int2 *__restrict__ tasks, const int *__restrict__ F,
                                                    const int *__restrict__ W, const int *__restrict__ pat_ptr,
                                                    const int *__restrict__ pat_rows, const int64_t *__restrict__ vbase,
                                                    const T *__restrict__ Y, T *__restrict__ V, int n, int nrhs,
        const int *__restrict__ fsys, const int *__restrict__ cols) {
    const int2 tk = tasks[blockIdx.x];
    const int nc = cols ? min(nrhs, cols[fsys[tk.x]]) : nrhs;  // right-hand sides the system needs
    if (nc == 0) return;
    const int g = tk.x, f = F[g], w = W[g];
    const int i = tk.y * 256 + threadIdx.x;
    if (i < w || i >= f) return;
    const int row = pat_rows[pat_ptr[g] + i];
    T *Vg = V + vbase[g] * nrhs;
    for (int r = 0; r < nc; ++r) Vg[(int64_t)r * f + i] = Y[(int64_t)r * n + row];
}

// Backward footer, stage 2: Z[own c] = y_c / d_c - sum of the partials (fixed order).
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_foot_red(const int4 *__restrict__ tasks, const int *__restrict__ W,
                                                      const int *__restrict__ firstcol, const double *__restrict__ D,
                                                      const T *__restrict__ Y, const T *__restrict__ part,
                                                      T *__restrict__ Z, int n, int nrhs,
        const int *__restrict__ fsys, const int *__restrict__ cols) {
    const int4 tk = tasks[blockIdx.x];
    const int nc = cols ? min(nrhs, cols[fsys[tk.x]]) : nrhs;  // right-hand sides the system needs
    if (nc == 0) return;
    const int g = tk.x, cb0 = tk.y * kPanel, tb = tk.z, nch = tk.w;
    const int ncol = min(kPanel, W[g] - cb0), fc = firstcol[g] + cb0;
    for (int idx = threadIdx.x; idx < ncol * nc; idx += 256) {
        const int col = idx % ncol, rr = idx / ncol;
        const T *pc = part + ((int64_t)tb * kPanel + col) * nrhs + rr;  // chunk ch at pc[ch * kPanel * nrhs]
        const int64_t e = (int64_t)rr * n + fc + col;
        const T yd = Y[e] / (T)D[fc + col];
        T s = T(0);
        for (int ch0 = 0; ch0 < nch; ch0 += 8) {  // eight loads in flight, then the sums in chunk order
            T v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = ch0 + j < nch ? pc[(int64_t)(ch0 + j) * kPanel * nrhs] : T(0);
#pragma unroll
            for (int j = 0; j < 8; ++j) s += v[j];
        }
        Z[e] = yd - s;
    }
}

// ---- block solves of the large fronts --------------------------------------------
//
// At the end of the factorization the unit-lower diagonal block L_JJ of every kSBlock-column
// block J of a front (sub-blocks of 64 columns, the last one possibly narrower) is inverted:
// X = L_JJ^{-1} with X_aa = Li_a (the sub-blocks' stored inverses) and, for a > b,
// X_ab = -Li_a sum_{c=b}^{a-1} L_ac X_cb.  X[k][m] (k > m) is kept at front row m, column k,
// the transposed layout of the 64-column inverses, which are its diagonal blocks: the
// sub-diagonal blocks X_ab fill the block's otherwise unused upper triangle.  One launch per
// distance d = a - b: X_ab reads only blocks of smaller distance, written by earlier launches.
// The solves then take one dependent step per kSBlock columns instead of per 64.
template <typename T>
__global__ __launch_bounds__(256, 2) void k_inv_blk(const int4 *__restrict__ tasks, T *__restrict__ arena,
                                                 const int *__restrict__ F, const int *__restrict__ W,
                                                 const int64_t *__restrict__ off) {
    __shared__ __align__(16) T As[kPanel][kSchurLd<T>];
    __shared__ __align__(16) T Bs[kPanel][kSchurLd<T>];
    const int4 tk = tasks[blockIdx.x];
    const int g = tk.x, a = tk.z, b = tk.w;
    const int f = F[g], w = W[g];
    const int bs = tk.y * kSBlock, be = min(bs + kSBlock, w);
    T *M = arena + off[g];
    const int pa = bs + kPanel * a, pb = bs + kPanel * b;
    const int na = min(kPanel, be - pa), nbb = min(kPanel, be - pb);
    const int tid = threadIdx.x, l = tid & 63, q4 = tid >> 6;
    typename MfmaAcc<T>::type acc[2][2];
    schur_acc_zero<T>(acc);
    // S = sum_{c=b}^{a-1} L_ac X_cb: As[k][r] = L_ac[r][k] = M[(pc + k) f + pa + r],
    // Bs[k][m] = X_cb[k][m] = M[(pc + k) f + pb + m] (c = b: Li_b, unit diagonal, zero above);
    // the next c's loads are in flight during this one's products
    T av[kPanel / 4], bv[kPanel / 4];
    auto load = [&](int c) {
        const int pc = bs + kPanel * c, nc = min(kPanel, be - pc);
#pragma unroll
        for (int j = 0; j < kPanel / 4; ++j) {
            const T *col = M + (int64_t)(pc + min(q4 + 4 * j, nc - 1)) * f;
            av[j] = col[pa + min(l, na - 1)];
            bv[j] = col[pb + min(l, nbb - 1)];
        }
    };
    load(b);
    for (int c = b; c < a; ++c) {
        const int pc = bs + kPanel * c, nc = min(kPanel, be - pc);
        __syncthreads();  // the previous product's reads are done
#pragma unroll
        for (int j = 0; j < kPanel / 4; ++j) {
            const int k = q4 + 4 * j;
            T x = bv[j];
            if (c == b) x = l < k ? x : (l == k ? T(1) : T(0));
            As[k][sch_col<T>(k, l)] = (k < nc && l < na) ? av[j] : T(0);
            Bs[k][sch_col<T>(k, l)] = (k < nc && l < nbb) ? x : T(0);
        }
        __syncthreads();
        if (c + 1 < a) load(c + 1);
        schur_acc_tile<T>(As, Bs, acc);
    }
    // X_ab = -Li_a S: As[k][r] = Li_a[r][k] = M[(pa + r) f + pa + k] (k < r; unit diagonal),
    // Bs[k][m] = S[k][m] from the accumulators
    const int lane = tid & 63, wave = tid >> 6;
    const int rq = (wave & 1) * 32, cq = (wave >> 1) * 32, l16 = lane & 15, lk = lane >> 4;
    T li[kPanel / 4];
#pragma unroll
    for (int j = 0; j < kPanel / 4; ++j) li[j] = M[(int64_t)(pa + min(q4 + 4 * j, na - 1)) * f + pa + min(l, na - 1)];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kk = rq + 16 * x + l16;
                Bs[kk][sch_col<T>(kk, cq + 16 * y + MfmaAcc<T>::row(lk, q))] = acc[x][y][q];
            }
#pragma unroll
    for (int j = 0; j < kPanel / 4; ++j) {
        const int r = q4 + 4 * j;  // li[j] = M[(pa + r) f + pa + l] = Li_a[r][l] for l < r
        As[l][sch_col<T>(l, r)] = (r < na && l < na) ? (l < r ? li[j] : (l == r ? T(1) : T(0))) : T(0);
    }
    __syncthreads();
    schur_acc_zero<T>(acc);
    schur_acc_tile<T>(As, Bs, acc);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = rq + 16 * x + l16, m = cq + 16 * y + MfmaAcc<T>::row(lk, q);
                if (r < na && m < nbb) M[(int64_t)(pa + r) * f + pb + m] = -acc[x][y][q];
            }
}

// The block solves feed the matrix cores from registers: no LDS staging and no barriers.  A
// task walks up to kSSub 64-deep tiles along k, with the next kSPre tiles' loads in flight.  In
// a tile, MFMA step s = 4 i + e takes on lane group lk = lane / 16 the index k = 16 i + 4 lk + e,
// so that a lane's four k of one i are consecutive in memory (one vector load where the operand
// runs along k).  Every kernel sums over k in this order.  The right-hand sides are the MFMA
// rows (16 per pass), the block's rows or columns the MFMA columns (16 per wave).
// NS = 1 for the tasks of blocks of at most 64 columns, kSSub otherwise (separate launches).
constexpr int kSSub = kSBlock / kPanel;  // 64-column sub-blocks of a solve block
template <typename T>
constexpr int kSPre = sizeof(T) == 4 ? 4 : 2;  // tiles in flight per wave
// Split tasks (KS = 2 or 4, for blocks wider than 64 columns): the task's 64 outputs (four
// groups of 16) spread over KS workgroups of 4 / KS groups each, and the KS waves of a group
// split the tiles along k (wave ks sums tiles ks, ks + KS, ...); the group's ks = 0 wave adds
// the others' partial sums in wave order.  The chain of dependent loads per wave is 1 / KS of
// the block's.  KS = 1: one workgroup per task, wave v takes group v over every tile.
// Workgroups go to the 8 XCDs in turn, so the KS workgroups of a task take ids 8 apart: one XCD
// and one L2, where a 128-byte line of a row-strided operand serves all of them.  The grid is
// padded to a multiple of 8 KS; ids past the tasks return.
constexpr int kSSplit = 4;  // the diagonal-block solves and the backward update
// The wide forward update splits k over four groups of one workgroup when a launch has at most
// this many tasks (one per CU); above it the tasks alone fill the chip (C3: 4.89 ms of forward
// kernels per step with the split at most 256 tasks, 5.27 never split, 5.09 always).
constexpr int kSolveTasksPerChip = 256;
template <int KS>
static inline int split_grid(int ntask) { return KS == 1 ? ntask : (ntask + 7) / 8 * 8 * KS; }
// This wave's task, output group (16 outputs of the task's 64) and k share.
template <int KS>
__device__ __forceinline__ void split_task(int &task, int &grp, int &ks) {
    const int wv = threadIdx.x >> 6;
    if constexpr (KS == 1) {
        task = blockIdx.x;
        grp = wv;
        ks = 0;
    } else {
        constexpr int G = 4 / KS;  // groups per workgroup
        const int b = blockIdx.x, r = b >> 3;
        task = ((r / KS) << 3) | (b & 7);
        grp = (r % KS) * G + wv % G;
        ks = wv / G;
    }
}

// acc (wave ks's partial sum) -> the group's ks = 0 wave holds the sum over its KS waves, added
// in wave order; red holds KS - 1 waves' accumulators per group of the workgroup.
template <typename T, int KS>
__device__ __forceinline__ void split_reduce(typename MfmaAcc<T>::type &acc, T (*red)[4 / KS][4][64], int ks) {
    constexpr int G = 4 / KS;
    const int lane = threadIdx.x & 63, gl = (threadIdx.x >> 6) % G;
    if (ks > 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) red[ks - 1][gl][q][lane] = acc[q];
    __syncthreads();
    if (ks == 0)
#pragma unroll
        for (int v = 0; v < KS - 1; ++v)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += red[v][gl][q][lane];
    __syncthreads();  // red is free for the next pass
}

// v[4 i + e] = p[16 i + 4 lk + e] where that index is < nk (a vector starting at or past nk reads
// p[0 .. 3] instead).  Vector loads; the four values past a column's end (or nk) that one of
// them may touch are in the same allocation (the next column, or the next buffer of the
// workspace, or the factor store's padding).  kvec_mask zeroes the entries at or past nk; it is
// applied where the tile is used, so that the loads of several tiles stay in flight (a mask
// right after the load would wait for it).
template <typename T>
__device__ __forceinline__ void kvec_load(T (&v)[16], const T *__restrict__ p, int nk) {
    const int lk = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int kb = 16 * i + 4 * lk, kc = kb < nk ? kb : 0;
        const T *q = p + kc;
        T x[4];
        if constexpr (sizeof(T) == 4) {
            const mg_u4u u = *reinterpret_cast<const mg_u4u *>(q);
            __builtin_memcpy(x, &u, 16);
        } else {
            const mg_u4u u0 = *reinterpret_cast<const mg_u4u *>(q), u1 = *reinterpret_cast<const mg_u4u *>(q + 2);
            __builtin_memcpy(x, &u0, 16);
            __builtin_memcpy(x + 2, &u1, 16);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) v[4 * i + e] = x[e];
    }
}
template <typename T>
__device__ __forceinline__ void kvec_mask(T (&v)[16], int nk) {
    const int lk = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) v[4 * i + e] = 16 * i + 4 * lk + e < nk ? v[4 * i + e] : T(0);
}
// v[4 i + e] = p[(16 i + 4 lk + e) * ld] with the index clamped to nk - 1 (values past nk are
// used only against zero operands).
template <typename T>
__device__ __forceinline__ void kstride_load(T (&v)[16], const T *__restrict__ p, int64_t ld, int nk) {
    const int lk = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) v[4 * i + e] = p[(int64_t)min(16 * i + 4 * lk + e, nk - 1) * ld];
}
template <typename T>
__device__ __forceinline__ void tile_mfma(typename MfmaAcc<T>::type &acc, const T (&a)[16], const T (&b)[16]) {
#pragma unroll
    for (int s = 0; s < 16; ++s) acc = MfmaAcc<T>::step(a[s], b[s], acc);
}
// Tiles [t0, t1) (t0 < t1) of a task through the matrix cores with P tiles' loads in flight, in
// named registers (no dynamic indexing): load(t, a, b) fills tile t's operands, prep(t, a, b)
// adjusts them before use (the masks past the tile's end, the diagonal tile's mask).
template <typename T, int P, typename Load, typename Prep>
__device__ __forceinline__ void tile_chain(typename MfmaAcc<T>::type &acc, int t0, int t1, Load &&load, Prep &&prep) {
    static_assert(P >= 1 && P <= 4, "up to four tiles in flight");
    T a0[16], b0[16], a1[16], b1[16], a2[16], b2[16], a3[16], b3[16];
    load(t0, a0, b0);
    if constexpr (P > 1)
        if (t0 + 1 < t1) load(t0 + 1, a1, b1);
    if constexpr (P > 2)
        if (t0 + 2 < t1) load(t0 + 2, a2, b2);
    if constexpr (P > 3)
        if (t0 + 3 < t1) load(t0 + 3, a3, b3);
    auto step = [&](int t, T(&av)[16], T(&bv)[16]) {
        if (t >= t1) return;  // uniform
        prep(t, av, bv);
        tile_mfma<T>(acc, av, bv);
        if (t + P < t1) load(t + P, av, bv);
    };
    for (int t = t0; t < t1; t += P) {
        step(t, a0, b0);
        if constexpr (P > 1) step(t + 1, a1, b1);
        if constexpr (P > 2) step(t + 2, a2, b2);
        if constexpr (P > 3) step(t + 3, a3, b3);
    }
}

// Tiles [0, nu) (nu <= NT) of a task through the matrix cores, every tile's loads issued first.
template <typename T, int NT, typename Load, typename Prep>
__device__ __forceinline__ void tile_all(typename MfmaAcc<T>::type &acc, int nu, Load &&load, Prep &&prep) {
    T a[NT][16], b[NT][16];
#pragma unroll
    for (int u = 0; u < NT; ++u)
        if (u < nu) load(u, a[u], b[u]);
#pragma unroll
    for (int u = 0; u < NT; ++u)
        if (u < nu) {
            prep(u, a[u], b[u]);
            tile_mfma<T>(acc, a[u], b[u]);
        }
}
// A wave's nu tiles (at most NT): all in flight at once where NT <= 4, a chain with up to P in
// flight otherwise.
template <typename T, int NT, int P, typename Load, typename Prep>
__device__ __forceinline__ void tile_run(typename MfmaAcc<T>::type &acc, int nu, Load &&load, Prep &&prep) {
    if constexpr (NT <= 4) {
        tile_all<T, NT>(acc, nu, load, prep);
    } else {
        if (nu > 0) tile_chain<T, P>(acc, 0, nu, load, prep);
    }
}

// Backward footer, stage 1: partial[task][col][rr] = sum over one chunk of footer rows of
// L[i, 64cb+col] x[row i] (x of the footer rows are final: ancestors are solved first).
// On the matrix cores: (partial)^T = x^T L[chunk, block] in 16x16x4 tiles.  Wave v takes the
// chunk's 64-row tile v for all 64 columns (four 16-column products sharing its x operand);
// both operands run along k (the footer rows of a column of L and of x), so they go from
// memory straight into the MFMA registers, every load of the chunk in flight at once.  Wave v
// then adds the four waves' sums for columns 16 v .. 16 v + 15 in wave order.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_foot_part(const int4 *__restrict__ tasks, const T *__restrict__ arena,
                                                       const int *__restrict__ F, const int *__restrict__ W,
                                                       const int64_t *__restrict__ off,
                                                       const int64_t *__restrict__ vbase, const T *__restrict__ V,
                                                       T *__restrict__ part, int nrhs,
        const int *__restrict__ fsys, const int *__restrict__ cols) {
    static_assert(kRhsChunk == 16, "one 16-row MFMA tile of right-hand sides per pass");
    static_assert(kFootChunk == 4 * kPanel, "one 64-row tile per wave");
    __shared__ T red[4][4][4][64];  // [wave][column group][q][lane]
    const int4 tk = tasks[blockIdx.x];
    const int nc = cols ? min(nrhs, cols[fsys[tk.x]]) : nrhs;  // right-hand sides the system needs
    if (nc == 0) return;
    const int g = tk.x, f = F[g], w = W[g];
    const int cb0 = tk.y * kPanel, ncol = min(kPanel, w - cb0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, l16 = lane & 15, lk = lane >> 4;
    const int i0 = w + tk.z * kFootChunk + kPanel * wv, nk = min(kPanel, f - i0);  // this wave's rows
    const T *M = arena + off[g];
    const T *Vg = V + vbase[g] * nrhs;
    for (int c0 = 0; c0 < nc; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nc - c0);
        typename MfmaAcc<T>::type acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[t][q] = T(0);
        if (nk > 0) {
            T a[16], b[4][16];
            kvec_load<T>(a, Vg + (int64_t)(c0 + min(l16, nr - 1)) * f + i0, nk);  // x[row i][c0 + l16]
#pragma unroll
            for (int t = 0; t < 4; ++t)  // L[i][cb0 + 16 t + l16]; column groups past ncol skipped
                if (16 * t < ncol) kvec_load<T>(b[t], M + (int64_t)(cb0 + min(16 * t + l16, ncol - 1)) * f + i0, nk);
            kvec_mask<T>(a, nk);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (16 * t < ncol) {
                    kvec_mask<T>(b[t], nk);
                    tile_mfma<T>(acc[t], a, b[t]);
                }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) red[wv][t][q][lane] = acc[t][q];
        __syncthreads();
        T sum[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sum[q] = red[0][wv][q][lane];
#pragma unroll
            for (int v = 1; v < 4; ++v) sum[q] += red[v][wv][q][lane];
        }
        const int col = 16 * wv + l16;
        if (col < ncol)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = MfmaAcc<T>::row(lk, q);
                if (rr < nr) part[((int64_t)blockIdx.x * kPanel + col) * nrhs + c0 + rr] = sum[q];
            }
        __syncthreads();  // red is free for the next pass
    }
}

// Forward, step 1 of block J: y of sub-block a, y[pa + j] = sum_{k <= j} X[pa + j][k] v[k] over
// the block's sub-blocks c <= a (X[pa + j][pc + k] at front row pc + k, column pa + j: one
// vector load along k), as y^T = v^T X^T; y goes to Y.
template <typename T, int NS, int KS>
__global__ __launch_bounds__(256) void k_fwd_diag(const PanelTask *__restrict__ tasks, const T *__restrict__ arena,
                                                  const T *__restrict__ V, T *__restrict__ Y, int J, int n, int nrhs,
                                                  const int *__restrict__ cols, int ntask) {
    static_assert(kRhsChunk == 16, "one 16-row MFMA tile of right-hand sides per pass");
    static_assert(KS == 1 || KS == 2 || KS == 4, "1, 2 or 4 waves per output group");
    constexpr int NT = (NS + KS - 1) / KS;  // tiles per wave, at most
    constexpr int P = NT < kSPre<T> ? NT : kSPre<T>;
    __shared__ T red[KS > 1 ? KS - 1 : 1][4 / KS][4][64];
    int task, grp, ks;
    split_task<KS>(task, grp, ks);
    if (task >= ntask) return;
    const PanelTask tk = tasks[task];
    const int nc = cols ? min(nrhs, cols[tk.sys]) : nrhs;  // right-hand sides the system needs
    if (nc == 0) return;
    const int a = NS == 1 ? 0 : tk.blk, f = tk.f, w = tk.w;
    const int bs = J * kSBlock, be = min(bs + kSBlock, w);
    const int pa = bs + kPanel * a, na = min(kPanel, be - pa);
    const T *M = arena + tk.off;
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const int nu = (a + 1 - ks + KS - 1) / KS;  // this wave's tiles c = ks + KS u <= a
    const int j = 16 * grp + l16;  // this lane's output row of the sub-block (MFMA column)
    if (KS == 1 && 16 * grp >= na) return;  // a wave past the sub-block's columns (no barriers when KS = 1)
    const T *xcol = M + (int64_t)(pa + min(j, na - 1)) * f;  // X[pa + j][.] along the rows of column pa + j
    for (int c0 = 0; c0 < nc; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nc - c0);
        const T *vcol = V + tk.vbase * nrhs + (int64_t)(c0 + min(l16, nr - 1)) * f;  // v[.][c0 + l16]
        auto load = [&](int u, T(&av)[16], T(&bv)[16]) {
            const int pc = bs + kPanel * (ks + KS * u), ncc = min(kPanel, be - pc);
            kvec_load<T>(av, vcol + pc, ncc);
            kvec_load<T>(bv, xcol + pc, ncc);
        };
        auto prep = [&](int u, T(&av)[16], T(&bv)[16]) {
            const int ncc = min(kPanel, be - bs - kPanel * (ks + KS * u));
            kvec_mask<T>(av, ncc);
            kvec_mask<T>(bv, ncc);
            if (ks + KS * u != a) return;  // X_aa = Li_a: unit diagonal, zero above it
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int k = 16 * (s >> 2) + 4 * lk + (s & 3);
                bv[s] = k < j ? bv[s] : (k == j ? T(1) : T(0));
            }
        };
        typename MfmaAcc<T>::type acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = T(0);
        tile_run<T, NT, P>(acc, nu, load, prep);
        if constexpr (KS > 1) split_reduce<T, KS>(acc, red, ks);
        // acc[q] = y[pa + j][c0 + rr], rr = row(lk, q)
        if (ks == 0 && j < na)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = MfmaAcc<T>::row(lk, q);
                if (rr < nr) Y[(int64_t)(c0 + rr) * n + tk.fc + pa + j] = acc[q];
            }
    }
}

// Forward, step 2 of block J: the 64 rows r0 + i below the block, V[r0 + i] -= sum_k
// L[r0 + i][k] y[k] over the block's columns, as (V_rows)^T -= y^T L[rows, block]^T (y along
// k by vector loads, L down the block's columns).
template <typename T, int NS, int KS>
__global__ __launch_bounds__(256) void k_fwd_upd(const PanelTask *__restrict__ tasks, const T *__restrict__ arena,
                                                 T *__restrict__ V, const T *__restrict__ Y, int J, int n, int nrhs,
                                                 const int *__restrict__ cols, int ntask) {
    static_assert(kRhsChunk == 16, "one 16-row MFMA tile of right-hand sides");
    static_assert(KS == 1 || KS == 2 || KS == 4, "1, 2 or 4 waves per output group");
    constexpr int NT = (NS + KS - 1) / KS;
    constexpr int P = NT < kSPre<T> ? NT : kSPre<T>;
    __shared__ T red[KS > 1 ? KS - 1 : 1][4 / KS][4][64];
    int task, grp, ks;
    split_task<KS>(task, grp, ks);
    if (task >= ntask) return;
    const PanelTask tk = tasks[task];
    const int nc = cols ? min(nrhs, cols[tk.sys]) : nrhs;
    if (nc == 0) return;
    const int f = tk.f, w = tk.w;
    const int bs = J * kSBlock, be = min(bs + kSBlock, w), nk = be - bs;
    const int nt = NS == 1 ? 1 : (nk + kPanel - 1) / kPanel;
    const int r0 = be + tk.blk * kTile, nrow = min(kTile, f - r0);
    const T *M = arena + tk.off;
    T *Vg = V + tk.vbase * nrhs;
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const int nu = (nt - ks + KS - 1) / KS;  // this wave's tiles ks + KS u < nt
    const int i = 16 * grp + l16;  // this lane's row below the block (MFMA column)
    if (KS == 1 && 16 * grp >= nrow) return;  // a wave past the strip's rows (no barriers when KS = 1)
    const T *lrow = M + (int64_t)bs * f + r0 + min(i, nrow - 1);  // L[r0 + i][bs + k] = lrow[k f]
    for (int c0 = 0; c0 < nc; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nc - c0);
        const T *ycol = Y + (int64_t)(c0 + min(l16, nr - 1)) * n + tk.fc + bs;  // y[bs + .][c0 + l16]
        T old[4] = {};
        if (ks == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                old[q] = Vg[(int64_t)(c0 + min(MfmaAcc<T>::row(lk, q), nr - 1)) * f + r0 + min(i, nrow - 1)];
        auto load = [&](int u, T(&av)[16], T(&bv)[16]) {
            const int kc = kPanel * (ks + KS * u), nkc = min(kPanel, nk - kc);
            kvec_load<T>(av, ycol + kc, nkc);
            kstride_load<T>(bv, lrow + (int64_t)kc * f, f, nkc);
        };
        typename MfmaAcc<T>::type acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = T(0);
        auto prep = [&](int u, T(&av)[16], T(&)[16]) { kvec_mask<T>(av, min(kPanel, nk - kPanel * (ks + KS * u))); };
        tile_run<T, NT, P>(acc, nu, load, prep);
        if constexpr (KS > 1) split_reduce<T, KS>(acc, red, ks);
        if (ks == 0 && i < nrow)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = MfmaAcc<T>::row(lk, q);
                if (rr < nr) Vg[(int64_t)(c0 + rr) * f + r0 + i] = old[q] - acc[q];
            }
    }
}

// Forward, step 2 of block J for blocks wider than 64 columns: the same product with the L
// tiles staged through LDS.  The L operand runs along rows (down a column of the front), so a
// group of 256 threads copies each 64 x 64 tile with 16-byte loads, four threads per column
// (256 contiguous bytes), into Ls[k][i], from where a lane reads its row i along k.  Two tiles
// are in flight in registers and the LDS copy is double-buffered (one barrier per tile); y
// comes from memory one tile ahead.  KQ = 4 (launches with few tasks): four such groups in one
// workgroup take every fourth tile each, and group 0 adds the four sums in group order.  Rows
// of the last 64-row strip past the front's end are read (at most 63 elements past the column,
// inside the arena's tail padding) and never stored; columns past the block's end are clamped
// to its last column and meet zero y.
template <typename T, int KQ>
__global__ __launch_bounds__(256 * KQ) void k_fwd_upd_wide(const PanelTask *__restrict__ tasks,
                                                           const T *__restrict__ arena, T *__restrict__ V,
                                                           const T *__restrict__ Y, int J, int n, int nrhs,
                                                           const int *__restrict__ cols) {
    static_assert(kRhsChunk == 16, "one 16-row MFMA tile of right-hand sides");
    static_assert(KQ == 1 || (KQ == 4 && sizeof(T) == 4), "one group, or four (FP32: 136 KB of LDS)");
    constexpr int LD = kPanel + 16 / (int)sizeof(T);  // row of Ls: 16-byte aligned, reads without conflicts
    constexpr int VW = 16 / (int)sizeof(T), NV = 16 / VW;  // elements per vector, vectors per thread and tile
    __shared__ __align__(16) T Ls[KQ][2][kPanel][LD];
    const PanelTask tk = tasks[blockIdx.x];
    const int nc = cols ? min(nrhs, cols[tk.sys]) : nrhs;
    if (nc == 0) return;
    const int f = tk.f, w = tk.w;
    const int bs = J * kSBlock, be = min(bs + kSBlock, w), nk = be - bs;
    const int nt = (nk + kPanel - 1) / kPanel;
    const int r0 = be + tk.blk * kTile, nrow = min(kTile, f - r0);
    const T *M = arena + tk.off;
    T *Vg = V + tk.vbase * nrhs;
    const int tid = threadIdx.x & 255, kq = threadIdx.x >> 8;
    const int lane = tid & 63, wv = tid >> 6, l16 = lane & 15, lk = lane >> 4;
    const int i = 16 * wv + l16;  // this lane's row below the block (MFMA column)
    const int nu = (nt - kq + KQ - 1) / KQ;  // this group's tiles u = kq + KQ j, j < nu
    const int nj = (nt + KQ - 1) / KQ;       // steps (group 0 has the most tiles)
    // The loads are unconditional (under a branch a load lands in a temporary whose copy waits
    // for it); past the group's tiles every lane reads the strip's first 64 bytes (one cached line).
    const int lc = tid >> 2, lr = 16 * (tid & 3);  // staging: column lc of a tile, rows lr .. lr + 15
    const T *lbase = M + (int64_t)bs * f + r0, *lsrc = lbase + lr;
    auto gload = [&](int j, mg_u4u(&r)[NV]) {
        const T *p = j < nu ? lsrc + (int64_t)min(kPanel * (kq + KQ * j) + lc, nk - 1) * f : lbase;
#pragma unroll
        for (int v = 0; v < NV; ++v) r[v] = *reinterpret_cast<const mg_u4u *>(p + VW * v);
    };
    for (int c0 = 0; c0 < nc; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nc - c0);
        const T *ycol = Y + (int64_t)(c0 + min(l16, nr - 1)) * n + tk.fc + bs;  // y[bs + .][c0 + l16]
        auto yload = [&](int j, T(&yv)[16]) {
            const int kc = j < nu ? kPanel * (kq + KQ * j) : 0;
            kvec_load<T>(yv, ycol + kc, min(kPanel, nk - kc));
        };
        typename MfmaAcc<T>::type acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = T(0);
        mg_u4u ra[NV], rb[NV];
        T ya[16], yb[16];
        gload(0, ra);
        gload(1, rb);
        yload(0, ya);
        // step j: the copy of R (tile j of the group) to Ls[kq][j % 2], then the loads of tile
        // j + 2 into R and of y(j + 1) into yn, then the products with yc.  Every thread takes
        // every barrier; a group past its tiles only waits.
        auto step = [&](int j, mg_u4u(&R)[NV], T(&yc)[16], T(&yn)[16]) {
            const int b = j & 1;
            if (j < nu)
#pragma unroll
                for (int v = 0; v < NV; ++v) *reinterpret_cast<mg_u4u *>(&Ls[kq][b][lc][lr + VW * v]) = R[v];
            __syncthreads();
            gload(j + 2, R);
            yload(j + 1, yn);
            if (j < nu) {
                T bv[16];
#pragma unroll
                for (int s = 0; s < 16; ++s) bv[s] = Ls[kq][b][16 * (s >> 2) + 4 * lk + (s & 3)][i];
                kvec_mask<T>(yc, min(kPanel, nk - kPanel * (kq + KQ * j)));
                tile_mfma<T>(acc, yc, bv);
            }
        };
        for (int j = 0; j < nj; j += 2) {
            step(j, ra, ya, yb);
            if (j + 1 < nj) step(j + 1, rb, yb, ya);
        }
        T old[4] = {};  // loaded here, in flight during the groups' sum
        if (kq == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                old[q] = Vg[(int64_t)(c0 + min(MfmaAcc<T>::row(lk, q), nr - 1)) * f + r0 + min(i, nrow - 1)];
        if constexpr (KQ > 1) {
            __syncthreads();  // every group is done with Ls
            T(*red)[4][4][64] = reinterpret_cast<T(*)[4][4][64]>(&Ls[0][0][0][0]);  // [group][wave][q][lane]
#pragma unroll
            for (int q = 0; q < 4; ++q) red[kq][wv][q][lane] = acc[q];
            __syncthreads();
            if (kq == 0)
#pragma unroll
                for (int g = 1; g < KQ; ++g)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] += red[g][wv][q][lane];
        }
        if (kq == 0 && i < nrow)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = MfmaAcc<T>::row(lk, q);
                if (rr < nr) Vg[(int64_t)(c0 + rr) * f + r0 + i] = old[q] - acc[q];
            }
        __syncthreads();  // Ls is free for the next pass
    }
}

// Backward, step 1 of block K: x of sub-block b, x[pb + m] = sum_{k >= m} X[k][pb + m] z[k] over
// the block's sub-blocks c >= b (z final: the footer and every block to the right are applied;
// X[pc + k][pb + m] at front row pb + m, column pc + k), as x^T = z^T X; x goes to Y.
template <typename T, int NS, int KS>
__global__ __launch_bounds__(256) void k_bwd_diag(const PanelTask *__restrict__ tasks, const T *__restrict__ arena,
                                                  T *__restrict__ Y, const T *__restrict__ Z, int K, int n, int nrhs,
                                                  const int *__restrict__ cols, int ntask) {
    static_assert(kRhsChunk == 16, "one 16-row MFMA tile of right-hand sides per pass");
    static_assert(KS == 1 || KS == 2 || KS == 4, "1, 2 or 4 waves per output group");
    constexpr int NT = (NS + KS - 1) / KS;
    constexpr int P = NT < kSPre<T> ? NT : kSPre<T>;
    __shared__ T red[KS > 1 ? KS - 1 : 1][4 / KS][4][64];
    int task, grp, ks;
    split_task<KS>(task, grp, ks);
    if (task >= ntask) return;
    const PanelTask tk = tasks[task];
    const int nc = cols ? min(nrhs, cols[tk.sys]) : nrhs;
    if (nc == 0) return;
    const int b = NS == 1 ? 0 : tk.blk, f = tk.f, w = tk.w;
    const int bs = K * kSBlock, be = min(bs + kSBlock, w);
    const int pb = bs + kPanel * b, nbb = min(kPanel, be - pb);
    const int nsub = NS == 1 ? 1 : (be - bs + kPanel - 1) / kPanel;
    const T *M = arena + tk.off;
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const int nu = (nsub - b - ks + KS - 1) / KS;  // this wave's tiles c = b + ks + KS u < nsub
    const int m = 16 * grp + l16;  // this lane's output row of the sub-block (MFMA column)
    if (KS == 1 && 16 * grp >= nbb) return;  // a wave past the sub-block's columns (no barriers when KS = 1)
    const T *xrow = M + pb + min(m, nbb - 1);  // X[k][pb + m] = xrow[k f]
    for (int c0 = 0; c0 < nc; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nc - c0);
        const T *zcol = Z + (int64_t)(c0 + min(l16, nr - 1)) * n + tk.fc;  // z[.][c0 + l16]
        auto load = [&](int u, T(&av)[16], T(&bv)[16]) {
            const int pc = bs + kPanel * (b + ks + KS * u), ncc = min(kPanel, be - pc);
            kvec_load<T>(av, zcol + pc, ncc);
            // masked here: with the sixteen strided loads of X per tile the registers are tight,
            // and the tiles' loads wait for each other anyway (measured faster)
            kvec_mask<T>(av, ncc);
            kstride_load<T>(bv, xrow + (int64_t)pc * f, f, ncc);
        };
        auto prep = [&](int u, T(&)[16], T(&bv)[16]) {
            if (ks + KS * u != 0) return;  // X_bb = Li_b: unit diagonal, zero above it (k < m)
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int k = 16 * (s >> 2) + 4 * lk + (s & 3);
                bv[s] = k > m ? bv[s] : (k == m ? T(1) : T(0));
            }
        };
        typename MfmaAcc<T>::type acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = T(0);
        tile_run<T, NT, P>(acc, nu, load, prep);
        if constexpr (KS > 1) split_reduce<T, KS>(acc, red, ks);
        if (ks == 0 && m < nbb)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = MfmaAcc<T>::row(lk, q);
                if (rr < nr) Y[(int64_t)(c0 + rr) * n + tk.fc + pb + m] = acc[q];
            }
    }
}

// Backward, step 2 of block K: every 64-column block cb to its left, z[64 cb + j] -= sum_k
// L[bs + k][64 cb + j] x[bs + k] over the block's rows, as (z_cb)^T -= x^T L[block rows, cb]
// (both operands along k by vector loads: L down column 64 cb + j).
template <typename T, int NS, int KS>
__global__ __launch_bounds__(256) void k_bwd_upd(const PanelTask *__restrict__ tasks, const T *__restrict__ arena,
                                                 const T *__restrict__ Y, T *__restrict__ Z, int K, int n, int nrhs,
                                                 const int *__restrict__ cols, int ntask) {
    static_assert(kRhsChunk == 16, "one 16-row MFMA tile of right-hand sides");
    static_assert(KS == 1 || KS == 2 || KS == 4, "1, 2 or 4 waves per output group");
    constexpr int NT = (NS + KS - 1) / KS;
    constexpr int P = NT < kSPre<T> ? NT : kSPre<T>;
    __shared__ T red[KS > 1 ? KS - 1 : 1][4 / KS][4][64];
    int task, grp, ks;
    split_task<KS>(task, grp, ks);
    if (task >= ntask) return;
    const PanelTask tk = tasks[task];
    const int nc = cols ? min(nrhs, cols[tk.sys]) : nrhs;
    if (nc == 0) return;
    const int cb = tk.blk, f = tk.f, w = tk.w;
    const int bs = K * kSBlock, be = min(bs + kSBlock, w), nk = be - bs;
    const int nt = NS == 1 ? 1 : (nk + kPanel - 1) / kPanel;
    const T *M = arena + tk.off;
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const int nu = (nt - ks + KS - 1) / KS;             // this wave's tiles ks + KS u < nt
    const int jj = kPanel * cb + 16 * grp + l16;       // this lane's column (MFMA column), < 64 cb + 64 <= bs
    const T *lcol = M + (int64_t)jj * f + bs;          // L[bs + k][jj] = lcol[k]
    for (int c0 = 0; c0 < nc; c0 += kRhsChunk) {
        const int nr = min(kRhsChunk, nc - c0);
        const T *xcol = Y + (int64_t)(c0 + min(l16, nr - 1)) * n + tk.fc + bs;  // x[bs + .][c0 + l16]
        T old[4] = {};
        if (ks == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                old[q] = Z[(int64_t)(c0 + min(MfmaAcc<T>::row(lk, q), nr - 1)) * n + tk.fc + jj];
        auto load = [&](int u, T(&av)[16], T(&bv)[16]) {
            const int kc = kPanel * (ks + KS * u), nkc = min(kPanel, nk - kc);
            kvec_load<T>(av, xcol + kc, nkc);
            kvec_load<T>(bv, lcol + kc, nkc);
        };
        typename MfmaAcc<T>::type acc;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = T(0);
        auto prep = [&](int u, T(&av)[16], T(&bv)[16]) {
            const int nkc = min(kPanel, nk - kPanel * (ks + KS * u));
            kvec_mask<T>(av, nkc);
            kvec_mask<T>(bv, nkc);
        };
        tile_run<T, NT, P>(acc, nu, load, prep);
        if constexpr (KS > 1) split_reduce<T, KS>(acc, red, ks);
        if (ks == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = MfmaAcc<T>::row(lk, q);
                if (rr < nr) Z[(int64_t)(c0 + rr) * n + tk.fc + jj] = old[q] - acc[q];
            }
    }
}

// ---- refinement helpers ------------------------------------------------------

// R = B - A X over the full CSR (LANES lanes per row, fixed xor-tree per row, all
// right-hand sides of a chunk of kResCols per pass over A), and per-block partial
// sums of R^2 and B^2 for the relative residual: part[(col * gridDim.x + block) * 2 + {0,1}].
// One row per LANES-lane group (grid = n * LANES / 256 blocks), so the partial sums and
// their order depend only on n and LANES.
constexpr int kResColsMax = 8;
static inline int residual_blocks(int n, int lanes) { return (int)(((int64_t)n * lanes + 255) / 256); }
template <int LANES, int kResCols, bool kNorms>
__global__ __launch_bounds__(256) void k_residual(const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const double *__restrict__ v, int n, int nrhs,
                                                  const double *__restrict__ B, int ldb, const double *__restrict__ X,
                                                  int ldx, double *__restrict__ R, double *__restrict__ part,
                                                  const int *__restrict__ row_sys, const int *__restrict__ cols) {
    __shared__ double sr[kResCols][256 / 64], sb[kResCols][256 / 64];
    constexpr int RPB = 256 / LANES;  // rows per block pass
    // one system with every column converged: nothing to recompute (k_residual_final keeps
    // the previous norms)
    if (kNorms && cols && cols[0] <= 0) return;
    const int sub = threadIdx.x % LANES;
    for (int c0 = 0; c0 < nrhs; c0 += kResCols) {
        const int nc = min(kResCols, nrhs - c0);
        double rr[kResCols], bb[kResCols];
#pragma unroll
        for (int c = 0; c < kResCols; ++c) rr[c] = bb[c] = 0.0;
        for (int row = blockIdx.x * RPB + threadIdx.x / LANES; row < n; row += gridDim.x * RPB) {
            // batches: the columns past the row's system count are left alone (zero, or
            // converged and unchanged; the norms of those columns are not recomputed)
            const int ncr = cols ? min(nc, cols[row_sys[row]] - c0) : nc;
            if (ncr <= 0) continue;  // the row's LANES lanes together
            double acc[kResCols];
#pragma unroll
            for (int c = 0; c < kResCols; ++c) acc[c] = 0.0;
            // the next entry's value and column are loaded with the current entry's X values
            // (indices clamped to an entry of the matrix, which holds at least one)
            const int ke = rp[row + 1], klast = max(ke - 1, 0);
            int k = rp[row] + sub;
            double a = v[min(k, klast)];
            int64_t col = ci[min(k, klast)];
#pragma unroll 1
            for (; k < ke; k += LANES) {
                const double an = v[min(k + LANES, klast)];
                const int64_t cn = ci[min(k + LANES, klast)];
                // columns past nc re-read the last one (unconditional loads issue together)
#pragma unroll
                for (int c = 0; c < kResCols; ++c) acc[c] += a * X[(int64_t)(c0 + min(c, ncr - 1)) * ldx + col];
                a = an;
                col = cn;
            }
#pragma unroll
            for (int c = 0; c < kResCols; ++c)
                for (int o = LANES / 2; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, LANES);
            if (sub == 0) {
#pragma unroll
                for (int c = 0; c < kResCols; ++c)
                    if (c < ncr) {
                        const double b = B[(int64_t)(c0 + c) * ldb + row];
                        const double res = b - acc[c];
                        R[(int64_t)(c0 + c) * n + row] = res;
                        if (kNorms) {
                            rr[c] += res * res;
                            bb[c] += b * b;
                        }
                    }
            }
        }
        if (!kNorms) continue;  // per-system norms come from R afterwards
        // block reduction: wave xor-tree, then the 4 waves in fixed order
#pragma unroll
        for (int c = 0; c < kResCols; ++c)
            for (int o = 32; o > 0; o >>= 1) {
                rr[c] += __shfl_xor(rr[c], o, 64);
                bb[c] += __shfl_xor(bb[c], o, 64);
            }
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < kResCols; ++c) {
                sr[c][wv] = rr[c];
                sb[c][wv] = bb[c];
            }
        __syncthreads();
        if (threadIdx.x < nc) {
            const int c = threadIdx.x;
            part[((int64_t)(c0 + c) * gridDim.x + blockIdx.x) * 2 + 0] = (sr[c][0] + sr[c][1]) + (sr[c][2] + sr[c][3]);
            part[((int64_t)(c0 + c) * gridDim.x + blockIdx.x) * 2 + 1] = (sb[c][0] + sb[c][1]) + (sb[c][2] + sb[c][3]);
        }
        __syncthreads();
    }
}

// relres[c] = sqrt(sum R^2) / sqrt(sum B^2) per column (one workgroup per column); a column
// past cols[0] (left out of the last solve: converged) keeps its previous value.
__global__ __launch_bounds__(256) void k_residual_final(const double *__restrict__ part, int nblocks,
                                                        double *__restrict__ relres, const int *__restrict__ cols) {
    __shared__ double sr[256], sb[256];
    const int c = blockIdx.x;
    if (cols && c >= cols[0]) return;
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
}

// relres[t * nrhs + c] = ||R[rows of system t, c]|| / ||B[rows of system t, c]|| (one workgroup
// per system and column; ||B|| = 0 counts as 1)
__global__ __launch_bounds__(256) void k_residual_seg(const double *__restrict__ R, int n,
                                                      const double *__restrict__ B, int ldb,
                                                      const int *__restrict__ sys_begin, int nrhs,
                                                      double *__restrict__ relres, const int *__restrict__ cols) {
    __shared__ double sr[256], sb[256];
    const int t = blockIdx.x, c = blockIdx.y;
    if (cols && c >= cols[t]) return;  // column left alone: it keeps its previous value
    double rr = 0.0, bb = 0.0;
    for (int i = sys_begin[t] + threadIdx.x; i < sys_begin[t + 1]; i += 256) {
        const double r = R[(int64_t)c * n + i], b = B[(int64_t)c * ldb + i];
        rr += r * r;
        bb += b * b;
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
        relres[(int64_t)t * nrhs + c] = sqrt(sr[0]) / (bn > 0.0 ? bn : 1.0);
    }
}

// R[rows of system t, all columns] = 0 for the systems whose every column has converged
// (one workgroup per system and column)
// R[c] = 0 on system t's rows for every column c >= cols[t] (columns a correction solve
// leaves out: their correction must be zero).
__global__ __launch_bounds__(256) void k_zero_past(double *__restrict__ R, int n, const int *__restrict__ sys_begin,
                                                   const int *__restrict__ cols) {
    const int t = blockIdx.x, c = blockIdx.y;
    if (c < cols[t]) return;
    for (int i = sys_begin[t] + threadIdx.x; i < sys_begin[t + 1]; i += 256) R[(int64_t)c * n + i] = 0.0;
}

// cols[t] = 1 + the last column with a nonzero right-hand-side entry on system t's rows (0 if
// none): the columns past it are zero (padding of uneven right-hand-side counts) and their
// solutions are zero, so the solves leave them out.
// Four columns per step, last first: each thread's loads of a step are issued together.
__global__ __launch_bounds__(1024) void k_sys_cols(const double *__restrict__ B, int ldb, int nrhs,
                                                   const int *__restrict__ sys_begin, int *__restrict__ cols) {
    __shared__ int top;
    const int t = blockIdx.x, r0 = sys_begin[t], r1 = sys_begin[t + 1];
    for (int c1 = nrhs; c1 > 0; c1 -= 4) {
        const int c0 = max(c1 - 4, 0);
        int hi = 0;  // one past the highest nonzero column of [c0, c1) among this thread's rows
#pragma unroll 2
        for (int i = r0 + threadIdx.x; i < r1; i += 1024) {
            double x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = B[(int64_t)min(c0 + u, c1 - 1) * ldb + i];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (c0 + u < c1 && x[u] != 0.0) hi = max(hi, c0 + u + 1);  // NaN counts as nonzero
        }
        if (threadIdx.x == 0) top = 0;
        __syncthreads();
        if (hi) atomicMax(&top, hi);
        __syncthreads();
        const int h = top;
        if (h) {
            if (threadIdx.x == 0) cols[t] = h;
            return;  // uniform
        }
        __syncthreads();  // every thread has read top before the next step clears it
    }
    if (threadIdx.x == 0) cols[t] = 0;
}

// Start of a solve with refinement: relres = 0 (columns no solve touches are zero), the
// round flags cleared, and for one system every column active.
__global__ __launch_bounds__(256) void k_refine_init(double *__restrict__ rel, int nrel, int *__restrict__ act,
                                                     int nact, int *__restrict__ cols, int single_nrhs) {
    for (int e = threadIdx.x; e < nrel; e += 256) rel[e] = 0.0;
    for (int e = threadIdx.x; e < nact; e += 256) act[e] = 0;
    if (single_nrhs >= 0 && threadIdx.x == 0) cols[0] = single_nrhs;
}

// Before a correction solve: cols[t] = 1 + the last column of system t whose relative
// residual is not <= tol (NaN included), 0 when all of them are; *active = 1 when some
// system has a column left (the round does work).
__global__ __launch_bounds__(1024) void k_refine_cols(const double *__restrict__ rel, int nsys, int nrhs, double tol,
                                                      int *__restrict__ cols, int *__restrict__ active) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    int mine = 0;
    for (int t = threadIdx.x; t < nsys; t += 1024) {
        int c1 = 0;
        for (int c = nrhs - 1; c >= 0; --c)
            if (!(rel[(int64_t)t * nrhs + c] <= tol)) {
                c1 = c + 1;
                break;
            }
        cols[t] = c1;
        mine |= c1 > 0;
    }
    if (mine) any = 1;  // every writer stores the same value
    __syncthreads();
    if (threadIdx.x == 0) *active = any;
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
// propagates NaN; 1 when the maximum is not positive), two-stage over a fixed grid.
__device__ inline double nanmax(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : fmax(a, b); }

__global__ __launch_bounds__(256) void k_diag_scale_part(const int *__restrict__ dpos, const double *__restrict__ vals,
                                                         int n, double *__restrict__ part) {
    __shared__ double sm[256];
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        m = nanmax(m, dpos[i] >= 0 ? fabs(vals[dpos[i]]) : 0.0);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void k_diag_scale_final(const double *__restrict__ part, int nparts,
                                                          double *__restrict__ out) {
    __shared__ double sm[256];
    double m = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) m = nanmax(m, part[b]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0] > 0.0 ? sm[0] : 1.0;
}

// max |A_ii| of every system of a batch (one workgroup per system), 1 when zero
__global__ __launch_bounds__(256) void k_diag_scale_seg(const int *__restrict__ dpos, const double *__restrict__ vals,
                                                        const int *__restrict__ sys_begin, double *__restrict__ out) {
    __shared__ double sm[256];
    const int t = blockIdx.x;
    double m = 0.0;
    for (int i = sys_begin[t] + threadIdx.x; i < sys_begin[t + 1]; i += 256)
        m = nanmax(m, dpos[i] >= 0 ? fabs(vals[dpos[i]]) : 0.0);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] = nanmax(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = sm[0] > 0.0 ? sm[0] : 1.0;
}

// Device workspace of one solve call: every buffer the solve and refinement kernels use
// besides the factor, carved out of one block (sub-buffers 256-byte aligned).  The kernels
// never rely on its contents at entry.
constexpr int kActSlots = 64;  // refinement rounds whose "did work" flags a workspace holds
struct SolveWs {
    void *Y = nullptr, *V = nullptr, *Z = nullptr, *part = nullptr;  // factor precision
    double *R = nullptr, *Part = nullptr, *Rel = nullptr;
    int *Cols = nullptr;  // per system: leading right-hand sides a solve still needs
    int *Act = nullptr;   // per refinement round (mod kActSlots): 1 when the round did work
};
// Bytes of the block for nrhs right-hand sides, and the view of a block at base.
size_t solve_ws_bytes(const Numeric &N, int nrhs);
SolveWs solve_ws_view(const Numeric &N, int nrhs, void *base);

// A pooled solve workspace: its block and the event recorded after the last solve that used
// it (a later solve on another stream waits for that event before touching the block).
struct WsSlot {
    DevArray<unsigned char> buf;
    hipEvent_t done = nullptr;
    ~WsSlot() {
        if (done) (void)hipEventDestroy(done);
    }
};

// A captured solve: the launches of solve_rounds for one set of arguments.
struct SolveGraph {
    const void *b = nullptr, *x = nullptr, *ws = nullptr;
    int ldb = 0, ldx = 0, nrhs = 0, rounds = -1, check = 0;
    double tol = 0.0;
    hipGraphExec_t exec = nullptr;
    bool same(const double *b_, int ldb_, const double *x_, int ldx_, int nrhs_, const void *ws_, int rounds_,
              int check_, double tol_) const {
        return exec && b == b_ && ldb == ldb_ && x == x_ && ldx == ldx_ && nrhs == nrhs_ && ws == ws_ &&
               rounds == rounds_ && check == check_ && tol == tol_;
    }
};

struct Numeric {
    std::shared_ptr<DevicePlan> plan;
    const Symbolic *sym = nullptr;
    int precision = MG_FP64;
    int n = 0;
    DevArray<unsigned char> arena;  // f^2 fronts, element type by precision
    DevArray<double> diag;
    DevArray<unsigned char> flags;
    DevArray<int> counts;
    DevArray<double> mags, scale, wsScale, wsStat;
    DevArray<double> values;  // upper CSR values (float64), kept for refinement
    Matrix A;                 // full CSR for residuals
    mg_numeric_info info{};
    // set only when the last factorization into this handle succeeded: a handle whose
    // factorization threw (e.g. a refused pivot) is never solved with
    bool factorized = false;
    // Solve workspaces, one per concurrent solve: the factor itself is read-only during
    // solves, so solves of one handle from several threads (on their own streams) are
    // independent (cf. UtduNumericFactorization.cs:101,121, per-call scratch).
    std::mutex ws_mu;
    std::vector<std::unique_ptr<WsSlot>> ws_free;
    hipGraphExec_t gexec = nullptr;  // captured numeric factorization (refactorizations replay it)
    hipStream_t cap_stream = nullptr;
    // captured solves (mg_options.device_graph of the factorization): one graph for the last
    // set of solve arguments; the refinement rounds a solve runs before its first host read
    // are the number the previous solve of this handle needed
    bool solve_graph = false;
    std::mutex sg_mu;
    SolveGraph sgraph;
    std::atomic<int> refine_hint{2};
    int graph_precision = -1;
    // pivot options baked into the captured launches; a refactorization with other values
    // captures again
    double graph_tol = 0.0, graph_reg = 0.0;
    int graph_allow = -1;
    hipStream_t side = nullptr;  // lookahead Schur updates (default priority)
    hipEvent_t ev_fork = nullptr, ev_side[2] = {nullptr, nullptr};
    void ensure_side() {
        if (side) return;
        MG_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        MG_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        MG_HIP(hipEventCreateWithFlags(&ev_side[0], hipEventDisableTiming));
        MG_HIP(hipEventCreateWithFlags(&ev_side[1], hipEventDisableTiming));
    }
    ~Numeric() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (sgraph.exec) (void)hipGraphExecDestroy(sgraph.exec);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (hipEvent_t ev : {ev_fork, ev_side[0], ev_side[1]})
            if (ev) (void)hipEventDestroy(ev);
        if (side) (void)hipStreamDestroy(side);
    }
};

// hipGraph capture (mg_options.device_graph): 1 captures the factorization launches and the
// solves, 2 only the solves (the factorization then keeps its lookahead stream).
static inline bool opts_use_graph(const mg_options &o) { return o.device_graph == 1; }
static inline bool opts_graph_solves(const mg_options &o) { return o.device_graph == 1 || o.device_graph == 2; }
static bool profile_on();

constexpr int kScaleParts = 256;

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
        for (const char *nm : {"assemble_pull", "assemble_values", "panel_diag", "panel_rows", "schur", "forward", "backward",
                               "small_fronts"}) {
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
enum { KP_PULL = 0, KP_VALUES, KP_PDIAG, KP_PROWS, KP_SCHUR, KP_FWD, KP_BWD, KP_SMALL };

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

static bool profile_on() { return g_prof.on; }
// MG_NO_LOOKAHEAD=1 in the environment runs every Schur step on the caller's stream (the
// schedule of timed launches), e.g. for per-dispatch profiles without concurrent dispatches
static bool lookahead_allowed() {
    static const bool off = [] {
        const char *v = std::getenv("MG_NO_LOOKAHEAD");
        return v && v[0] == '1';
    }();
    return !off;
}

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
    // Lookahead: the part of a middle or outer Schur step that the next 256-column block's panels do not
    // touch runs on a second stream, overlapping those panels (off while launches are timed)
    const bool look = !profile_on() && lookahead_allowed();
    if (look) N.ensure_side();
    hipEvent_t prev_side = nullptr;
    int side_k = 0;
    for (int L = 0; L < P.nlevels; ++L) {
        const int np = P.lvl_pull_ptr[L + 1] - P.lvl_pull_ptr[L];
        if (np > 0) {
            ProfScope ps_(st, KP_PULL, 0.0, e * P.lvl_pull_elems[L] + 4.0 * P.lvl_pull_index[L]);
            hipLaunchKernelGGL(k_assemble_pull<T>, dim3(np), dim3(256), 0, st, P.pull_tasks.p + P.lvl_pull_ptr[L], arena,
                               P.f.p, P.w.p, P.off.p, P.coff.p, P.child_ptr.p, P.children.p, P.inv_off.p, P.inv_map.p,
                               P.inv_col.p);
        }
        for (int c = 0; c < 3; ++c) {  // fronts of at most 8 / 16 / 32 rows
            const int m0 = P.lvl_small_ptr[4 * L + c], m1 = P.lvl_small_ptr[4 * L + c + 1];
            if (m1 <= m0) continue;
            ProfScope pa_(st, KP_PULL, 0.0, 0.0);
            const int per_block = 4 * (8 >> c);
            auto kern = c == 0 ? k_assemble_small<T, 8> : (c == 1 ? k_assemble_small<T, 16> : k_assemble_small<T, 32>);
            hipLaunchKernelGGL(kern, dim3((m1 - m0 + per_block - 1) / per_block), dim3(256), 0, st,
                               P.small_recs.p + m0, m1 - m0, arena, P.child_recs.p, P.inv_map.p, P.inv_col.p);
        }
        const int64_t na = P.lvl_asm_ptr[L + 1] - P.lvl_asm_ptr[L];
        if (na > 0) {
            ProfScope ps_(st, KP_VALUES, 0.0, (double)na * (8.0 + 4.0 + 8.0 + 2 * e));
            hipLaunchKernelGGL(k_assemble_values<T>, dim3(blocks_for(na, 256)), dim3(256), 0, st,
                               P.asm_pos.p + P.lvl_asm_ptr[L], P.asm_vidx.p + P.lvl_asm_ptr[L], na, N.values.p, arena);
        }
        for (int c = 0; c < 3; ++c) {  // fronts of at most 8 / 16 / 32 rows: factored whole by lane groups
            const int m0 = P.lvl_small_ptr[4 * L + c], m1 = P.lvl_small_ptr[4 * L + c + 1];
            if (m1 <= m0) continue;
            ProfScope pf_(st, KP_SMALL, 0.0, 0.0);
            const int per_block = 4 * (8 >> c);  // 4 waves of 64 / FR fronts
            auto kern = c == 0 ? k_factor_small<T, 8> : (c == 1 ? k_factor_small<T, 16> : k_factor_small<T, 32>);
            hipLaunchKernelGGL(kern, dim3((m1 - m0 + per_block - 1) / per_block), dim3(256), 0, st, P.small_nodes.p + m0,
                               m1 - m0, arena, P.f.p, P.w.p, P.off.p, P.coff.p, P.first.p, N.scale.p, P.front_sys.p,
                               opts.pivot_tolerance, opts.diagonal_regularization, opts.allow_diagonal_regularization,
                               N.diag.p, N.flags.p);
        }
        for (int q = 0; q < P.lvl_npanels[L]; ++q) {
            const int ps = q * kPanel;
            const int t0 = P.lvl_ptask_ptr[L][q], t1 = P.lvl_ptask_ptr[L][q + 1];
            const int b0 = P.lvl_panel_ptr[L][q], b1 = P.lvl_panel_ptr[L][q + 1];
            if (t1 > t0) {
                ProfScope ps_(st, KP_PDIAG, 0.0, 0.0);
                hipLaunchKernelGGL(k_panel<T>, dim3(t1 - t0), dim3(256), 0, st, P.ptasks.p + t0, arena, P.f.p, P.w.p,
                                   P.off.p, P.first.p, ps, N.scale.p, P.front_sys.p, opts.pivot_tolerance,
                                   opts.diagonal_regularization, opts.allow_diagonal_regularization, N.diag.p,
                                   N.flags.p);
            }
            if (b1 > b0) {
                ProfScope ps_(st, KP_PROWS, 0.0, 0.0);
                hipLaunchKernelGGL(k_panel_rows<T>, dim3(b1 - b0), dim3(256), 0, st, P.panel_tasks.p + b0, arena, P.f.p,
                                   P.w.p, P.off.p, P.first.p, ps, N.diag.p);
            }
            const int *b64 = &P.lvl_ub64[L][q * kUb64Stride];
            const int *b128 = &P.lvl_ub128[L][q * kUb128Stride];
            // launches of one part: the inner and middle steps, then the outer steps in one launch
            // per run of consecutive buckets with the same tile shape
            auto schur_part = [&](int part, hipStream_t s) {
                const int *u64 = b64 + part * (kSchurBuckets + 1), *u128 = b128 + part * kSchurBuckets;
                if (u64[1] > u64[0])
                    hipLaunchKernelGGL(k_schur<T>, dim3(u64[1] - u64[0]), dim3(256), 0, s, P.upd_tasks.p + u64[0],
                                       arena, q, N.diag.p);
                const int *o64 = u64 + 1;
                const int *big = P.schur128[sizeof(T) == 4];
                for (int b = 0; b < kSchurBuckets;) {
                    int c = b + 1;
                    while (c < kSchurBuckets && big[c] == big[b]) ++c;
                    if (big[b]) {
                        if (u128[c] > u128[b])
                            hipLaunchKernelGGL(k_schur128<T>, dim3(u128[c] - u128[b]), dim3(256), 0, s,
                                               P.upd128_tasks.p + u128[b], arena, q, N.diag.p);
                    } else if (o64[c] > o64[b]) {
                        hipLaunchKernelGGL(k_schur<T>, dim3(o64[c] - o64[b]), dim3(256), 0, s, P.upd_tasks.p + o64[b],
                                           arena, q, N.diag.p);
                    }
                    b = c;
                }
            };
            const bool has1 = b64[2 * (kSchurBuckets + 1)] > b64[kSchurBuckets + 1];
            if (b64[2 * (kSchurBuckets + 1)] > b64[0]) {
                ProfScope ps_(st, KP_SCHUR, P.lvl_upd_flops[L][q], e * P.lvl_upd_elems[L][q]);
                if (look && has1) {
                    // part 1 on the side stream once this panel's rows are final; part 0, which
                    // writes the columns the previous part 1 also writes, waits for that one
                    MG_HIP(hipEventRecord(N.ev_fork, st));
                    MG_HIP(hipStreamWaitEvent(N.side, N.ev_fork, 0));
                    schur_part(1, N.side);
                    hipEvent_t evb = N.ev_side[side_k & 1];
                    MG_HIP(hipEventRecord(evb, N.side));
                    if (prev_side) MG_HIP(hipStreamWaitEvent(st, prev_side, 0));
                    schur_part(0, st);
                    prev_side = evb;
                    ++side_k;
                } else {
                    // a step past its block writes columns an earlier overlapped part may still write
                    if (prev_side && P.lvl_beyond[L][q]) {
                        MG_HIP(hipStreamWaitEvent(st, prev_side, 0));
                        prev_side = nullptr;
                    }
                    schur_part(0, st);
                    schur_part(1, st);
                }
            }
        }
        // the next level's assembly reads every footer this level wrote
        if (prev_side) {
            MG_HIP(hipStreamWaitEvent(st, prev_side, 0));
            prev_side = nullptr;
        }
    }
    // inverses of the solve blocks' diagonal blocks, by distance of the sub-blocks
    for (int d = 0; d < kSBlock / kPanel - 1; ++d) {
        const int t0 = P.inv_ptr[d], t1 = P.inv_ptr[d + 1];
        if (t1 > t0)
            hipLaunchKernelGGL(k_inv_blk<T>, dim3(t1 - t0), dim3(256), 0, st, P.inv_tasks.p + t0, arena, P.f.p,
                               P.w.p, P.off.p);
    }
    MG_HIP(hipGetLastError());
}

void factorize_device(Symbolic &S, const double *d_values, int precision, const mg_options &opts, hipStream_t st,
                      Numeric *&num) {
    if (precision != MG_FP64 && precision != MG_FP32) throw Error(MG_E_ARG, "Unknown factor precision.");
    ensure_plan(S);
    // the reference's limit (UtduNumericFactorization.cs:298-304): the packed form of the
    // largest front must fit in one array, f (f + 1) / 2 <= int.MaxValue (the device's fused
    // fronts are checked the same way)
    const int64_t max_front_length = (int64_t)S.max_front * (S.max_front + 1) / 2;
    if (max_front_length > INT32_MAX || (int64_t)S.plan->max_f * (S.plan->max_f + 1) / 2 > INT32_MAX)
        throw Error(MG_E_TOO_DENSE,
                    "The largest frontal matrix (" + std::to_string(S.max_front) + " x " + std::to_string(S.max_front) +
                        ") cannot be represented by a single array. The problem is too dense for the direct solver in "
                        "this formulation. Check the model for equations that connect too many unknowns.");
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
            N->scale.alloc(std::max(S.plan->nsys, 1));
            N->wsScale.alloc(kScaleParts);
            N->wsStat.alloc(5 * kStatParts);
            N->values.alloc(std::max(S.matrix_nnz, 1));
            N->A.build(S.n, S.a_rowptr.data(), S.a_colidx.data());
        }
        N->factorized = false;
        N->info = mg_numeric_info{};
        N->info.precision = precision;
        if (S.n > 0) {
            if (S.matrix_nnz > 0)
                MG_HIP(hipMemcpyAsync(N->values.p, d_values, sizeof(double) * S.matrix_nnz, hipMemcpyDeviceToDevice, st));
            N->A.set_values_dev(N->values.p, st);
            const bool graph = opts_use_graph(opts) && !profile_on();
            N->solve_graph = opts_graph_solves(opts);
            const bool same_opts = N->graph_tol == opts.pivot_tolerance && N->graph_reg == opts.diagonal_regularization &&
                                   N->graph_allow == opts.allow_diagonal_regularization;
            if (graph && N->gexec && N->graph_precision == precision && same_opts) {
                MG_HIP(hipGraphLaunch(N->gexec, st));
            } else {
                // capture on a private stream (the null stream cannot be captured); the graph
                // is then launched on the caller's stream
                hipStream_t cs = st;
                if (graph) {
                    if (N->gexec) {
                        (void)hipGraphExecDestroy(N->gexec);
                        N->gexec = nullptr;
                    }
                    if (!N->cap_stream) MG_HIP(hipStreamCreateWithFlags(&N->cap_stream, hipStreamNonBlocking));
                    cs = N->cap_stream;
                    MG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                }
                try {
                    if (S.plan->nsys > 1) {
                        hipLaunchKernelGGL(k_diag_scale_seg, dim3(S.plan->nsys), dim3(256), 0, cs, N->A.dpos.p,
                                           N->A.vals.p, S.plan->sys_begin.p, N->scale.p);
                    } else {
                        hipLaunchKernelGGL(k_diag_scale_part, dim3(kScaleParts), dim3(256), 0, cs, N->A.dpos.p,
                                           N->A.vals.p, S.n, N->wsScale.p);
                        hipLaunchKernelGGL(k_diag_scale_final, dim3(1), dim3(256), 0, cs, N->wsScale.p, kScaleParts,
                                           N->scale.p);
                    }
                    if (precision == MG_FP32) run_factor<float>(*N, opts, cs);
                    else run_factor<double>(*N, opts, cs);
                } catch (...) {
                    if (graph) {
                        hipGraph_t g = nullptr;
                        (void)hipStreamEndCapture(cs, &g);
                        if (g) (void)hipGraphDestroy(g);
                    }
                    throw;
                }
                if (graph) {
                    hipGraph_t g = nullptr;
                    MG_HIP(hipStreamEndCapture(cs, &g));
                    const hipError_t ie = hipGraphInstantiate(&N->gexec, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    MG_HIP(ie);
                    N->graph_precision = precision;
                    N->graph_tol = opts.pivot_tolerance;
                    N->graph_reg = opts.diagonal_regularization;
                    N->graph_allow = opts.allow_diagonal_regularization;
                    MG_HIP(hipGraphLaunch(N->gexec, st));
                }
            }
            hipLaunchKernelGGL(k_pivot_stats_part, dim3(kStatParts), dim3(256), 0, st, N->diag.p, N->flags.p, S.n,
                               N->wsStat.p);
            hipLaunchKernelGGL(k_pivot_stats_final, dim3(1), dim3(256), 0, st, N->wsStat.p, N->counts.p, N->mags.p);
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
        N->factorized = true;
    } catch (...) {
        if (created) delete N;
        throw;
    }
    num = N;
}

static void drop_graph(Numeric *N) {
    if (N->gexec) {
        (void)hipGraphExecDestroy(N->gexec);
        N->gexec = nullptr;
    }
}

template <typename T>
static double factor_ms(Numeric &N, const mg_options &o, hipStream_t st) {
    hipEvent_t e0, e1;
    MG_HIP(hipEventCreate(&e0));
    MG_HIP(hipEventCreate(&e1));
    MG_HIP(hipEventRecord(e0, st));
    run_factor<T>(N, o, st);
    MG_HIP(hipEventRecord(e1, st));
    MG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    MG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

void autotune_device(Symbolic &S, const double *d_values, const mg_options &opts, hipStream_t st, Numeric *&num,
                     int *tiles, double *ms) {
    mg_options o = opts;
    o.device_graph = 0;
    if (num) o.precision = num->precision;  // an existing handle keeps its precision
    // a plain factorization first: it sets the values, the pivot scale and the workspaces
    factorize_device(S, d_values, o.precision, o, st, num);
    Numeric &N = *num;
    DevicePlan &P = *N.plan;
    const bool fp32 = N.precision == MG_FP32;
    int *big = P.schur128[fp32];
    // run_factor rebuilds every front from the values, so repeated runs are independent
    auto best_of_3 = [&] {
        double t = 1e300;
        for (int r = 0; r < 3; ++r) t = std::min(t, fp32 ? factor_ms<float>(N, o, st) : factor_ms<double>(N, o, st));
        return t;
    };
    double cur = best_of_3();
    for (int b = 0; b < kSchurBuckets; ++b) {
        if (P.bucket_tasks[b] == 0) continue;
        const int keep = big[b];
        big[b] = 1 - keep;
        const double alt = best_of_3();
        const double t64 = keep ? alt : cur, t128 = keep ? cur : alt;
        big[b] = t128 < 0.98 * t64 ? 1 : 0;  // 128x128 only for a clear gain
        cur = big[b] ? t128 : t64;
    }
    drop_graph(&N);
    o.device_graph = opts.device_graph;
    factorize_device(S, d_values, N.precision, o, st, num);
    if (tiles)
        for (int b = 0; b < kSchurBuckets; ++b) tiles[b] = big[b] ? 128 : 64;
    if (ms) *ms = cur;
}

void set_schur_tiles(Numeric *N, const int *tiles) {
    for (int b = 0; b < kSchurBuckets; ++b) {
        if (tiles[b] != 64 && tiles[b] != 128) throw Error(MG_E_ARG, "Schur tiles must be 64 or 128.");
        N->plan->schur128[N->precision == MG_FP32][b] = tiles[b] == 128;
    }
    drop_graph(N);
}

void numeric_release(Numeric *num) { delete num; }

void numeric_info(const Numeric *N, mg_numeric_info &info, double *diag_host) {
    info = N->info;
    if (diag_host && N->n > 0) MG_HIP(hipMemcpy(diag_host, N->diag.p, sizeof(double) * N->n, hipMemcpyDeviceToHost));
}

template <typename T>
static void run_solve_once(const Numeric &N, const SolveWs &ws, int nrhs, const double *d_rhs, int ldr, double *d_x,
                           int ldx, int accumulate, hipStream_t st, const int *cols = nullptr) {
    const DevicePlan &P = *N.plan;
    const int n = N.n;
    T *Y = reinterpret_cast<T *>(ws.Y);
    T *V = reinterpret_cast<T *>(ws.V);
    const T *arena = reinterpret_cast<const T *>(N.arena.p);
    const int *rsys = cols ? P.row_sys.p : nullptr;  // cols: batches only
    hipLaunchKernelGGL(k_permute_in<T>, dim3((n + 255) / 256), dim3(256), 0, st, d_rhs, ldr, P.perm.p, n, nrhs, Y,
                       rsys, cols);
    T *Z = reinterpret_cast<T *>(ws.Z);
    T *part = reinterpret_cast<T *>(ws.part);
    const SmallFwdArgs sfa{P.child_recs.p, P.inv_map.p, n, nrhs, cols};
    for (int L = 0; L < P.nlevels; ++L) {
        const int i0 = P.lvl_init_ptr[L], i1 = P.lvl_init_ptr[L + 1];
        const double fl = 2.0 * nrhs * (P.lvl_solve_fac[L] - P.lvl_solve_w[L]);
        const double by = sizeof(T) * P.lvl_solve_fac[L] + 4.0 * P.lvl_solve_f[L] +
                          sizeof(T) * nrhs * (1.5 * P.lvl_solve_f[L] + 0.5 * P.lvl_solve_w[L]);
        ProfScope pi_(st, KP_FWD, fl, by);
        for (int c = 0; c < 4; ++c) {
            const int m0 = P.lvl_small_ptr[4 * L + c], m1 = P.lvl_small_ptr[4 * L + c + 1];
            if (m1 <= m0) continue;
            const int per_block = 4 * (8 >> c);  // 4 waves of 64 / FR fronts
            auto kern = c == 0 ? k_fwd_small<T, 8>
                               : (c == 1 ? k_fwd_small<T, 16> : (c == 2 ? k_fwd_small<T, 32> : k_fwd_small<T, 64>));
            hipLaunchKernelGGL(kern, dim3((m1 - m0 + per_block - 1) / per_block), dim3(256), 0, st,
                               P.small_recs.p + m0, m1 - m0, sfa, arena, V, Y);
        }
        if (i1 > i0)
            hipLaunchKernelGGL(k_fwd_init<T>, dim3(i1 - i0), dim3(256), 0, st, P.init_tasks.p + i0, P.f.p, P.w.p,
                               P.first.p, P.child_ptr.p, P.children.p, P.child_recs.p, P.inv_off.p, P.inv_map.p,
                               P.vbase.p, V, Y, n, nrhs, P.front_sys.p, cols);
        for (int J = 0; 2 * J + 2 < (int)P.lvl_fa_ptr[L].size(); ++J) {
            const int *pa = &P.lvl_fa_ptr[L][2 * J], *pu = &P.lvl_fb_ptr[L][2 * J];
            if (pa[1] > pa[0])
                hipLaunchKernelGGL((k_fwd_diag<T, 1, 1>), dim3(pa[1] - pa[0]), dim3(256), 0, st, P.fa_tasks.p + pa[0],
                                   arena, V, Y, J, n, nrhs, cols, pa[1] - pa[0]);
            if (pa[2] > pa[1])
                hipLaunchKernelGGL((k_fwd_diag<T, kSSub, kSSplit>), dim3(split_grid<kSSplit>(pa[2] - pa[1])), dim3(256), 0, st,
                                   P.fa_tasks.p + pa[1], arena, V, Y, J, n, nrhs, cols, pa[2] - pa[1]);
            if (pu[1] > pu[0])
                hipLaunchKernelGGL((k_fwd_upd<T, 1, 1>), dim3(pu[1] - pu[0]), dim3(256), 0, st, P.fb_tasks.p + pu[0],
                                   arena, V, Y, J, n, nrhs, cols, pu[1] - pu[0]);
            if (pu[2] > pu[1])
            {
                const int nt = pu[2] - pu[1];
                if (sizeof(T) == 4 && nt <= kSolveTasksPerChip)  // few tasks: k split in four groups
                    hipLaunchKernelGGL((k_fwd_upd_wide<T, sizeof(T) == 4 ? 4 : 1>), dim3(nt), dim3(sizeof(T) == 4 ? 1024 : 256),
                                       0, st, P.fb_tasks.p + pu[1], arena, V, Y, J, n, nrhs, cols);
                else
                    hipLaunchKernelGGL((k_fwd_upd_wide<T, 1>), dim3(nt), dim3(256), 0, st, P.fb_tasks.p + pu[1], arena,
                                       V, Y, J, n, nrhs, cols);
            }
        }
    }
    for (int L = P.nlevels - 1; L >= 0; --L) {
        const double fl = 2.0 * nrhs * (P.lvl_solve_fac[L] - P.lvl_solve_w[L]) + nrhs * P.lvl_solve_w[L];
        const double by = sizeof(T) * P.lvl_solve_fac[L] + 4.0 * P.lvl_solve_f[L] +
                          sizeof(T) * nrhs * (1.5 * P.lvl_solve_f[L] + 0.5 * P.lvl_solve_w[L]);
        ProfScope pb_(st, KP_BWD, fl, by);
        const int f0 = P.lvl_foot_ptr[L], f1 = P.lvl_foot_ptr[L + 1];
        if (f1 > f0) {
            const int i0 = P.lvl_init_ptr[L], i1 = P.lvl_init_ptr[L + 1];
            hipLaunchKernelGGL(k_bwd_gather<T>, dim3(i1 - i0), dim3(256), 0, st, P.init_tasks.p + i0, P.f.p, P.w.p,
                               P.pat_ptr.p, P.pat_rows.p, P.vbase.p, Y, V, n, nrhs, P.front_sys.p, cols);
            hipLaunchKernelGGL(k_bwd_foot_part<T>, dim3(f1 - f0), dim3(256), 0, st, P.foot_tasks.p + f0, arena, P.f.p,
                               P.w.p, P.off.p, P.vbase.p, V, part, nrhs, P.front_sys.p, cols);
        }
        const int r0 = P.lvl_footr_ptr[L], r1 = P.lvl_footr_ptr[L + 1];
        if (r1 > r0)
            hipLaunchKernelGGL(k_bwd_foot_red<T>, dim3(r1 - r0), dim3(256), 0, st, P.footr_tasks.p + r0, P.w.p,
                               P.first.p, N.diag.p, Y, part, Z, n, nrhs, P.front_sys.p, cols);
        for (int K = ((int)P.lvl_ba_ptr[L].size() - 1) / 2 - 1; K >= 0; --K) {
            const int *pa = &P.lvl_ba_ptr[L][2 * K], *pu = &P.lvl_bb_ptr[L][2 * K];
            if (pa[1] > pa[0])
                hipLaunchKernelGGL((k_bwd_diag<T, 1, 1>), dim3(pa[1] - pa[0]), dim3(256), 0, st, P.ba_tasks.p + pa[0],
                                   arena, Y, Z, K, n, nrhs, cols, pa[1] - pa[0]);
            if (pa[2] > pa[1])
                hipLaunchKernelGGL((k_bwd_diag<T, kSSub, kSSplit>), dim3(split_grid<kSSplit>(pa[2] - pa[1])), dim3(256),
                                   0, st, P.ba_tasks.p + pa[1], arena, Y, Z, K, n, nrhs, cols, pa[2] - pa[1]);
            if (pu[1] > pu[0])
                hipLaunchKernelGGL((k_bwd_upd<T, 1, 1>), dim3(pu[1] - pu[0]), dim3(256), 0, st, P.bb_tasks.p + pu[0],
                                   arena, Y, Z, K, n, nrhs, cols, pu[1] - pu[0]);
            if (pu[2] > pu[1])
                hipLaunchKernelGGL((k_bwd_upd<T, kSSub, kSSplit>), dim3(split_grid<kSSplit>(pu[2] - pu[1])), dim3(256), 0, st,
                                   P.bb_tasks.p + pu[1], arena, Y, Z, K, n, nrhs, cols, pu[2] - pu[1]);
        }
        for (int c = 0; c < 4; ++c) {
            const int m0 = P.lvl_small_ptr[4 * L + c], m1 = P.lvl_small_ptr[4 * L + c + 1];
            if (m1 <= m0) continue;
            const int per_block = 4 * (8 >> c);
            auto kern = c == 0 ? k_bwd_small<T, 8>
                               : (c == 1 ? k_bwd_small<T, 16> : (c == 2 ? k_bwd_small<T, 32> : k_bwd_small<T, 64>));
            hipLaunchKernelGGL(kern, dim3((m1 - m0 + per_block - 1) / per_block), dim3(256), 0, st,
                               P.small_recs.p + m0, m1 - m0, P.pat_rows.p, arena, N.diag.p, Y, n, nrhs, cols);
        }
    }
    hipLaunchKernelGGL(k_permute_out<T>, dim3((n + 255) / 256), dim3(256), 0, st, Y, P.iperm.p, n, nrhs, d_x, ldx,
                       accumulate, rsys, cols);
    MG_HIP(hipGetLastError());
}

size_t solve_ws_bytes(const Numeric &N, int nrhs) {
    const size_t esz = N.precision == MG_FP32 ? sizeof(float) : sizeof(double);
    const size_t n = std::max(N.n, 1), k = std::max(nrhs, 1), nsys = std::max(N.plan->nsys, 1);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return al(n * k * esz) + al((size_t)std::max<int64_t>(N.plan->vsum, 1) * k * esz) + al(n * k * esz) +
           al((size_t)std::max(N.plan->max_foot_tasks, 1) * kPanel * k * esz) + al(n * k * sizeof(double)) +
           al((size_t)residual_blocks((int)n, N.A.lanes) * k * 2 * sizeof(double)) + al(k * nsys * sizeof(double)) +
           al(nsys * sizeof(int)) + al(kActSlots * sizeof(int));
}

SolveWs solve_ws_view(const Numeric &N, int nrhs, void *base) {
    const size_t esz = N.precision == MG_FP32 ? sizeof(float) : sizeof(double);
    const size_t n = std::max(N.n, 1), k = std::max(nrhs, 1), nsys = std::max(N.plan->nsys, 1);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    unsigned char *p = static_cast<unsigned char *>(base);
    auto take = [&](size_t bytes) {
        void *q = p;
        p += al(bytes);
        return q;
    };
    SolveWs w;
    w.Y = take(n * k * esz);
    w.V = take((size_t)std::max<int64_t>(N.plan->vsum, 1) * k * esz);
    w.Z = take(n * k * esz);
    w.part = take((size_t)std::max(N.plan->max_foot_tasks, 1) * kPanel * k * esz);
    w.R = static_cast<double *>(take(n * k * sizeof(double)));
    w.Part = static_cast<double *>(take((size_t)residual_blocks((int)n, N.A.lanes) * k * 2 * sizeof(double)));
    w.Rel = static_cast<double *>(take(k * nsys * sizeof(double)));
    w.Cols = static_cast<int *>(take(nsys * sizeof(int)));
    w.Act = static_cast<int *>(take(kActSlots * sizeof(int)));
    return w;
}

// A workspace from the handle's pool for one solve on stream st: the smallest free block that
// is large enough, else a free block grown, else a new one.  The block is re-entered into
// the pool (with an event recorded on st) when the lease ends, also after an error; a grown
// block replaces the old one only once its allocation has succeeded.
struct WsLease {
    Numeric &N;
    std::unique_ptr<WsSlot> slot;
    hipStream_t st;
    WsLease(Numeric &num, int nrhs, hipStream_t s) : N(num), st(s) {
        const size_t need = solve_ws_bytes(N, nrhs);
        {
            std::lock_guard<std::mutex> lk(N.ws_mu);
            size_t best = (size_t)-1;
            int pick = -1;
            for (int i = 0; i < (int)N.ws_free.size(); ++i) {
                const size_t have = N.ws_free[i]->buf.n;
                if (have >= need && have < best) best = have, pick = i;
            }
            if (pick < 0 && !N.ws_free.empty()) pick = (int)N.ws_free.size() - 1;
            if (pick >= 0) {
                slot = std::move(N.ws_free[pick]);
                N.ws_free.erase(N.ws_free.begin() + pick);
            }
        }
        if (!slot) slot = std::make_unique<WsSlot>();
        try {
            if (slot->buf.n < need) {
                if (slot->done) MG_HIP(hipEventSynchronize(slot->done));  // its last solve is over
                DevArray<unsigned char> grown(need);
                slot->buf = std::move(grown);
            }
            if (!slot->done) MG_HIP(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming));
            else MG_HIP(hipStreamWaitEvent(st, slot->done, 0));  // the previous solve may run elsewhere
        } catch (...) {
            give_back(false);
            throw;
        }
    }
    void give_back(bool record) {
        if (!slot) return;
        if (record && slot->done) (void)hipEventRecord(slot->done, st);
        std::lock_guard<std::mutex> lk(N.ws_mu);
        N.ws_free.push_back(std::move(slot));
    }
    ~WsLease() { give_back(true); }
};

// ws.Rel = per-column ||B - A X|| / ||B|| (per system and column in a batch); ws.R receives the
// residuals.  Columns past cols[t] (left out of the last solve) keep their previous values.
// Device work only: no host synchronization.
static void residual_dev(const Numeric &N, const SolveWs &ws, int nrhs, const double *d_b, int ldb,
                         const double *d_x, int ldx, hipStream_t st, const int *cols) {
    const int n = N.n;
    // one system: 8 right-hand sides per pass over A with the norm partials; a batch takes its
    // per-system norms from R afterwards, kResColsMax columns per pass, each row only up to its
    // system's count, and rows of at most 4 lanes' worth of entries take 2 lanes each
    const bool seg = N.plan->nsys > 1;
    const int lanes = seg && N.A.lanes == 4 ? 2 : N.A.lanes;
    const int nb = residual_blocks(n, lanes);
    const Matrix &A = N.A;
    const int *rsys = N.plan->row_sys.p;
#define MG_RES(LN)                                                                                                  \
    if (seg)                                                                                                        \
        hipLaunchKernelGGL((k_residual<LN, kResColsMax, false>), dim3(nb), dim3(256), 0, st, A.rowptr.p, A.colidx.p, \
                           A.vals.p, n, nrhs, d_b, ldb, d_x, ldx, ws.R, ws.Part, rsys, cols);                       \
    else                                                                                                            \
        hipLaunchKernelGGL((k_residual<LN, 8, true>), dim3(nb), dim3(256), 0, st, A.rowptr.p, A.colidx.p, A.vals.p, \
                           n, nrhs, d_b, ldb, d_x, ldx, ws.R, ws.Part, rsys, cols)
    switch (lanes) {
        case 2: MG_RES(2); break;
        case 4: MG_RES(4); break;
        case 8: MG_RES(8); break;
        case 16: MG_RES(16); break;
        case 32: MG_RES(32); break;
        default: MG_RES(64); break;
    }
#undef MG_RES
    const int nsys = N.plan->nsys;
    if (nsys > 1)
        hipLaunchKernelGGL(k_residual_seg, dim3(nsys, nrhs), dim3(256), 0, st, ws.R, n, d_b, ldb, N.plan->sys_begin.p,
                           nrhs, ws.Rel, cols);
    else
        hipLaunchKernelGGL(k_residual_final, dim3(nrhs), dim3(256), 0, st, ws.Part, nb, ws.Rel, cols);
}

// One refinement round k: the columns still above tol (per system) are found on the device
// (ws.Act[k mod kActSlots] = 1 when there are any), the residual columns past them are zeroed,
// x += solve(r) on those columns, and their residuals and norms are recomputed.  A round after
// convergence leaves everything as it is (every kernel finds no column to work on).
template <typename Once>
static void refine_round(const Numeric &N, const SolveWs &ws, int nrhs, const double *d_b, int ldb,
                         const double *d_x, int ldx, double tol, int k, hipStream_t st, Once &&once) {
    const DevicePlan &P = *N.plan;
    hipLaunchKernelGGL(k_refine_cols, dim3(1), dim3(1024), 0, st, ws.Rel, P.nsys, nrhs, tol, ws.Cols,
                       ws.Act + k % kActSlots);
    hipLaunchKernelGGL(k_zero_past, dim3(P.nsys, nrhs), dim3(256), 0, st, ws.R, N.n, P.sys_begin.p, ws.Cols);
    once(ws.R, N.n, 1, ws.Cols, st);
    residual_dev(N, ws, nrhs, d_b, ldb, d_x, ldx, st, ws.Cols);
}

// The first solve, its residual and the first `rounds` refinement rounds, with no host
// synchronization: what a captured solve graph holds.
template <typename Once>
static void solve_rounds(const Numeric &N, const SolveWs &ws, int nrhs, const double *d_b, int ldb, double *d_x,
                         int ldx, double tol, bool check, int rounds, hipStream_t st, Once &&once) {
    const DevicePlan &P = *N.plan;
    const int nsys = P.nsys;
    hipLaunchKernelGGL(k_refine_init, dim3(1), dim3(256), 0, st, ws.Rel, nsys * nrhs, ws.Act, kActSlots, ws.Cols,
                       nsys > 1 ? -1 : nrhs);
    // in a batch, every system's solves stop at its last nonzero right-hand side (uneven
    // counts padded to nrhs columns): the columns past it stay zero
    if (nsys > 1)
        hipLaunchKernelGGL(k_sys_cols, dim3(nsys), dim3(1024), 0, st, d_b, ldb, nrhs, P.sys_begin.p, ws.Cols);
    once(d_b, ldb, 0, ws.Cols, st);
    if (!check) return;
    residual_dev(N, ws, nrhs, d_b, ldb, d_x, ldx, st, ws.Cols);
    for (int k = 0; k < rounds; ++k) refine_round(N, ws, nrhs, d_b, ldb, d_x, ldx, tol, k, st, once);
}


// UtduNumericFactorization.Solve (:90-112) for nrhs columns, followed by FP64 iterative
// refinement: x += solve(b - A x) until every column's relres <= tol.  The convergence test is
// made on the device: the first solve and the number of rounds the previous solve of this
// handle needed run without host synchronization (replayed from a captured graph when the
// factorization was captured, mg_options.device_graph); the host reads the norms once, and
// further rounds (each followed by a read) run only if some column is still above tol.  Reads
// the handle only (besides its graph cache); every buffer it writes besides x is in ws.
static void solve_with_ws(Numeric &N, const SolveWs &ws, int nrhs, const double *d_b, int ldb, double *d_x,
                          int ldx, double tol, int max_refine, hipStream_t st, mg_solve_report *rep) {
    auto once = [&](const double *rhs, int ldr, int acc, const int *cols, hipStream_t s) {
        if (N.precision == MG_FP32) run_solve_once<float>(N, ws, nrhs, rhs, ldr, d_x, ldx, acc, s, cols);
        else run_solve_once<double>(N, ws, nrhs, rhs, ldr, d_x, ldx, acc, s, cols);
    };
    const int nsys = N.plan->nsys;
    const bool check = rep || max_refine > 0;
    const int rounds = check ? std::min({std::max(max_refine, 0), N.refine_hint.load(), kActSlots}) : 0;
    bool ran = false;
    if (N.solve_graph && !profile_on()) {
        std::unique_lock<std::mutex> lk(N.sg_mu, std::try_to_lock);  // a concurrent solve runs eagerly
        if (lk.owns_lock()) {
            SolveGraph &G = N.sgraph;
            if (!G.same(d_b, ldb, d_x, ldx, nrhs, ws.Y, rounds, check, tol)) {
                if (G.exec) (void)hipGraphExecDestroy(G.exec);
                G = SolveGraph{};
                if (!N.cap_stream) MG_HIP(hipStreamCreateWithFlags(&N.cap_stream, hipStreamNonBlocking));
                hipStream_t cs = N.cap_stream;
                MG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                try {
                    solve_rounds(N, ws, nrhs, d_b, ldb, d_x, ldx, tol, check, rounds, cs, once);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(cs, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hipGraph_t g = nullptr;
                MG_HIP(hipStreamEndCapture(cs, &g));
                hipGraphExec_t ex = nullptr;
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                MG_HIP(ie);
                G = SolveGraph{d_b, d_x, ws.Y, ldb, ldx, nrhs, rounds, check, tol, ex};
            }
            MG_HIP(hipGraphLaunch(G.exec, st));
            ran = true;
        }
    }
    if (!ran) solve_rounds(N, ws, nrhs, d_b, ldb, d_x, ldx, tol, check, rounds, st, once);
    if (!check) return;
    std::vector<double> rel((size_t)nrhs * nsys);
    int act[kActSlots];
    auto read_back = [&] {
        MG_HIP(hipMemcpyAsync(rel.data(), ws.Rel, sizeof(double) * rel.size(), hipMemcpyDeviceToHost, st));
        MG_HIP(hipMemcpyAsync(act, ws.Act, sizeof(act), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
    };
    auto worst_of = [](const std::vector<double> &v) {
        double m = 0.0;
        for (double x : v) m = (std::isnan(x) || std::isnan(m)) ? NAN : std::max(m, x);
        return m;
    };
    read_back();
    // rounds after the one that found every column converged do no work
    int steps = 0;
    while (steps < rounds && act[steps]) ++steps;
    double worst = worst_of(rel);
    const bool more = steps == rounds;  // every round so far did work
    while (more && steps < max_refine && !(worst <= tol)) {
        refine_round(N, ws, nrhs, d_b, ldb, d_x, ldx, tol, steps, st, once);
        read_back();
        if (!act[steps % kActSlots]) break;
        ++steps;
        worst = worst_of(rel);
    }
    N.refine_hint.store(steps);
    if (rep) {
        rep->refinement_steps = steps;
        rep->relative_residual = worst;
        rep->converged = worst <= tol;
    }
}

static void solve_checks(const Numeric &N, int nrhs, int ldb, int ldx) {
    if (!N.factorized)
        throw Error(MG_E_ARG, "The numeric factorization did not complete; factorize the matrix again before solving.");
    if (nrhs < 0) throw Error(MG_E_ARG, "The number of right-hand sides cannot be negative.");
    if (N.n > 0 && nrhs > 0 && (ldb < N.n || ldx < N.n))
        throw Error(MG_E_ARG, "The right-hand side length does not match the matrix size.");
}

void solve_device(Numeric *Np, int nrhs, const double *d_b, int ldb, double *d_x, int ldx, double tol, int max_refine,
                  hipStream_t st, mg_solve_report *rep) {
    Numeric &N = *Np;
    solve_checks(N, nrhs, ldb, ldx);
    if (rep) *rep = mg_solve_report{0, 1, 0.0};
    if (N.n == 0 || nrhs == 0) return;
    WsLease lease(N, nrhs, st);
    solve_with_ws(N, solve_ws_view(N, nrhs, lease.slot->buf.p), nrhs, d_b, ldb, d_x, ldx, tol, max_refine, st, rep);
}

// ---- allocation-free solve --------------------------------------------------
//
// For callers that cannot allocate device memory (code under stream capture, tight serving
// loops, hosts that forbid it): the solve workspace comes from one file-scope device array
// instead of the handle's pool.  Each device holds its own copy of the array for the life of
// the process; its device address is looked up once per device (hipGetSymbolAddress) and only
// that address is handed to kernels.  Every allocation-free solve on one device shares the
// array: calls that overlap on one device race, and a call larger than the array is refused.
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static void *scratch_base() {
    static std::mutex mu;
    static void *base[64] = {};
    int dev = 0;
    MG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error(MG_E_DEVICE, "Device index out of range for the solve scratch.");
    std::lock_guard<std::mutex> lk(mu);
    if (!base[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(MG_E_DEVICE, std::string("HIP error '") + hipGetErrorString(e) +
                                         "' looking up the device address of the solve scratch.");
        base[dev] = p;
    }
    return base[dev];
}

void solve_device_noalloc(Numeric *Np, int nrhs, const double *d_b, int ldb, double *d_x, int ldx, double tol,
                          int max_refine, hipStream_t st, mg_solve_report *rep) {
    Numeric &N = *Np;
    solve_checks(N, nrhs, ldb, ldx);
    if (rep) *rep = mg_solve_report{0, 1, 0.0};
    if (N.n == 0 || nrhs == 0) return;
    const size_t need = solve_ws_bytes(N, nrhs);
    if (need > kScratchBytes)
        throw Error(MG_E_RANGE, "The allocation-free solve needs " + std::to_string(need) + " bytes of workspace for " +
                                    std::to_string(nrhs) + " right-hand sides, more than its " +
                                    std::to_string(kScratchBytes) + "-byte device scratch; use fewer right-hand "
                                    "sides per call or mg_solve_dev.");
    solve_with_ws(N, solve_ws_view(N, nrhs, scratch_base()), nrhs, d_b, ldb, d_x, ldx, tol, max_refine, st, rep);
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
    const int64_t cnt = (int64_t)fg * P.h_w[g];  // the pivot columns (factor store)
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
    // each 64x64 diagonal block holds Li = L11^{-1} (transposed, upper triangle) and its
    // assembled lower triangle: L11 = Li^{-1}, unit diagonal
    const int wg = P.h_w[g];
    for (int ps = 0; ps < wg; ps += kPanel) {
        const int nb = std::min(kPanel, wg - ps);
        auto li = [&](int r, int c) { return front[(int64_t)(ps + r) * fg + ps + c]; };  // Li[r][c], c < r
        std::vector<double> Lb((size_t)nb * nb, 0.0);
        for (int c = 0; c < nb; ++c) {
            Lb[(size_t)c * nb + c] = 1.0;
            for (int r = c + 1; r < nb; ++r) {
                double acc = li(r, c);
                for (int m = c + 1; m < r; ++m) acc += li(r, m) * Lb[(size_t)m * nb + c];
                Lb[(size_t)r * nb + c] = -acc;
            }
        }
        for (int c = 0; c < nb; ++c)
            for (int r = c; r < nb; ++r) front[(int64_t)(ps + c) * fg + ps + r] = Lb[(size_t)r * nb + c];
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
