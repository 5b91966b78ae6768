/*
 * oracle.c -- CPU restatement of the MAGLA UCHOL / AMD / CGGP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * product library and the CPU timing baseline; the product never links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Reference files (MAGLA repository, CAESolvers/...) are cited per function.
 * Operation order follows the C# exactly so that, compiled with
 * -ffp-contract=off -fwrapv, the FP64 factor equals the reference's bit for
 * bit (the reference's per-element sums never depend on Vector<double>.Count;
 * the one exception, DenseKernels.Dot in the backward solve, is restated for
 * Vector<double>.Count == 4, see or_dot4).
 *
 * Parity pins (tests/test_oracle_*.py): CgComparisonData and workbook values,
 * closed forms, dense LDL^T / solve in numpy, brute-force symbolic
 * elimination for etree / column counts / supernode patterns, hand traces of
 * the reference AMD on small graphs.  "Parity unpinned" (no .NET here):
 * bit-exactness of the AMD permutation against the C# binary itself, and the
 * last bit of the backward-solve dot product.
 */
#include "oracle.h"

#include <float.h>
#include <time.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EMPTY (-1)

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz ? sz : 1); }
static void *xmalloc(size_t n) { return malloc(n ? n : 1); }

/* ======================================================================
 * SymmetricCSRMatrix
 * ====================================================================== */

typedef struct {
    int row, col;
    int64_t order;
    double value;
} coo_item;

static int coo_cmp(const void *pa, const void *pb)
{
    const coo_item *a = (const coo_item *)pa, *b = (const coo_item *)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order ? 1 : 0);
}

/* SymmetricCSRMatrix.cs:36-90 -- normalise (min,max) (:40-52), sum duplicate
 * positions in input order (:54-56, Enumerable.Sum adds sequentially from
 * 0.0), drop |sum| <= 1e-15 (:57, NaN is dropped too), columns ascending
 * (:59), diagonal index table (:66-87).  Out-of-range indices are rejected as
 * SymmetricCSRMatrixBuilder.AddToElement does (SymmetricCSRMatrixBuilder.cs:24-25). */
int or_csr_from_coo(int n, int64_t count, const int *rows, const int *cols,
                    const double *vals, or_csr **out)
{
    if (!out || (count > 0 && (!rows || !cols || !vals))) return OR_E_NULL;
    if (n < 0 || count < 0) return OR_E_ARG;
    for (int64_t k = 0; k < count; k++)
        if (rows[k] < 0 || rows[k] >= n || cols[k] < 0 || cols[k] >= n) return OR_E_RANGE;

    coo_item *items = (coo_item *)xmalloc((size_t)count * sizeof(coo_item));
    for (int64_t k = 0; k < count; k++) {
        int r = rows[k], c = cols[k];
        if (r > c) { int t = r; r = c; c = t; }
        items[k].row = r; items[k].col = c; items[k].order = k; items[k].value = vals[k];
    }
    qsort(items, (size_t)count, sizeof(coo_item), coo_cmp);

    /* reduce groups */
    int64_t kept = 0;
    for (int64_t k = 0; k < count;) {
        int64_t e = k;
        double sum = 0.0;
        while (e < count && items[e].row == items[k].row && items[e].col == items[k].col) {
            sum += items[e].value;
            e++;
        }
        if (fabs(sum) > 1e-15) {
            items[kept] = items[k];
            items[kept].value = sum;
            kept++;
        }
        k = e;
    }

    or_csr *a = (or_csr *)xcalloc(1, sizeof(or_csr));
    a->n = n;
    a->nnz = (int)kept;
    a->rowptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
    a->colidx = (int *)xmalloc((size_t)kept * sizeof(int));
    a->values = (double *)xmalloc((size_t)kept * sizeof(double));
    a->diag = (int *)xmalloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) a->diag[i] = EMPTY;
    for (int64_t k = 0; k < kept; k++) {
        a->rowptr[items[k].row + 1]++;
        a->colidx[k] = items[k].col;
        a->values[k] = items[k].value;
        if (items[k].row == items[k].col) a->diag[items[k].row] = (int)k;
    }
    for (int i = 0; i < n; i++) a->rowptr[i + 1] += a->rowptr[i];
    free(items);
    *out = a;
    return OR_OK;
}

int or_csr_create(int n, const int *rowptr, const int *colidx, const double *values, or_csr **out)
{
    if (!out || !rowptr || (n > 0 && (!colidx || !values))) return OR_E_NULL;
    int nnz = rowptr[n];
    or_csr *a = (or_csr *)xcalloc(1, sizeof(or_csr));
    a->n = n;
    a->nnz = nnz;
    a->rowptr = (int *)xmalloc(((size_t)n + 1) * sizeof(int));
    a->colidx = (int *)xmalloc((size_t)nnz * sizeof(int));
    a->values = (double *)xmalloc((size_t)nnz * sizeof(double));
    a->diag = (int *)xmalloc((size_t)n * sizeof(int));
    memcpy(a->rowptr, rowptr, ((size_t)n + 1) * sizeof(int));
    if (nnz) {
        memcpy(a->colidx, colidx, (size_t)nnz * sizeof(int));
        memcpy(a->values, values, (size_t)nnz * sizeof(double));
    }
    for (int i = 0; i < n; i++) {
        a->diag[i] = EMPTY;
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
            if (colidx[k] == i) a->diag[i] = k;
    }
    *out = a;
    return OR_OK;
}

void or_csr_free(or_csr *a)
{
    if (!a) return;
    free(a->rowptr); free(a->colidx); free(a->values); free(a->diag);
    free(a);
}

/* SymmetricCSRMatrix.cs:293-324 -- row accumulator plus mirrored scatter. */
void or_spmv(const or_csr *a, const double *x, double *y)
{
    int n = a->n;
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int row = 0; row < n; row++) {
        double xrow = x[row];
        double acc = 0.0;
        int end = a->rowptr[row + 1];
        for (int k = a->rowptr[row]; k < end; k++) {
            int col = a->colidx[k];
            double v = a->values[k];
            acc += v * x[col];
            if (col != row) y[col] += v * xrow;
        }
        y[row] += acc;
    }
}

/* SymmetricCSRMatrix.cs:125-167 */
int or_linecross(or_csr *a, double *rhs, double prescribed, int index)
{
    if (!a || !rhs) return OR_E_NULL;
    if (index < 0 || index >= a->n) return OR_E_RANGE;
    int d = a->diag[index];
    if (d < 0) return OR_E_NOT_PD; /* InvalidOperationException: diagonal absent */
    for (int row = 0; row < a->n; row++) {
        for (int pos = a->rowptr[row]; pos < a->rowptr[row + 1]; pos++) {
            int col = a->colidx[pos];
            if (row == col) continue;
            int other;
            if (row == index) other = col;
            else if (col == index) other = row;
            else continue;
            rhs[other] -= a->values[pos] * prescribed;
            a->values[pos] = 0.0;
        }
    }
    a->values[d] = 1.0;
    rhs[index] = prescribed;
    return OR_OK;
}

double or_norm2(int n, const double *v)
{
    /* ResidualCalculator.cs:22-30 */
    double s = 0.0;
    for (int i = 0; i < n; i++) s += v[i] * v[i];
    return sqrt(s);
}

/* ======================================================================
 * AMD -- AMD/ApproximateMinimumDegreeOrdering.cs and its helper classes.
 * Signed arithmetic on wflg / w[] wraps as C# unchecked int does (-fwrapv).
 * ====================================================================== */

typedef struct {
    int n;
    /* FactorGraph (AMD/FactorGraph.cs) */
    int *iw;
    int iwlen;
    int pfree;
    int *pe, *len, *elen;
    int *corder, *ckeys;
    /* WorkSpace arrays */
    int *nv, *degree, *w;
    int wflg, wbig;
    /* DegreeBuckets (AMD/DegreeBuckets.cs) */
    int *dhead, *dnext, *dlast;
    int mindeg;
    /* HashBuckets (AMD/HashBuckets.cs) */
    int *hhead, *hnext, *hprev, *hused;
    int hcount;
    /* MemberLists (AMD/MemberLists.cs) */
    int *mtail, *mnext;
    /* output */
    int *perm;
    int pcur;
    int *dense;
    int ndense;
} amd_state;

static inline int flip(int i) { return -i - 2; }

/* DegreeBuckets.cs:56-69 */
static void deg_insert(amd_state *s, int i, int d)
{
    int following = s->dhead[d];
    if (following != EMPTY) s->dlast[following] = i;
    s->dnext[i] = following;
    s->dlast[i] = EMPTY;
    s->dhead[d] = i;
    if (d < s->mindeg) s->mindeg = d;
}

/* DegreeBuckets.cs:77-89 */
static void deg_remove(amd_state *s, int i)
{
    int d = s->degree[i];
    int previous = s->dlast[i], following = s->dnext[i];
    if (following != EMPTY) s->dlast[following] = previous;
    if (previous != EMPTY) s->dnext[previous] = following;
    else s->dhead[d] = following;
}

/* DegreeBuckets.cs:96-114 */
static int deg_pop_min(amd_state *s, int maxdeg)
{
    for (int d = s->mindeg; d <= maxdeg; d++) {
        int cand = s->dhead[d];
        if (cand == EMPTY) continue;
        s->mindeg = d;
        int following = s->dnext[cand];
        if (following != EMPTY) s->dlast[following] = EMPTY;
        s->dhead[d] = following;
        return cand;
    }
    return EMPTY;
}

/* HashBuckets.cs:54-67 */
static void hash_add(amd_state *s, int64_t hash, int i)
{
    int bucket = (int)(hash % s->n);
    int old = s->hhead[bucket];
    if (old == EMPTY) s->hused[s->hcount++] = bucket;
    else s->hprev[old] = i;
    s->hnext[i] = old;
    s->hprev[i] = EMPTY;
    s->hhead[bucket] = i;
}

/* HashBuckets.cs:86-94 */
static void hash_unlink(amd_state *s, int i)
{
    int p = s->hprev[i], after = s->hnext[i];
    s->hnext[p] = after;
    if (after != EMPTY) s->hprev[after] = p;
}

/* MemberLists.cs:42-46 (head[i] == i for every i, never rewritten) */
static void member_append(amd_state *s, int target, int source)
{
    s->mnext[s->mtail[target]] = source;
    s->mtail[target] = s->mtail[source];
}

/* ApproximateMinimumDegreeOrdering.cs:666-670 */
static void emit(amd_state *s, int rep)
{
    for (int v = rep; v != EMPTY; v = s->mnext[v]) s->perm[s->pcur++] = v;
}

typedef struct { int key, val; } kv_pair;
static int kv_cmp(const void *pa, const void *pb)
{
    int a = ((const kv_pair *)pa)->key, b = ((const kv_pair *)pb)->key;
    return a < b ? -1 : (a > b);
}

/* FactorGraph.cs:95-143 -- live lists (Pe>=0, Len>0) moved left in offset
 * order, then the half-built element; grow when still short of room. */
static int fg_compact(amd_state *s, int *first_entry)
{
    int n = s->n, count = 0;
    kv_pair *pairs = (kv_pair *)xmalloc((size_t)n * sizeof(kv_pair));
    for (int j = 0; j < n; j++)
        if (s->pe[j] >= 0 && s->len[j] > 0) { pairs[count].key = s->pe[j]; pairs[count].val = j; count++; }
    qsort(pairs, (size_t)count, sizeof(kv_pair), kv_cmp);
    int dest = 0;
    for (int t = 0; t < count; t++) {
        int j = pairs[t].val, src = s->pe[j], length = s->len[j];
        if (dest != src) memmove(s->iw + dest, s->iw + src, (size_t)length * sizeof(int));
        s->pe[j] = dest;
        dest += length;
    }
    free(pairs);
    int partial = s->pfree - *first_entry;
    if (partial > 0 && dest != *first_entry)
        memmove(s->iw + dest, s->iw + *first_entry, (size_t)partial * sizeof(int));
    *first_entry = dest;
    s->pfree = dest + partial;
    int required = s->pfree + n + 16;
    if (required > s->iwlen) {
        int grown = s->iwlen + s->iwlen / 2;
        s->iwlen = required > grown ? required : grown;
        int *niw = (int *)realloc(s->iw, (size_t)s->iwlen * sizeof(int));
        if (!niw) return OR_E_NOMEM;
        s->iw = niw;
    }
    return OR_OK;
}

/* ApproximateMinimumDegreeOrdering.cs:647-659 */
static void ensure_mark_capacity(amd_state *s, int required)
{
    if (s->wflg + required < s->wbig) return;
    for (int x = 0; x < s->n; x++)
        if (s->w[x] != 0) s->w[x] = 1;
    s->wflg = 2;
}

/* ApproximateMinimumDegreeOrdering.cs:231-269 */
static int init_degree_lists(amd_state *s)
{
    int n = s->n;
    int dense = (int)(10.0 * sqrt((double)n));
    if (dense < 16) dense = 16;
    if (dense > n) dense = n;
    int eliminated = 0;
    for (int i = 0; i < n; i++) {
        int d = s->degree[i];
        if (d == 0) {
            s->pe[i] = EMPTY; s->w[i] = 0; s->nv[i] = 0; s->elen[i] = EMPTY;
            eliminated++;
            emit(s, i);
        } else if (d > dense) {
            s->pe[i] = EMPTY; s->nv[i] = 0; s->elen[i] = EMPTY;
            eliminated++;
            s->dense[s->ndense++] = i;
        } else {
            deg_insert(s, i, d);
        }
    }
    return eliminated;
}

/* ApproximateMinimumDegreeOrdering.cs:285-304 */
static void build_from_variables(amd_state *s, int pivot, int *edeg, int *fe, int *le)
{
    int degsum = 0, first = s->pe[pivot], last = first - 1;
    int end = first + s->len[pivot];
    for (int p = first; p < end; p++) {
        int i = s->iw[p], nvi = s->nv[i];
        if (nvi <= 0) continue;
        degsum += nvi;
        s->nv[i] = -nvi;
        s->iw[++last] = i;
        deg_remove(s, i);
    }
    *edeg = degsum; *fe = first; *le = last;
}

/* ApproximateMinimumDegreeOrdering.cs:313-393 */
static int build_from_elements(amd_state *s, int pivot, int pivot_elen, int *edeg, int *fe, int *le)
{
    int degsum = 0;
    int p = s->pe[pivot];
    int first = s->pfree;
    int pvc = s->len[pivot] - pivot_elen;
    for (int outer = 1; outer <= pivot_elen + 1; outer++) {
        int e, cursor, count;
        int own = outer > pivot_elen;
        if (own) { e = pivot; cursor = p; count = pvc; }
        else { e = s->iw[p++]; cursor = s->pe[e]; count = s->len[e]; }
        for (int inner = 1; inner <= count; inner++) {
            int i = s->iw[cursor++], nvi = s->nv[i];
            if (nvi <= 0) continue;
            if (s->pfree >= s->iwlen) {
                if (!own) {
                    s->pe[pivot] = p;
                    s->len[pivot] = (pivot_elen - outer) + pvc;
                    if (s->len[pivot] == 0) s->pe[pivot] = EMPTY;
                }
                s->pe[e] = cursor;
                s->len[e] = count - inner;
                if (s->len[e] == 0) s->pe[e] = EMPTY;
                int rc = fg_compact(s, &first);
                if (rc) return rc;
                cursor = s->pe[e];
                p = own ? cursor : s->pe[pivot];
            }
            degsum += nvi;
            s->nv[i] = -nvi;
            s->iw[s->pfree++] = i;
            deg_remove(s, i);
        }
        if (e != pivot) { s->pe[e] = flip(pivot); s->w[e] = 0; }
    }
    *edeg = degsum; *fe = first; *le = s->pfree - 1;
    return OR_OK;
}

/* ApproximateMinimumDegreeOrdering.cs:402-428 */
static void element_set_differences(amd_state *s, int fe, int le)
{
    for (int pos = fe; pos <= le; pos++) {
        int i = s->iw[pos], ec = s->elen[i];
        if (ec <= 0) continue;
        int nvi = -s->nv[i], wnvi = s->wflg - nvi;
        int end = s->pe[i] + ec;
        for (int p = s->pe[i]; p < end; p++) {
            int e = s->iw[p], we = s->w[e];
            if (we >= s->wflg) we -= nvi;
            else if (we != 0) we = s->degree[e] + wnvi;
            s->w[e] = we;
        }
    }
}

/* ApproximateMinimumDegreeOrdering.cs:436-522 */
static void approximate_degrees(amd_state *s, int pivot, int fe, int le, int *edeg, int *pnv, int *elim)
{
    s->hcount = 0;
    for (int pos = fe; pos <= le; pos++) {
        int i = s->iw[pos];
        int ls = s->pe[i], eend = ls + s->elen[i], wr = ls;
        int64_t hash = 0;
        int d = 0;
        for (int p = ls; p < eend; p++) {
            int e = s->iw[p];
            if (s->w[e] == 0) continue;
            int ext = s->w[e] - s->wflg;
            if (ext > 0) { d += ext; s->iw[wr++] = e; hash += e; }
            else if (ext == 0) { s->pe[e] = flip(pivot); s->w[e] = 0; }
        }
        s->elen[i] = wr - ls + 1;
        int vs = wr, lend = ls + s->len[i];
        for (int p = eend; p < lend; p++) {
            int j = s->iw[p], nvj = s->nv[j];
            if (nvj <= 0) continue;
            d += nvj;
            s->iw[wr++] = j;
            hash += j;
        }
        if (s->elen[i] == 1 && vs == wr) {
            s->pe[i] = flip(pivot);
            int nvi = -s->nv[i];
            *edeg -= nvi; *pnv += nvi; *elim += nvi;
            s->nv[i] = 0;
            s->elen[i] = EMPTY;
            member_append(s, pivot, i);
        } else {
            if (d < s->degree[i]) s->degree[i] = d;
            s->iw[wr] = s->iw[vs];
            s->iw[vs] = s->iw[ls];
            s->iw[ls] = pivot;
            s->len[i] = wr - ls + 1;
            hash_add(s, hash, i);
        }
    }
}

/* ApproximateMinimumDegreeOrdering.cs:531-592 */
static void detect_indistinguishable(amd_state *s)
{
    for (int b = 0; b < s->hcount; b++) {
        int bucket = s->hused[b];
        for (int i = s->hhead[bucket]; i != EMPTY; i = s->hnext[i]) {
            if (s->nv[i] >= 0) continue;
            int length = s->len[i], ec = s->elen[i];
            int start = s->pe[i] + 1, end = s->pe[i] + length;
            for (int p = start; p < end; p++) s->w[s->iw[p]] = s->wflg;
            int j = s->hnext[i];
            while (j != EMPTY) {
                int same = s->nv[j] < 0 && s->len[j] == length && s->elen[j] == ec;
                if (same) {
                    int js = s->pe[j] + 1, je = s->pe[j] + length;
                    for (int p = js; p < je; p++)
                        if (s->w[s->iw[p]] != s->wflg) { same = 0; break; }
                }
                int next = s->hnext[j];
                if (same) {
                    s->pe[j] = flip(i);
                    s->nv[i] += s->nv[j];
                    s->nv[j] = 0;
                    s->elen[j] = EMPTY;
                    member_append(s, i, j);
                    hash_unlink(s, j);
                }
                j = next;
            }
            s->wflg++;
        }
        s->hhead[bucket] = EMPTY;
    }
}

/* ApproximateMinimumDegreeOrdering.cs:599-640 */
static void restore_degree_lists(amd_state *s, int pivot, int pivot_elen, int edeg, int pnv, int elim, int fe, int le)
{
    int wr = fe, remaining = s->n - elim;
    for (int pos = fe; pos <= le; pos++) {
        int i = s->iw[pos], nvi = -s->nv[i];
        if (nvi <= 0) continue;
        s->nv[i] = nvi;
        int d = s->degree[i] + edeg - nvi;
        if (remaining - nvi < d) d = remaining - nvi;
        if (d < 0) d = 0;
        s->degree[i] = d;
        deg_insert(s, i, d);
        s->iw[wr++] = i;
    }
    s->nv[pivot] = pnv;
    s->len[pivot] = wr - fe;
    if (s->len[pivot] == 0) { s->pe[pivot] = EMPTY; s->w[pivot] = 0; }
    if (pivot_elen != 0) s->pfree = wr;
}

/* ApproximateMinimumDegreeOrdering.cs:71-84 + WorkSpace ctor (:129-155),
 * Run (:157-224); SymmetricPatternGraph.FromMatrix (AMD/SymmetricPatternGraph.cs:71-121);
 * FactorGraph ctor (AMD/FactorGraph.cs:48-78). */
int or_amd(int n, const int *rowptr, const int *colidx, int *perm_out)
{
    if (n == 0) return OR_OK;
    /* pattern graph: degree pass then placement pass, both in CSR order */
    int *gptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
    for (int row = 0; row < n; row++)
        for (int k = rowptr[row]; k < rowptr[row + 1]; k++) {
            int col = colidx[k];
            if (col == row) continue;
            gptr[row + 1]++;
            gptr[col + 1]++;
        }
    for (int i = 0; i < n; i++) gptr[i + 1] += gptr[i];
    int edges = gptr[n];
    int *cursor = (int *)xmalloc((size_t)n * sizeof(int));
    memcpy(cursor, gptr, (size_t)n * sizeof(int));

    amd_state S;
    memset(&S, 0, sizeof(S));
    amd_state *s = &S;
    s->n = n;
    s->iwlen = edges + edges / 5 + n + 16;
    s->iw = (int *)xmalloc((size_t)s->iwlen * sizeof(int));
    for (int row = 0; row < n; row++)
        for (int k = rowptr[row]; k < rowptr[row + 1]; k++) {
            int col = colidx[k];
            if (col == row) continue;
            s->iw[cursor[row]++] = col;
            s->iw[cursor[col]++] = row;
        }
    free(cursor);
    s->pfree = edges;
    s->pe = (int *)xmalloc((size_t)n * sizeof(int));
    s->len = (int *)xmalloc((size_t)n * sizeof(int));
    s->elen = (int *)xcalloc((size_t)n, sizeof(int));
    s->nv = (int *)xmalloc((size_t)n * sizeof(int));
    s->degree = (int *)xmalloc((size_t)n * sizeof(int));
    s->w = (int *)xmalloc((size_t)n * sizeof(int));
    s->dhead = (int *)xmalloc(((size_t)n + 1) * sizeof(int));
    s->dnext = (int *)xmalloc((size_t)n * sizeof(int));
    s->dlast = (int *)xmalloc((size_t)n * sizeof(int));
    s->hhead = (int *)xmalloc((size_t)n * sizeof(int));
    s->hnext = (int *)xmalloc((size_t)n * sizeof(int));
    s->hprev = (int *)xmalloc((size_t)n * sizeof(int));
    s->hused = (int *)xmalloc((size_t)n * sizeof(int));
    s->mtail = (int *)xmalloc((size_t)n * sizeof(int));
    s->mnext = (int *)xmalloc((size_t)n * sizeof(int));
    s->dense = (int *)xmalloc((size_t)n * sizeof(int));
    s->perm = perm_out;
    for (int i = 0; i <= n; i++) s->dhead[i] = EMPTY;
    for (int i = 0; i < n; i++) {
        s->pe[i] = gptr[i];
        s->len[i] = gptr[i + 1] - gptr[i];
        s->nv[i] = 1;
        s->w[i] = 1;
        s->degree[i] = s->len[i];
        s->dnext[i] = s->dlast[i] = EMPTY;
        s->hhead[i] = s->hnext[i] = s->hprev[i] = EMPTY;
        s->mtail[i] = i;
        s->mnext[i] = EMPTY;
    }
    free(gptr);
    s->wbig = 2147483647 - n - 1;
    s->mindeg = 0;

    int rc = OR_OK;
    int eliminated = init_degree_lists(s);
    s->wflg = 2;
    s->mindeg = 0;
    int max_element = 0;
    while (eliminated < n) {
        ensure_mark_capacity(s, 2 * n + 2);
        int pivot = deg_pop_min(s, n);
        if (pivot == EMPTY) { rc = OR_E_INTERNAL; break; }
        int pivot_elen = s->elen[pivot];
        int pnv = s->nv[pivot];
        eliminated += pnv;
        s->nv[pivot] = -pnv;
        int edeg, fe, le;
        if (pivot_elen == 0) build_from_variables(s, pivot, &edeg, &fe, &le);
        else {
            rc = build_from_elements(s, pivot, pivot_elen, &edeg, &fe, &le);
            if (rc) break;
        }
        s->pe[pivot] = fe;
        s->len[pivot] = le - fe + 1;
        element_set_differences(s, fe, le);
        approximate_degrees(s, pivot, fe, le, &edeg, &pnv, &eliminated);
        s->degree[pivot] = edeg;
        s->elen[pivot] = flip(pnv + edeg);
        if (edeg > max_element) max_element = edeg;
        detect_indistinguishable(s);
        restore_degree_lists(s, pivot, pivot_elen, edeg, pnv, eliminated, fe, le);
        emit(s, pivot);
        s->wflg += max_element + 1;
    }
    if (rc == OR_OK) {
        for (int t = 0; t < s->ndense; t++) emit(s, s->dense[t]);
        if (s->pcur != n) rc = OR_E_INTERNAL;
    }
    free(s->iw); free(s->pe); free(s->len); free(s->elen); free(s->nv); free(s->degree); free(s->w);
    free(s->dhead); free(s->dnext); free(s->dlast); free(s->hhead); free(s->hnext); free(s->hprev);
    free(s->hused); free(s->mtail); free(s->mnext); free(s->dense);
    return rc;
}

/* ======================================================================
 * Symbolic phase -- UCHOL/UtduSymbolicFactorization.cs, PermutedSymmetricPattern.cs,
 * EliminationTree.cs, SupernodalStructure.cs
 * ====================================================================== */

typedef struct {
    int n;
    int *ptr, *rows, *lower_start, *lvptr, *lvidx, *dvidx;
} perm_pattern;

static int int_cmp(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return x < y ? -1 : (x > y);
}

/* PermutedSymmetricPattern.cs:81-185 */
static void pattern_create(const or_csr *a, const int *iperm, perm_pattern *pp)
{
    int n = a->n;
    pp->n = n;
    pp->ptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
    pp->lvptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
    pp->dvidx = (int *)xmalloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) pp->dvidx[i] = EMPTY;
    for (int row = 0; row < n; row++) {
        int prow = iperm[row];
        for (int k = a->rowptr[row]; k < a->rowptr[row + 1]; k++) {
            int col = a->colidx[k];
            if (col == row) { pp->dvidx[prow] = k; continue; }
            int pcol = iperm[col];
            pp->ptr[prow + 1]++;
            pp->ptr[pcol + 1]++;
            pp->lvptr[(prow < pcol ? prow : pcol) + 1]++;
        }
    }
    for (int j = 0; j < n; j++) { pp->ptr[j + 1] += pp->ptr[j]; pp->lvptr[j + 1] += pp->lvptr[j]; }
    pp->rows = (int *)xmalloc((size_t)pp->ptr[n] * sizeof(int));
    pp->lvidx = (int *)xmalloc((size_t)pp->lvptr[n] * sizeof(int));
    int *lrows = (int *)xmalloc((size_t)pp->lvptr[n] * sizeof(int));
    int *cur = (int *)xmalloc((size_t)n * sizeof(int));
    int *lcur = (int *)xmalloc((size_t)n * sizeof(int));
    memcpy(cur, pp->ptr, (size_t)n * sizeof(int));
    memcpy(lcur, pp->lvptr, (size_t)n * sizeof(int));
    for (int row = 0; row < n; row++) {
        int prow = iperm[row];
        for (int k = a->rowptr[row]; k < a->rowptr[row + 1]; k++) {
            int col = a->colidx[k];
            if (col == row) continue;
            int pcol = iperm[col];
            pp->rows[cur[prow]++] = pcol;
            pp->rows[cur[pcol]++] = prow;
            int low = prow < pcol ? prow : pcol, high = prow + pcol - low;
            int pos = lcur[low]++;
            lrows[pos] = high;
            pp->lvidx[pos] = k;
        }
    }
    /* column sorts (:485-496); keys are unique, so any sort gives one answer */
    kv_pair *tmp = NULL;
    size_t tmpcap = 0;
    for (int j = 0; j < n; j++) {
        int st = pp->ptr[j], ln = pp->ptr[j + 1] - st;
        if (ln > 1) qsort(pp->rows + st, (size_t)ln, sizeof(int), int_cmp);
        int ls = pp->lvptr[j], ll = pp->lvptr[j + 1] - ls;
        if (ll > 1) {
            if ((size_t)ll > tmpcap) { tmpcap = (size_t)ll; tmp = (kv_pair *)realloc(tmp, tmpcap * sizeof(kv_pair)); }
            for (int t = 0; t < ll; t++) { tmp[t].key = lrows[ls + t]; tmp[t].val = pp->lvidx[ls + t]; }
            qsort(tmp, (size_t)ll, sizeof(kv_pair), kv_cmp);
            for (int t = 0; t < ll; t++) { lrows[ls + t] = tmp[t].key; pp->lvidx[ls + t] = tmp[t].val; }
        }
    }
    free(tmp);
    pp->lower_start = (int *)xmalloc((size_t)n * sizeof(int));
    for (int j = 0; j < n; j++) pp->lower_start[j] = pp->ptr[j + 1] - (pp->lvptr[j + 1] - pp->lvptr[j]);
    free(lrows); free(cur); free(lcur);
}

static void pattern_free(perm_pattern *pp)
{
    free(pp->ptr); free(pp->rows); free(pp->lower_start); free(pp->lvptr); free(pp->lvidx); free(pp->dvidx);
}

/* EliminationTree.cs:39-71 */
static void etree_build(int n, const int *ptr, const int *rows, int *parent)
{
    int *anc = (int *)xmalloc((size_t)n * sizeof(int));
    for (int k = 0; k < n; k++) {
        parent[k] = EMPTY;
        anc[k] = EMPTY;
        for (int p = ptr[k]; p < ptr[k + 1]; p++) {
            int i = rows[p];
            if (i >= k) break;
            while (i != EMPTY && i < k) {
                int next = anc[i];
                anc[i] = k;
                if (next == EMPTY) parent[i] = k;
                i = next;
            }
        }
    }
    free(anc);
}

/* EliminationTree.cs:80-139 */
static void etree_postorder(int n, const int *parent, int *post)
{
    int *head = (int *)xmalloc((size_t)n * sizeof(int));
    int *next = (int *)xmalloc((size_t)n * sizeof(int));
    int *stack = (int *)xmalloc((size_t)n * sizeof(int));
    for (int j = 0; j < n; j++) head[j] = EMPTY;
    for (int j = n - 1; j >= 0; j--) {
        if (parent[j] == EMPTY) continue;
        next[j] = head[parent[j]];
        head[parent[j]] = j;
    }
    int written = 0;
    for (int j = 0; j < n; j++) {
        if (parent[j] != EMPTY) continue;
        int top = 0;
        stack[0] = j;
        while (top >= 0) {
            int node = stack[top], child = head[node];
            if (child == EMPTY) { top--; post[written++] = node; }
            else { head[node] = next[child]; stack[++top] = child; }
        }
    }
    free(head); free(next); free(stack);
}

/* EliminationTree.cs:152-249 (Gilbert-Ng-Peyton) */
static void etree_colcounts(int n, const int *ptr, const int *rows, const int *parent, const int *post, int *delta)
{
    int *anc = (int *)xmalloc((size_t)n * sizeof(int));
    int *maxfirst = (int *)xmalloc((size_t)n * sizeof(int));
    int *prevleaf = (int *)xmalloc((size_t)n * sizeof(int));
    int *first = (int *)xmalloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) { maxfirst[i] = EMPTY; prevleaf[i] = EMPTY; first[i] = EMPTY; }
    for (int k = 0; k < n; k++) {
        int j = post[k];
        delta[j] = first[j] == EMPTY ? 1 : 0;
        for (; j != EMPTY && first[j] == EMPTY; j = parent[j]) first[j] = k;
    }
    for (int i = 0; i < n; i++) anc[i] = i;
    for (int k = 0; k < n; k++) {
        int j = post[k];
        if (parent[j] != EMPTY) delta[parent[j]]--;
        for (int p = ptr[j]; p < ptr[j + 1]; p++) {
            int i = rows[p];
            if (i <= j || first[j] <= maxfirst[i]) continue;
            maxfirst[i] = first[j];
            int jprev = prevleaf[i];
            prevleaf[i] = j;
            if (jprev == EMPTY) { delta[j]++; continue; }
            int root = jprev;
            while (root != anc[root]) root = anc[root];
            for (int q = jprev; q != root;) { int nx = anc[q]; anc[q] = root; q = nx; }
            delta[j]++;
            delta[root]--;
        }
        if (parent[j] != EMPTY) anc[j] = parent[j];
    }
    for (int j = 0; j < n; j++)
        if (parent[j] != EMPTY) delta[parent[j]] += delta[j];
    free(anc); free(maxfirst); free(prevleaf); free(first);
}

/* SupernodalStructure.cs:271-286 */
static int merge_worthwhile(int columns, int64_t added, int64_t stored)
{
    if (added <= 0) return 1;
    double fraction = (double)added / (double)stored;
    if (columns <= 4) return fraction <= 0.8;
    if (columns <= 16) return fraction <= 0.1;
    if (columns <= 48) return fraction <= 0.05;
    return 0;
}

void or_options_default(or_options *o)
{
    o->ordering = 1;
    o->amalgamation_limit = 64;
    o->panel_width = 32;
    o->allow_regularization = 1;
    o->pivot_tolerance = 1e-13;
    o->regularization = 1e-10;
    o->time_budget_s = 0.0;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void or_symbolic_free(or_symbolic *s)
{
    if (!s) return;
    free(s->amd); free(s->perm); free(s->iperm); free(s->etree); free(s->colcount);
    free(s->pptr); free(s->prows); free(s->lower_start); free(s->lvptr); free(s->lvidx); free(s->dvidx);
    free(s->first); free(s->super_of); free(s->sparent); free(s->child_ptr); free(s->children);
    free(s->pat_ptr); free(s->pat_rows); free(s->block_off);
    free(s);
}

/* UtduSymbolicFactorization.Analyze (UtduSymbolicFactorization.cs:240-295),
 * SupernodalStructure.Build (:102-166), FindFundamentalSupernodes (:175-193),
 * Amalgamate (:203-262), BuildPatterns (:296-355), ComputeCosts (:370-418),
 * FactorStorageLayout.Build size check (FactorStorageLayout.cs:67-76). */
int or_analyze(const or_csr *a, const or_options *o, or_symbolic **out)
{
    if (!a || !o || !out) return OR_E_NULL;
    int n = a->n;
    or_symbolic *s = (or_symbolic *)xcalloc(1, sizeof(or_symbolic));
    s->n = n;
    s->matrix_nnz = a->nnz;
    s->amd = (int *)xmalloc((size_t)n * sizeof(int));
    if (o->ordering == 1) {
        int rc = or_amd(n, a->rowptr, a->colidx, s->amd);
        if (rc) { or_symbolic_free(s); return rc; }
    } else {
        for (int i = 0; i < n; i++) s->amd[i] = i;
    }
    int *inv = (int *)xmalloc((size_t)n * sizeof(int));
    for (int k = 0; k < n; k++) inv[s->amd[k]] = k;
    perm_pattern pp;
    pattern_create(a, inv, &pp);
    int *par1 = (int *)xmalloc((size_t)n * sizeof(int));
    int *post = (int *)xmalloc((size_t)n * sizeof(int));
    etree_build(n, pp.ptr, pp.rows, par1);
    etree_postorder(n, par1, post);
    pattern_free(&pp);
    s->perm = (int *)xmalloc((size_t)n * sizeof(int));
    s->iperm = (int *)xmalloc((size_t)n * sizeof(int));
    for (int k = 0; k < n; k++) s->perm[k] = s->amd[post[k]];
    for (int k = 0; k < n; k++) s->iperm[s->perm[k]] = k;
    free(inv); free(par1);

    pattern_create(a, s->iperm, &pp);
    s->pptr = pp.ptr; s->prows = pp.rows; s->lower_start = pp.lower_start;
    s->lvptr = pp.lvptr; s->lvidx = pp.lvidx; s->dvidx = pp.dvidx;
    s->etree = (int *)xmalloc((size_t)n * sizeof(int));
    s->colcount = (int *)xmalloc((size_t)n * sizeof(int));
    etree_build(n, s->pptr, s->prows, s->etree);
    etree_postorder(n, s->etree, post);
    etree_colcounts(n, s->pptr, s->prows, s->etree, post, s->colcount);
    free(post);

    int amalg = o->amalgamation_limit > 1 ? o->amalgamation_limit : 1;
    const int *parent = s->etree, *cc = s->colcount;
    int *childcount = (int *)xcalloc((size_t)n, sizeof(int));
    for (int j = 0; j < n; j++) if (parent[j] >= 0) childcount[parent[j]]++;
    int *fstart = (int *)xmalloc(((size_t)n + 1) * sizeof(int));
    int nf = 0;
    for (int j = 0; j < n; j++)
        if (j == 0 || parent[j - 1] != j || childcount[j] > 1 || cc[j - 1] != cc[j] + 1) fstart[nf++] = j;
    fstart[nf] = n;
    free(childcount);

    int64_t *ccp = (int64_t *)xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int j = 0; j < n; j++) ccp[j + 1] = ccp[j] + cc[j];
    int *bounds = (int *)xmalloc(((size_t)n + 1) * sizeof(int));
    int *fronts = (int *)xmalloc(((size_t)n + 1) * sizeof(int));
    int ns = 0;
    for (int idx = 0; idx < nf;) {
        int first = fstart[idx], last = fstart[idx + 1] - 1;
        int columns = last - first + 1;
        int fsize = cc[first];
        int64_t nz = ccp[last + 1] - ccp[first];
        while (idx + 1 < nf) {
            if (parent[last] != last + 1) break;
            int nlast = fstart[idx + 2] - 1, nfirst = fstart[idx + 1];
            int mcols = columns + (nlast - nfirst + 1);
            if (mcols > amalg) break;
            int64_t mfront = (int64_t)columns + cc[nfirst];
            int64_t stored = mfront * mcols - (int64_t)mcols * (mcols - 1) / 2;
            int64_t mnz = nz + (ccp[nlast + 1] - ccp[nfirst]);
            if (!merge_worthwhile(mcols, stored - mnz, stored)) break;
            last = nlast; columns = mcols; fsize = (int)mfront; nz = mnz;
            idx++;
        }
        bounds[ns] = first;
        fronts[ns] = fsize;
        ns++;
        idx++;
    }
    bounds[ns] = n;
    free(fstart); free(ccp);

    s->nsuper = ns;
    s->first = (int *)xmalloc(((size_t)ns + 1) * sizeof(int));
    memcpy(s->first, bounds, ((size_t)ns + 1) * sizeof(int));
    free(bounds);
    s->super_of = (int *)xmalloc((size_t)n * sizeof(int));
    for (int q = 0; q < ns; q++)
        for (int j = s->first[q]; j < s->first[q + 1]; j++) s->super_of[j] = q;
    s->sparent = (int *)xmalloc((size_t)ns * sizeof(int));
    int *scc = (int *)xcalloc((size_t)ns, sizeof(int));
    for (int q = 0; q < ns; q++) {
        int pc = parent[s->first[q + 1] - 1];
        s->sparent[q] = pc < 0 ? -1 : s->super_of[pc];
        if (s->sparent[q] >= 0) scc[s->sparent[q]]++;
    }
    s->child_ptr = (int *)xcalloc((size_t)ns + 1, sizeof(int));
    for (int q = 0; q < ns; q++) s->child_ptr[q + 1] = s->child_ptr[q] + scc[q];
    s->children = (int *)xmalloc((size_t)s->child_ptr[ns] * sizeof(int));
    int *ccur = (int *)xmalloc((size_t)ns * sizeof(int));
    memcpy(ccur, s->child_ptr, (size_t)ns * sizeof(int));
    for (int q = 0; q < ns; q++)
        if (s->sparent[q] >= 0) s->children[ccur[s->sparent[q]]++] = q;
    free(ccur); free(scc);
    s->pat_ptr = (int *)xcalloc((size_t)ns + 1, sizeof(int));
    for (int q = 0; q < ns; q++) s->pat_ptr[q + 1] = s->pat_ptr[q] + fronts[q];
    free(fronts);
    s->pat_rows = (int *)xmalloc((size_t)s->pat_ptr[ns] * sizeof(int));

    /* BuildPatterns */
    int rc = OR_OK;
    int *flag = (int *)xmalloc((size_t)n * sizeof(int));
    for (int j = 0; j < n; j++) flag[j] = -1;
    for (int q = 0; q < ns && rc == OR_OK; q++) {
        int first = s->first[q], last = s->first[q + 1] - 1;
        int wr = s->pat_ptr[q];
        for (int j = first; j <= last; j++) { flag[j] = q; s->pat_rows[wr++] = j; }
        for (int j = first; j <= last; j++)
            for (int p = s->lower_start[j]; p < s->pptr[j + 1]; p++) {
                int row = s->prows[p];
                if (flag[row] == q) continue;
                flag[row] = q;
                if (wr >= s->pat_ptr[q + 1]) { rc = OR_E_INTERNAL; break; }
                s->pat_rows[wr++] = row;
            }
        for (int c = s->child_ptr[q]; c < s->child_ptr[q + 1] && rc == OR_OK; c++) {
            int ch = s->children[c];
            int cw = s->first[ch + 1] - s->first[ch];
            for (int t = s->pat_ptr[ch] + cw; t < s->pat_ptr[ch + 1]; t++) {
                int row = s->pat_rows[t];
                if (flag[row] == q) continue;
                flag[row] = q;
                if (wr >= s->pat_ptr[q + 1]) { rc = OR_E_INTERNAL; break; }
                s->pat_rows[wr++] = row;
            }
        }
        if (rc == OR_OK && wr != s->pat_ptr[q + 1]) rc = OR_E_INTERNAL;
        if (rc == OR_OK)
            qsort(s->pat_rows + s->pat_ptr[q], (size_t)(s->pat_ptr[q + 1] - s->pat_ptr[q]), sizeof(int), int_cmp);
    }
    free(flag);
    if (rc) { or_symbolic_free(s); return rc; }

    /* ComputeCosts */
    int64_t entries = 0, ops = 0, live = 0, peak = 0;
    int maxf = 0, maxw = 0;
    int64_t *contrib = (int64_t *)xcalloc((size_t)ns, sizeof(int64_t));
    s->block_off = (int64_t *)xcalloc((size_t)ns + 1, sizeof(int64_t));
    for (int q = 0; q < ns; q++) {
        int64_t f = s->pat_ptr[q + 1] - s->pat_ptr[q];
        int64_t w = s->first[q + 1] - s->first[q];
        int64_t blk = f * w - w * (w - 1) / 2;
        if (blk > 2147483647LL) rc = OR_E_TOO_DENSE;
        entries += blk;
        s->block_off[q + 1] = s->block_off[q] + blk;
        for (int64_t k = 0; k < w; k++) ops += (f - k) * (f - k);
        if (f > maxf) maxf = (int)f;
        if (w > maxw) maxw = (int)w;
        if (live + f * (f + 1) / 2 > peak) peak = live + f * (f + 1) / 2;
        for (int c = s->child_ptr[q]; c < s->child_ptr[q + 1]; c++) live -= contrib[s->children[c]];
        int64_t t = f - w;
        contrib[q] = t * (t + 1) / 2;
        live += contrib[q];
        if (live > peak) peak = live;
    }
    free(contrib);
    s->factor_entries = entries;
    s->factor_ops = ops;
    s->max_front = maxf;
    s->max_width = maxw;
    s->peak_contribution = peak;
    int64_t strict = 0;
    for (int j = 0; j < n; j++) strict += s->colcount[j];
    s->strict_nonzeros = strict;
    if (rc) { or_symbolic_free(s); return rc; }
    *out = s;
    return OR_OK;
}

/* ======================================================================
 * Numeric phase -- UCHOL/UtduNumericFactorization.cs
 * ====================================================================== */

void or_numeric_free(or_numeric *f)
{
    if (!f) return;
    free(f->factor); free(f->diag); free(f);
}

/* SupernodalMultifrontalFactorizer.Factorize (:281-403) run serially in
 * ascending supernode order (the ready-stack order only changes WHEN a front is
 * processed; each front depends only on its children, so the numbers are the
 * same), FactorizeSupernode (:420-531), FactorizeFront (:541-601),
 * ApplyPanelUpdate / UpdateTrailingColumns / UpdateColumnTile / PanelDot
 * (:624-852): every trailing element gets F[i,j] -= sum_k (L[i,k]*d_k)*L[j,k],
 * the sum taken over the panel in ascending k from 0.0 and subtracted once. */
int or_factorize(const or_symbolic *s, const or_csr *a, const or_options *o, or_numeric **out)
{
    if (!s || !a || !o || !out) return OR_E_NULL;
    if (a->n != s->n || a->nnz != s->matrix_nnz) return OR_E_ARG; /* EnsureCompatible :306-319 */
    int n = s->n, ns = s->nsuper;
    or_numeric *F = (or_numeric *)xcalloc(1, sizeof(or_numeric));
    F->sym = s;
    F->diag = (double *)xcalloc((size_t)n, sizeof(double));
    F->factor = (double *)xmalloc((size_t)s->factor_entries * sizeof(double));
    if (ns == 0) { *out = F; return OR_OK; }
    int64_t maxlen = (int64_t)s->max_front * (s->max_front + 1) / 2;
    if (maxlen > 2147483647LL) { or_numeric_free(F); return OR_E_TOO_DENSE; }

    /* ComputeDiagonalScale :411-418 (Math.Max propagates NaN) */
    double scale = 0.0;
    for (int i = 0; i < n; i++) {
        double d = a->diag[i] >= 0 ? a->values[a->diag[i]] : 0.0;
        double m = fabs(d);
        if (isnan(m) || isnan(scale)) scale = NAN;
        else if (m > scale) scale = m;
    }
    if (!(scale > 0.0)) scale = 1.0;
    int pw = o->panel_width < 4 ? 4 : (o->panel_width > 128 ? 128 : o->panel_width);
    double thr = o->pivot_tolerance * scale;
    double regv = o->regularization * scale;

    double **updates = (double **)xcalloc((size_t)ns, sizeof(double *));
    int *offs = (int *)xmalloc(((size_t)s->max_front + 1) * sizeof(int));
    int *rowmap = (int *)xmalloc((size_t)n * sizeof(int));
    int *rel = (int *)xmalloc((size_t)s->max_front * sizeof(int) + sizeof(int));
    double *front = (double *)xmalloc((size_t)maxlen * sizeof(double));
    int regularized = 0, negative = 0;
    double smallest = DBL_MAX, largest = 0.0;
    int rc = OR_OK;
    const double t0 = now_s();
    int done = 0;
    int64_t ops_done = 0;

    for (int q = 0; q < ns && rc == OR_OK; q++) {
        if (o->time_budget_s > 0.0 && now_s() - t0 >= o->time_budget_s) break;
        int first = s->first[q];
        int w = s->first[q + 1] - first;
        int f = s->pat_ptr[q + 1] - s->pat_ptr[q];
        const int *prow = s->pat_rows + s->pat_ptr[q];
        int flen = (int)((int64_t)f * (f + 1) / 2);
        offs[0] = 0;
        for (int j = 0; j < f; j++) offs[j + 1] = offs[j] + (f - j);
        memset(front, 0, (size_t)flen * sizeof(double));
        for (int t = 0; t < f; t++) rowmap[prow[t]] = t;
        /* assembly of original entries :450-468 */
        for (int loc = 0; loc < w; loc++) {
            int col = first + loc;
            int di = s->dvidx[col];
            if (di >= 0) front[offs[loc]] += a->values[di];
            int lb = s->lvptr[col], st = s->lower_start[col], en = s->pptr[col + 1];
            int cb = offs[loc] - loc;
            for (int p = st; p < en; p++) {
                int row = rowmap[s->prows[p]];
                front[cb + row] += a->values[s->lvidx[lb + (p - st)]];
            }
        }
        /* extend-add :471-509 */
        for (int c = s->child_ptr[q]; c < s->child_ptr[q + 1]; c++) {
            int ch = s->children[c];
            double *upd = updates[ch];
            if (!upd) continue;
            int cw = s->first[ch + 1] - s->first[ch];
            int ct = (s->pat_ptr[ch + 1] - s->pat_ptr[ch]) - cw;
            const int *crow = s->pat_rows + s->pat_ptr[ch] + cw;
            for (int x = 0; x < ct; x++) rel[x] = rowmap[crow[x]];
            int src = 0;
            for (int x = 0; x < ct; x++) {
                int tc = rel[x];
                int tb = offs[tc] - tc;
                for (int y = x; y < ct; y++) front[tb + rel[y]] += upd[src + (y - x)];
                src += ct - x;
            }
            free(upd);
            updates[ch] = NULL;
        }
        /* FactorizeFront :541-601 */
        for (int ps = 0; ps < w && rc == OR_OK; ps += pw) {
            int pe = ps + pw < w ? ps + pw : w;
            for (int k = ps; k < pe; k++) {
                int co = offs[k], cl = f - k;
                double pivot = front[co];
                if (!(pivot > thr)) {
                    if (!o->allow_regularization) { rc = OR_E_NOT_PD; break; }
                    if (pivot < 0.0) negative++;
                    regularized++;
                    pivot = regv;
                }
                F->diag[first + k] = pivot;
                double mag = fabs(pivot);
                if (mag < smallest) smallest = mag;
                if (mag > largest) largest = mag;
                double inv = 1.0 / pivot;
                for (int t = 1; t < cl; t++) front[co + t] *= inv;
                for (int j = k + 1; j < pe; j++) {
                    double factor = pivot * front[co + (j - k)];
                    if (factor == 0.0) continue;
                    double *tgt = front + offs[j];
                    const double *src = front + co + (j - k);
                    for (int t = 0; t < f - j; t++) tgt[t] -= src[t] * factor;
                }
                front[co] = 1.0;
            }
            if (rc != OR_OK || pe >= f) continue;
            int np = pe - ps;
            /* panel x D, row-major per trailing row: buf[(i-pe)*np + kk] */
            int trail = f - pe;
            double *buf = (double *)xmalloc((size_t)trail * np * sizeof(double));
            for (int kk = 0; kk < np; kk++) {
                int col = ps + kk;
                double d = F->diag[first + col];
                const double *L = front + offs[col] - col;
                for (int i = pe; i < f; i++) buf[(size_t)(i - pe) * np + kk] = L[i] * d;
            }
            for (int j = pe; j < f; j++) {
                double *Fj = front + offs[j] - j;
                for (int i = j; i < f; i++) {
                    const double *bi = buf + (size_t)(i - pe) * np;
                    double sum = 0.0;
                    for (int kk = 0; kk < np; kk++) {
                        int col = ps + kk;
                        sum += bi[kk] * front[offs[col] - col + j];
                    }
                    Fj[i] -= sum;
                }
            }
            free(buf);
        }
        if (rc != OR_OK) break;
        int blen = offs[w];
        memcpy(F->factor + s->block_off[q], front, (size_t)blen * sizeof(double));
        if (f - w > 0) {
            int ulen = flen - blen;
            updates[q] = (double *)xmalloc((size_t)ulen * sizeof(double));
            memcpy(updates[q], front + blen, (size_t)ulen * sizeof(double));
        }
        for (int64_t k = 0; k < w; k++) ops_done += ((int64_t)f - k) * ((int64_t)f - k);
        done++;
    }
    for (int q = 0; q < ns; q++) free(updates[q]);
    free(updates); free(offs); free(rowmap); free(rel); free(front);
    if (rc) { or_numeric_free(F); return rc; }
    if (smallest == DBL_MAX) smallest = 0.0;
    F->regularized = regularized;
    F->negative = negative;
    F->smallest = smallest;
    F->largest = largest;
    F->supernodes_done = done;
    F->ops_done = ops_done;
    *out = F;
    return OR_OK;
}

/* DenseKernels.Dot (DenseKernels.cs:91-113) for Vector<double>.Count == 4; the
 * horizontal sum of the accumulator is taken as (a0+a1)+(a2+a3). */
static double or_dot4(const double *l, const double *r, int len)
{
    int t = 0;
    double sum = 0.0;
    if (len >= 4) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (; t <= len - 4; t += 4) {
            a0 += l[t] * r[t];
            a1 += l[t + 1] * r[t + 1];
            a2 += l[t + 2] * r[t + 2];
            a3 += l[t + 3] * r[t + 3];
        }
        sum = (a0 + a1) + (a2 + a3);
    }
    for (; t < len; t++) sum += l[t] * r[t];
    return sum;
}

/* UtduNumericFactorization.Solve (:90-112), SolvePermuted (:118-129),
 * ForwardSubstitution (:139-173), BackwardSubstitution (:179-207). */
int or_solve(const or_numeric *F, const double *b, double *x)
{
    if (!F || !b || !x) return OR_E_NULL;
    const or_symbolic *s = F->sym;
    int n = s->n;
    double *wk = (double *)xmalloc((size_t)n * sizeof(double));
    double *sc = (double *)xmalloc(((size_t)s->max_front + 1) * sizeof(double));
    for (int i = 0; i < n; i++) wk[i] = b[s->perm[i]];
    for (int q = 0; q < s->nsuper; q++) {
        int f = s->pat_ptr[q + 1] - s->pat_ptr[q], w = s->first[q + 1] - s->first[q];
        const int *pr = s->pat_rows + s->pat_ptr[q];
        const double *blk = F->factor + s->block_off[q];
        for (int t = 0; t < f; t++) sc[t] = wk[pr[t]];
        int off = 0;
        for (int k = 0; k < w; k++) {
            int cl = f - k;
            double v = sc[k];
            if (v != 0.0)
                for (int t = 0; t < cl - 1; t++) sc[k + 1 + t] -= blk[off + 1 + t] * v;
            off += cl;
        }
        for (int t = 0; t < f; t++) wk[pr[t]] = sc[t];
    }
    for (int i = 0; i < n; i++) wk[i] /= F->diag[i];
    for (int q = s->nsuper - 1; q >= 0; q--) {
        int f = s->pat_ptr[q + 1] - s->pat_ptr[q], w = s->first[q + 1] - s->first[q];
        const int *pr = s->pat_rows + s->pat_ptr[q];
        const double *blk = F->factor + s->block_off[q];
        for (int t = 0; t < f; t++) sc[t] = wk[pr[t]];
        for (int k = w - 1; k >= 0; k--) {
            int off = k * f - k * (k - 1) / 2;
            int cl = f - k;
            sc[k] -= or_dot4(blk + off + 1, sc + k + 1, cl - 1);
        }
        for (int t = 0; t < w; t++) wk[pr[t]] = sc[t];
    }
    for (int i = 0; i < n; i++) x[s->perm[i]] = wk[i];
    free(wk); free(sc);
    return OR_OK;
}

/* ======================================================================
 * CGGP -- CGGP/ConjugateGradientGaussPreSolver.cs:78-175,
 * CGGP/JacobiPreconditioner.cs:12-56, IterativeLinearSolver.cs:26-33
 * ====================================================================== */
int or_pcg(const or_csr *a, const double *b, const double *x0, double tol,
           int max_iterations, int use_jacobi, double *x, or_pcg_result *res)
{
    if (!a || !b || !x || !res) return OR_E_NULL;
    if (!isfinite(tol) || tol < 0.0) return OR_E_RANGE;
    if (max_iterations < 0) return OR_E_RANGE;
    int n = a->n;
    for (int i = 0; i < n; i++) x[i] = x0 ? x0[i] : 0.0;
    double bn = or_norm2(n, b);
    double scale = bn > 0.0 ? bn : 1.0;
    res->iterations = 0;
    res->converged = 1;
    res->residual_norm = 0.0;
    res->relative_residual = 0.0;
    if (n == 0) { res->converged = 0.0 / scale <= tol; return OR_OK; }
    double thr = tol * scale;
    double *r = (double *)xmalloc((size_t)n * sizeof(double));
    double *z = (double *)xmalloc((size_t)n * sizeof(double));
    double *p = (double *)xmalloc((size_t)n * sizeof(double));
    double *ap = (double *)xmalloc((size_t)n * sizeof(double));
    double *invd = NULL;
    int rc = OR_OK, iters = 0;
    or_spmv(a, x, ap);
    for (int i = 0; i < n; i++) r[i] = b[i] - ap[i];
    double rn = or_norm2(n, r);
    if (rn <= thr) goto done;
    if (use_jacobi) {
        invd = (double *)xmalloc((size_t)n * sizeof(double));
        for (int i = 0; i < n; i++) {
            double d = a->diag[i] >= 0 ? a->values[a->diag[i]] : 0.0;
            if (!isfinite(d) || fabs(d) <= 1e-300) { rc = OR_E_NOT_PD; goto done; }
            invd[i] = 1.0 / d;
        }
    }
    for (int i = 0; i < n; i++) z[i] = invd ? invd[i] * r[i] : r[i];
    memcpy(p, z, (size_t)n * sizeof(double));
    double rz = 0.0;
    for (int i = 0; i < n; i++) rz += r[i] * z[i];
    int maxit = max_iterations > 0 ? max_iterations : n;
    iters = maxit;
    for (int it = 1; it <= maxit; it++) {
        or_spmv(a, p, ap);
        double pap = 0.0;
        for (int i = 0; i < n; i++) pap += p[i] * ap[i];
        if (pap <= 0.0) { rc = OR_E_NOT_PD; goto done; }
        double alpha = rz / pap;
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
        rn = or_norm2(n, r);
        if (rn <= thr) { iters = it; goto done; }
        for (int i = 0; i < n; i++) z[i] = invd ? invd[i] * r[i] : r[i];
        double rz2 = 0.0;
        for (int i = 0; i < n; i++) rz2 += r[i] * z[i];
        double beta = rz2 / rz;
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rz2;
    }
done:
    if (rc == OR_OK) {
        res->iterations = iters;
        res->residual_norm = rn;
        res->relative_residual = rn / scale;
        res->converged = res->relative_residual <= tol;
    }
    free(r); free(z); free(p); free(ap); free(invd);
    return rc;
}

/* ---- learned preconditioner network (no reference counterpart; see oracle.h) ---- */

typedef struct {
    int n;
    int *ptr, *nbr;          /* full neighbour lists without the diagonal, ascending */
    double *ew;              /* e1_ij */
    double *diag, *x;        /* a_ii; node features [1, c_i, s_i] */
    int *gptr, *gcol;        /* pattern of G: row i = {j <= i within two hops}, ascending (diagonal last) */
    double *ge1, *ge2;       /* e1_ij, e2_ij on the pattern of G */
} mp_graph;

static void mp_graph_free(mp_graph *g) {
    free(g->ptr); free(g->nbr); free(g->ew); free(g->diag); free(g->x);
    free(g->gptr); free(g->gcol); free(g->ge1); free(g->ge2);
}

static int cmp_int(const void *a, const void *b) { return (*(const int *)a > *(const int *)b) - (*(const int *)a < *(const int *)b); }

static int mp_graph_build(const or_csr *a, mp_graph *g) {
    const int n = a->n;
    memset(g, 0, sizeof(*g));
    g->n = n;
    g->ptr = (int *)calloc((size_t)n + 1, sizeof(int));
    g->diag = (double *)calloc((size_t)n + 1, sizeof(double));
    g->x = (double *)calloc(3 * (size_t)n + 1, sizeof(double));
    g->gptr = (int *)calloc((size_t)n + 1, sizeof(int));
    if (!g->ptr || !g->diag || !g->x || !g->gptr) return OR_E_NOMEM;
    for (int i = 0; i < n; ++i)
        for (int k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) {
            const int j = a->colidx[k];
            if (j == i) g->diag[i] = a->values[k];
            else { g->ptr[i + 1]++; g->ptr[j + 1]++; }
        }
    for (int i = 0; i < n; ++i) g->ptr[i + 1] += g->ptr[i];
    const int m = g->ptr[n];
    g->nbr = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    g->ew = (double *)malloc(sizeof(double) * (size_t)(m + 1));
    int *fill = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    if (!g->nbr || !g->ew || !fill) { free(fill); return OR_E_NOMEM; }
    memcpy(fill, g->ptr, sizeof(int) * (size_t)n);
    /* ascending rows: lower neighbours (upper entries of earlier rows) first, then upper entries */
    for (int i = 0; i < n; ++i)
        for (int k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) {
            const int j = a->colidx[k];
            if (j == i) continue;
            g->nbr[fill[j]] = i; g->ew[fill[j]++] = a->values[k] / sqrt(g->diag[i] * g->diag[j]);
        }
    for (int i = 0; i < n; ++i)
        for (int k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) {
            const int j = a->colidx[k];
            if (j == i) continue;
            g->nbr[fill[i]] = j; g->ew[fill[i]++] = a->values[k] / sqrt(g->diag[i] * g->diag[j]);
        }
    free(fill);
    double mean = 0.0;
    for (int i = 0; i < n; ++i) mean += g->diag[i];
    mean = n ? mean / n : 1.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = g->ptr[i]; k < g->ptr[i + 1]; ++k) s += fabs(g->ew[k]);
        g->x[3 * i] = 1.0;
        g->x[3 * i + 1] = log(g->diag[i] / mean);
        g->x[3 * i + 2] = s;
    }
    /* pattern of G and its weights: dense scratch rows w1/w2 indexed by column */
    double *w1 = (double *)calloc((size_t)n + 1, sizeof(double)), *w2 = (double *)calloc((size_t)n + 1, sizeof(double));
    char *seen = (char *)calloc((size_t)n + 1, 1);
    int *cols = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int cap = 16 * n + 16, used = 0;
    g->gcol = (int *)malloc(sizeof(int) * (size_t)cap);
    g->ge1 = (double *)malloc(sizeof(double) * (size_t)cap);
    g->ge2 = (double *)malloc(sizeof(double) * (size_t)cap);
    if (!w1 || !w2 || !seen || !cols || !g->gcol || !g->ge1 || !g->ge2) {
        free(w1); free(w2); free(seen); free(cols);
        return OR_E_NOMEM;
    }
    for (int i = 0; i < n; ++i) {
        int nc = 0;
        seen[i] = 1; cols[nc++] = i;
        for (int k = g->ptr[i]; k < g->ptr[i + 1]; ++k) {
            const int j = g->nbr[k];
            if (j < i) { w1[j] += g->ew[k]; if (!seen[j]) { seen[j] = 1; cols[nc++] = j; } }
            for (int q = g->ptr[j]; q < g->ptr[j + 1]; ++q) {
                const int l = g->nbr[q];
                if (l >= i) continue;
                w2[l] += g->ew[k] * g->ew[q];
                if (!seen[l]) { seen[l] = 1; cols[nc++] = l; }
            }
        }
        qsort(cols, (size_t)nc, sizeof(int), cmp_int);
        if (used + nc > cap) {
            cap = 2 * (used + nc);
            g->gcol = (int *)realloc(g->gcol, sizeof(int) * (size_t)cap);
            g->ge1 = (double *)realloc(g->ge1, sizeof(double) * (size_t)cap);
            g->ge2 = (double *)realloc(g->ge2, sizeof(double) * (size_t)cap);
            if (!g->gcol || !g->ge1 || !g->ge2) { free(w1); free(w2); free(seen); free(cols); return OR_E_NOMEM; }
        }
        for (int t = 0; t < nc; ++t) {
            const int j = cols[t];
            g->gcol[used] = j;
            g->ge1[used] = w1[j];
            g->ge2[used] = w2[j];
            ++used;
            w1[j] = w2[j] = 0.0;
            seen[j] = 0;
        }
        g->gptr[i + 1] = used;
    }
    free(w1); free(w2); free(seen); free(cols);
    return OR_OK;
}

/* G values on its pattern */
typedef struct {
    double *gv;
} mp_factor;

#define MP_OFF_T (3 * OR_MP_W + 6 * OR_MP_W * OR_MP_W)

/* forward pass of the network: keeps h_l (4 x n x W), m_l and pre_l (3 x n x W); fills G */
static void mp_forward(const mp_graph *g, const double *P, double *H, double *Mb, double *Pre, mp_factor *G) {
    const int n = g->n, W = OR_MP_W;
    const double *Win = P, *wt = P + MP_OFF_T;
    for (int i = 0; i < n; ++i)
        for (int f = 0; f < W; ++f)
            H[(size_t)i * W + f] = Win[3 * f] * g->x[3 * i] + Win[3 * f + 1] * g->x[3 * i + 1] +
                                   Win[3 * f + 2] * g->x[3 * i + 2];
    for (int l = 0; l < 3; ++l) {
        const double *W1 = P + 3 * W + 2 * l * W * W, *W2 = W1 + W * W;
        const double *h = H + (size_t)l * n * W;
        double *ho = H + (size_t)(l + 1) * n * W, *m = Mb + (size_t)l * n * W, *pre = Pre + (size_t)l * n * W;
        for (int i = 0; i < n; ++i) {
            for (int f = 0; f < W; ++f) {
                double acc = 0.0;
                for (int k = g->ptr[i]; k < g->ptr[i + 1]; ++k) acc += g->ew[k] * h[(size_t)g->nbr[k] * W + f];
                m[(size_t)i * W + f] = acc;
            }
            for (int f = 0; f < W; ++f) {
                double acc = 0.0;
                for (int q = 0; q < W; ++q) acc += W1[f * W + q] * h[(size_t)i * W + q] + W2[f * W + q] * m[(size_t)i * W + q];
                pre[(size_t)i * W + f] = acc;
                ho[(size_t)i * W + f] = (l < 2 && !(acc > 0.0)) ? 0.0 : acc;
            }
        }
    }
    const double *h3 = H + (size_t)3 * n * W;
    for (int i = 0; i < n; ++i) {
        const double *hi = h3 + (size_t)i * W;
        for (int e = g->gptr[i]; e < g->gptr[i + 1]; ++e) {
            const int j = g->gcol[e];
            if (j == i) {
                double t = 0.0;
                for (int f = 0; f < W; ++f) t += wt[f] * hi[f];
                G->gv[e] = exp(t) / sqrt(g->diag[i]);
                continue;
            }
            const double *hj = h3 + (size_t)j * W;
            double p[2] = {0.0, 0.0};
            for (int s = 0; s < 2; ++s) {
                const double *u = P + MP_OFF_T + (1 + 3 * s) * W, *v = u + W, *w = v + W;
                for (int f = 0; f < W; ++f) p[s] += u[f] * hi[f] * hj[f] + v[f] * hi[f] + w[f] * hj[f];
            }
            G->gv[e] = (g->ge1[e] * p[0] + g->ge2[e] * p[1]) / sqrt(g->diag[j]);
        }
    }
}

/* y = G r ; z = G' y */
static void mp_apply_g(const mp_graph *g, const mp_factor *G, const double *r, double *y, double *z) {
    const int n = g->n;
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int e = g->gptr[i]; e < g->gptr[i + 1]; ++e) acc += G->gv[e] * r[g->gcol[e]];
        y[i] = acc;
    }
    for (int i = 0; i < n; ++i) z[i] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int e = g->gptr[i]; e < g->gptr[i + 1]; ++e) z[g->gcol[e]] += G->gv[e] * y[i];
}

/* scale != 0: ga += row i of G A (A_kk = a_kk, A_kj = e1_kj sqrt(a_kk a_jj)); scale == 0: clear
   the entries that row touches */
static void mp_row_ga(const mp_graph *g, const mp_factor *G, int i, double *ga, double scale) {
    for (int e = g->gptr[i]; e < g->gptr[i + 1]; ++e) {
        const int k = g->gcol[e];
        const double gik = G->gv[e];
        if (scale == 0.0) {
            ga[k] = 0.0;
            for (int q = g->ptr[k]; q < g->ptr[k + 1]; ++q) ga[g->nbr[q]] = 0.0;
        } else {
            ga[k] += gik * g->diag[k];
            for (int q = g->ptr[k]; q < g->ptr[k + 1]; ++q)
                ga[g->nbr[q]] += gik * g->ew[q] * sqrt(g->diag[k] * g->diag[g->nbr[q]]);
        }
    }
}

static void *mp_alloc(size_t count) { return malloc(sizeof(double) * (count + 1)); }

int or_mpnn_apply(const or_csr *a, const double *params, const double *r, double *z) {
    if (!a || !params || !r || !z) return OR_E_NULL;
    mp_graph g;
    if (mp_graph_build(a, &g)) { mp_graph_free(&g); return OR_E_NOMEM; }
    const size_t nw = (size_t)a->n * OR_MP_W;
    double *H = (double *)mp_alloc(4 * nw), *Mb = (double *)mp_alloc(3 * nw), *Pre = (double *)mp_alloc(3 * nw);
    double *y = (double *)mp_alloc((size_t)a->n);
    mp_factor G = {(double *)mp_alloc((size_t)g.gptr[a->n])};
    int rc = OR_E_NOMEM;
    if (H && Mb && Pre && y && G.gv) {
        mp_forward(&g, params, H, Mb, Pre, &G);
        mp_apply_g(&g, &G, r, y, z);
        rc = OR_OK;
    }
    free(H); free(Mb); free(Pre); free(y); free(G.gv);
    mp_graph_free(&g);
    return rc;
}

int or_mpnn_loss_grad(const or_csr *a, const double *params, int nsamples, const double *R, const double *E,
                      double kappa, double *loss, double *grad) {
    if (!a || !params || (nsamples > 0 && (!R || !E)) || !loss || !grad) return OR_E_NULL;
    const int n = a->n, W = OR_MP_W;
    mp_graph g;
    if (mp_graph_build(a, &g)) { mp_graph_free(&g); return OR_E_NOMEM; }
    const size_t nw = (size_t)n * W;
    const int ng = g.gptr[n];
    double *H = (double *)mp_alloc(4 * nw), *Mb = (double *)mp_alloc(3 * nw), *Pre = (double *)mp_alloc(3 * nw);
    double *dH = (double *)calloc(nw + 1, sizeof(double)), *dHp = (double *)mp_alloc(nw), *dM = (double *)mp_alloc(nw);
    double *v = (double *)mp_alloc(6 * (size_t)n), *dG = (double *)calloc((size_t)ng + 1, sizeof(double));
    mp_factor G = {(double *)mp_alloc((size_t)ng)};
    if (!H || !Mb || !Pre || !dH || !dHp || !dM || !v || !dG || !G.gv) {
        free(H); free(Mb); free(Pre); free(dH); free(dHp); free(dM); free(v); free(dG); free(G.gv);
        mp_graph_free(&g);
        return OR_E_NOMEM;
    }
    memset(grad, 0, sizeof(double) * OR_MP_P);
    mp_forward(&g, params, H, Mb, Pre, &G);
    double *y = v, *z = v + n, *az = v + 2 * (size_t)n, *gg = v + 3 * (size_t)n, *u = v + 4 * (size_t)n,
           *tmp = v + 5 * (size_t)n;
    double total = 0.0, kaporin_value = 0.0;
    for (int t = 0; t < nsamples; ++t) {
        const double *r = R + (size_t)t * n, *e = E + (size_t)t * n;
        mp_apply_g(&g, &G, r, y, z);
        or_spmv(a, z, az);
        double zaz = 0.0, zr = 0.0, re = 0.0;
        for (int i = 0; i < n; ++i) { zaz += z[i] * az[i]; zr += z[i] * r[i]; re += r[i] * e[i]; }
        total += (zaz - 2.0 * zr) / re + 1.0;
        for (int i = 0; i < n; ++i) gg[i] = 2.0 * (az[i] - r[i]) / re / nsamples;
        mp_apply_g(&g, &G, gg, u, tmp); /* u = G g */
        /* dL/dG_ij = y_i g_j + u_i r_j on the pattern of G */
        for (int i = 0; i < n; ++i)
            for (int q = g.gptr[i]; q < g.gptr[i + 1]; ++q) dG[q] += y[i] * gg[g.gcol[q]] + u[i] * r[g.gcol[q]];
    }
    /* Kaporin term: T = tr(G A G') = sum_i g_i' A g_i (g_i = row i of G);
       dL_K/dG_ij = (2/T) (G A)_ij - (2/n) [j == i] / G_ii */
    if (kappa != 0.0 && n > 0) {
        double *ga = tmp; /* dense row of G A; zero between rows */
        for (int i = 0; i < n; ++i) ga[i] = 0.0;
        double T = 0.0, logdiag = 0.0;
        for (int sweep = 0; sweep < 2; ++sweep)
            for (int i = 0; i < n; ++i) {
                mp_row_ga(&g, &G, i, ga, 1.0);
                for (int q = g.gptr[i]; q < g.gptr[i + 1]; ++q) {
                    const int j = g.gcol[q];
                    if (sweep == 0) {
                        T += G.gv[q] * ga[j];
                        if (j == i) logdiag += log(G.gv[q]);
                    } else {
                        dG[q] += kappa * (2.0 / T * ga[j] - (j == i ? 2.0 / n / G.gv[q] : 0.0));
                    }
                }
                mp_row_ga(&g, &G, i, ga, 0.0);
            }
        kaporin_value = log(T / n) - 2.0 / n * logdiag;
    }
    /* readouts */
    const double *h3 = H + (size_t)3 * n * W;
    const double *wt = params + MP_OFF_T;
    double *gwt = grad + MP_OFF_T;
    for (int i = 0; i < n; ++i) {
        const double *hi = h3 + (size_t)i * W;
        for (int q = g.gptr[i]; q < g.gptr[i + 1]; ++q) {
            const int j = g.gcol[q];
            if (j == i) {
                const double dt = dG[q] * G.gv[q];
                for (int f = 0; f < W; ++f) { gwt[f] += dt * hi[f]; dH[(size_t)i * W + f] += dt * wt[f]; }
                continue;
            }
            const double *hj = h3 + (size_t)j * W;
            for (int s = 0; s < 2; ++s) {
                const double dp = dG[q] * (s == 0 ? g.ge1[q] : g.ge2[q]) / sqrt(g.diag[j]);
                const double *uu = params + MP_OFF_T + (1 + 3 * s) * W, *vv = uu + W, *ww = vv + W;
                double *gu = grad + MP_OFF_T + (1 + 3 * s) * W, *gv = gu + W, *gw = gv + W;
                for (int f = 0; f < W; ++f) {
                    gu[f] += dp * hi[f] * hj[f];
                    gv[f] += dp * hi[f];
                    gw[f] += dp * hj[f];
                    dH[(size_t)i * W + f] += dp * (uu[f] * hj[f] + vv[f]);
                    dH[(size_t)j * W + f] += dp * (uu[f] * hi[f] + ww[f]);
                }
            }
        }
    }
    for (int l = 2; l >= 0; --l) {
        const double *W1 = params + 3 * W + 2 * l * W * W, *W2 = W1 + W * W;
        double *gW1 = grad + 3 * W + 2 * l * W * W, *gW2 = gW1 + W * W;
        const double *h = H + (size_t)l * n * W, *mm = Mb + (size_t)l * n * W, *pre = Pre + (size_t)l * n * W;
        for (int i = 0; i < n; ++i) {
            double dp[OR_MP_W];
            for (int f = 0; f < W; ++f) {
                const double p = pre[(size_t)i * W + f];
                dp[f] = (l < 2 && !(p > 0.0)) ? 0.0 : dH[(size_t)i * W + f];
            }
            for (int f = 0; f < W; ++f)
                for (int q = 0; q < W; ++q) {
                    gW1[f * W + q] += dp[f] * h[(size_t)i * W + q];
                    gW2[f * W + q] += dp[f] * mm[(size_t)i * W + q];
                }
            for (int q = 0; q < W; ++q) {
                double a1 = 0.0, a2 = 0.0;
                for (int f = 0; f < W; ++f) { a1 += W1[f * W + q] * dp[f]; a2 += W2[f * W + q] * dp[f]; }
                dHp[(size_t)i * W + q] = a1;
                dM[(size_t)i * W + q] = a2;
            }
        }
        /* transpose of the aggregation (e1 symmetric): dh_i += sum_j e1_ij dm_j */
        for (int i = 0; i < n; ++i)
            for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
                for (int q = 0; q < W; ++q) dHp[(size_t)i * W + q] += g.ew[k] * dM[(size_t)g.nbr[k] * W + q];
        memcpy(dH, dHp, sizeof(double) * nw);
    }
    for (int i = 0; i < n; ++i)
        for (int f = 0; f < W; ++f)
            for (int c = 0; c < 3; ++c) grad[3 * f + c] += dH[(size_t)i * W + f] * g.x[3 * i + c];
    *loss = (nsamples ? total / nsamples : 0.0) + kappa * kaporin_value;
    free(H); free(Mb); free(Pre); free(dH); free(dHp); free(dM); free(v); free(dG); free(G.gv);
    mp_graph_free(&g);
    return OR_OK;
}


/* ======================================================================
 * CSRMatrix, Ilu0Preconditioner, BiCgStabSolver
 * ====================================================================== */

/* CSRMatrix.cs:34-80 -- group by (row, col) and sum in input order (:39-41), drop
 * |sum| <= 1e-15 (:42), columns ascending per row (:44), diagonal table (:65-66).
 * Indices are checked as CSRMatrixBuilder.AddToElement does (CSRMatrixBuilder.cs:74-81). */
int or_gcsr_from_coo(int rows, int cols, int64_t count, const int *ri, const int *cj, const double *vals,
                     or_gcsr **out)
{
    if (!out || (count > 0 && (!ri || !cj || !vals))) return OR_E_NULL;
    if (rows < 0 || cols < 0 || count < 0) return OR_E_ARG;
    for (int64_t k = 0; k < count; k++)
        if (ri[k] < 0 || ri[k] >= rows || cj[k] < 0 || cj[k] >= cols) return OR_E_RANGE;
    coo_item *items = (coo_item *)xmalloc((size_t)count * sizeof(coo_item));
    for (int64_t k = 0; k < count; k++) {
        items[k].row = ri[k]; items[k].col = cj[k]; items[k].order = k; items[k].value = vals[k];
    }
    qsort(items, (size_t)count, sizeof(coo_item), coo_cmp);
    int64_t kept = 0;
    for (int64_t k = 0; k < count;) {
        int64_t e = k;
        double sum = 0.0;
        while (e < count && items[e].row == items[k].row && items[e].col == items[k].col) {
            sum += items[e].value;
            e++;
        }
        if (fabs(sum) > 1e-15) {
            items[kept] = items[k];
            items[kept].value = sum;
            kept++;
        }
        k = e;
    }
    or_gcsr *a = (or_gcsr *)xcalloc(1, sizeof(or_gcsr));
    a->rows = rows; a->cols = cols; a->nnz = (int)kept;
    a->rowptr = (int *)xcalloc((size_t)rows + 1, sizeof(int));
    a->colidx = (int *)xmalloc((size_t)kept * sizeof(int));
    a->values = (double *)xmalloc((size_t)kept * sizeof(double));
    a->diag = (int *)xmalloc((size_t)rows * sizeof(int));
    for (int i = 0; i < rows; i++) a->diag[i] = EMPTY;
    for (int64_t k = 0; k < kept; k++) {
        a->rowptr[items[k].row + 1]++;
        a->colidx[k] = items[k].col;
        a->values[k] = items[k].value;
        if (items[k].row == items[k].col) a->diag[items[k].row] = (int)k;
    }
    for (int i = 0; i < rows; i++) a->rowptr[i + 1] += a->rowptr[i];
    free(items);
    *out = a;
    return OR_OK;
}

int or_gcsr_create(int rows, int cols, const int *rowptr, const int *colidx, const double *values, or_gcsr **out)
{
    if (!out || !rowptr) return OR_E_NULL;
    int nnz = rowptr[rows];
    if (nnz > 0 && (!colidx || !values)) return OR_E_NULL;
    or_gcsr *a = (or_gcsr *)xcalloc(1, sizeof(or_gcsr));
    a->rows = rows; a->cols = cols; a->nnz = nnz;
    a->rowptr = (int *)xmalloc(((size_t)rows + 1) * sizeof(int));
    a->colidx = (int *)xmalloc((size_t)nnz * sizeof(int));
    a->values = (double *)xmalloc((size_t)nnz * sizeof(double));
    a->diag = (int *)xmalloc((size_t)rows * sizeof(int));
    memcpy(a->rowptr, rowptr, ((size_t)rows + 1) * sizeof(int));
    if (nnz) {
        memcpy(a->colidx, colidx, (size_t)nnz * sizeof(int));
        memcpy(a->values, values, (size_t)nnz * sizeof(double));
    }
    for (int i = 0; i < rows; i++) {
        a->diag[i] = EMPTY;
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
            if (colidx[k] == i) a->diag[i] = k;
    }
    *out = a;
    return OR_OK;
}

void or_gcsr_free(or_gcsr *a)
{
    if (!a) return;
    free(a->rowptr); free(a->colidx); free(a->values); free(a->diag); free(a);
}

/* CSRMatrix.Multiply (CSRMatrix.cs:230-248), also BiCgStabSolver.Multiply (:167-180) */
void or_gspmv(const or_gcsr *a, const double *x, double *y)
{
    for (int i = 0; i < a->rows; i++) {
        double sum = 0.0;
        for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; j++) sum += a->values[j] * x[a->colidx[j]];
        y[i] = sum;
    }
}

/* CSRMatrix.LineCross (CSRMatrix.cs:110-164) */
int or_gcsr_linecross(or_gcsr *a, double *rhs, double prescribed, int index)
{
    if (!a || !rhs) return OR_E_NULL;
    if (a->rows != a->cols) return OR_E_NOT_PD;
    if (index < 0 || index >= a->rows) return OR_E_RANGE;
    int di = a->diag[index];
    if (di < 0) return OR_E_NOT_PD;
    for (int row = 0; row < a->rows; row++)
        for (int p = a->rowptr[row]; p < a->rowptr[row + 1]; p++) {
            int col = a->colidx[p];
            if (row == index) {
                if (col != index) a->values[p] = 0.0;
            } else if (col == index) {
                rhs[row] -= a->values[p] * prescribed;
                a->values[p] = 0.0;
            }
        }
    a->values[di] = 1.0;
    rhs[index] = prescribed;
    return OR_OK;
}

/* Ilu0Preconditioner.FindPosition (:139-160): binary search of (row, col) */
static int ilu_find(const or_gcsr *a, int row, int col)
{
    int left = a->rowptr[row], right = a->rowptr[row + 1] - 1;
    while (left <= right) {
        int middle = left + (right - left) / 2;
        int c = a->colidx[middle];
        if (c == col) return middle;
        if (c < col) left = middle + 1;
        else right = middle - 1;
    }
    return -1;
}

static int ilu_bad_pivot(double pivot, double tol) { return !isfinite(pivot) || fabs(pivot) < tol; }

/* Ilu0Preconditioner(CSRMatrix, double) (:14-31): FindDiagonalIndices (:86-109) takes the
 * first position of column i in row i; Factorize (:111-137) eliminates row by row. */
int or_ilu0_create(const or_gcsr *a, double pivot_tolerance, or_ilu0 **out, int *bad_row)
{
    if (!a || !out) return OR_E_NULL;
    if (bad_row) *bad_row = -1;
    if (a->rows != a->cols) return OR_E_ARG;
    if (!isfinite(pivot_tolerance) || pivot_tolerance <= 0.0) return OR_E_RANGE;
    int n = a->rows;
    int *dg = (int *)xmalloc((size_t)n * sizeof(int));
    for (int row = 0; row < n; row++) {
        dg[row] = -1;
        for (int p = a->rowptr[row]; p < a->rowptr[row + 1]; p++)
            if (a->colidx[p] == row) { dg[row] = p; break; }
        if (dg[row] < 0) {
            if (bad_row) *bad_row = row;
            free(dg);
            return OR_E_NOT_PD;
        }
    }
    double *f = (double *)xmalloc((size_t)a->nnz * sizeof(double));
    if (a->nnz) memcpy(f, a->values, (size_t)a->nnz * sizeof(double));
    for (int row = 0; row < n; row++) {
        for (int p = a->rowptr[row]; p < dg[row]; p++) {
            int column = a->colidx[p];
            double pivot = f[dg[column]];
            if (ilu_bad_pivot(pivot, pivot_tolerance)) {
                if (bad_row) *bad_row = column;
                free(dg); free(f);
                return OR_E_NOT_PD;
            }
            double multiplier = f[p] / pivot;
            f[p] = multiplier;
            for (int u = dg[column] + 1; u < a->rowptr[column + 1]; u++) {
                int t = ilu_find(a, row, a->colidx[u]);
                if (t >= 0) f[t] -= multiplier * f[u];
            }
        }
        if (ilu_bad_pivot(f[dg[row]], pivot_tolerance)) {
            if (bad_row) *bad_row = row;
            free(dg); free(f);
            return OR_E_NOT_PD;
        }
    }
    or_ilu0 *pc = (or_ilu0 *)xcalloc(1, sizeof(or_ilu0));
    pc->a = a; pc->pivot_tolerance = pivot_tolerance; pc->factors = f; pc->diag = dg;
    *out = pc;
    return OR_OK;
}

/* Ilu0Preconditioner.Apply (:44-84) */
void or_ilu0_apply(const or_ilu0 *pc, const double *b, double *x)
{
    const or_gcsr *a = pc->a;
    int n = a->rows;
    double *w = (double *)xmalloc((size_t)n * sizeof(double));
    for (int row = 0; row < n; row++) {
        double sum = b[row];
        for (int p = a->rowptr[row]; p < pc->diag[row]; p++) sum -= pc->factors[p] * w[a->colidx[p]];
        w[row] = sum;
    }
    for (int row = n - 1; row >= 0; row--) {
        double sum = w[row];
        for (int p = pc->diag[row] + 1; p < a->rowptr[row + 1]; p++) sum -= pc->factors[p] * x[a->colidx[p]];
        x[row] = sum / pc->factors[pc->diag[row]];
    }
    free(w);
}

void or_ilu0_free(or_ilu0 *pc)
{
    if (!pc) return;
    free(pc->factors); free(pc->diag); free(pc);
}

static double bg_dot(int n, const double *l, const double *r)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += l[i] * r[i];
    return s;
}

static int bg_breakdown(double v, double tol) { return !isfinite(v) || fabs(v) < tol; }

/* BiCgStabSolver.SolveWithInitialGuess (BiCgStabSolver.cs:37-148) */
int or_bicgstab(const or_gcsr *a, const or_ilu0 *pre, int use_pre, const double *b, const double *x0, double tol,
                int max_iterations, double btol, double *x, or_pcg_result *res, int *fail_code, int *fail_it)
{
    if (fail_code) *fail_code = 0;
    if (fail_it) *fail_it = 0;
    if (!a || !b || !x || !res) return OR_E_NULL;
    if (!isfinite(tol) || tol < 0.0) return OR_E_RANGE;
    if (max_iterations < 0) return OR_E_RANGE;
    if (!isfinite(btol) || btol <= 0.0) return OR_E_RANGE;
    if (a->rows != a->cols) return OR_E_ARG;
    if (pre && pre->a != a) return OR_E_ARG;
    int n = a->rows;
    for (int i = 0; i < n; i++) x[i] = x0 ? x0[i] : 0.0;
    double bn = or_norm2(n, b);
    double scale = bn > 0.0 ? bn : 1.0;
    res->iterations = 0;
    res->residual_norm = 0.0;
    res->relative_residual = 0.0;
    res->converged = 1;
    if (n == 0) return OR_OK;
    double *r = (double *)xmalloc((size_t)n * sizeof(double));
    double *rh = (double *)xmalloc((size_t)n * sizeof(double));
    double *p = (double *)xmalloc((size_t)n * sizeof(double));
    double *ph = (double *)xmalloc((size_t)n * sizeof(double));
    double *v = (double *)xmalloc((size_t)n * sizeof(double));
    double *s = (double *)xmalloc((size_t)n * sizeof(double));
    double *sh = (double *)xmalloc((size_t)n * sizeof(double));
    double *t = (double *)xmalloc((size_t)n * sizeof(double));
    or_ilu0 *owned = NULL;
    int rc = OR_OK, iters = 0;
    or_gspmv(a, x, v);
    for (int i = 0; i < n; i++) r[i] = b[i] - v[i];
    memcpy(rh, r, (size_t)n * sizeof(double));
    double thr = tol * scale;
    double rn = or_norm2(n, r);
    if (rn <= thr) goto done;
    const or_ilu0 *M = pre;
    if (!M && use_pre) {
        int bad = -1;
        rc = or_ilu0_create(a, btol, &owned, &bad);
        if (rc != OR_OK) {
            if (fail_code) *fail_code = bad >= 0 && a->diag[bad] < 0 ? 6 : 7;
            if (fail_it) *fail_it = bad;
            goto done;
        }
        M = owned;
    }
    double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
    int maxit = max_iterations > 0 ? max_iterations : n;
    iters = maxit;
    for (int it = 1; it <= maxit; it++) {
        double rho = bg_dot(n, rh, r);
        if (bg_breakdown(rho, btol)) { rc = OR_E_NOT_PD; if (fail_code) *fail_code = 1; if (fail_it) *fail_it = it; goto done; }
        if (it == 1) {
            memcpy(p, r, (size_t)n * sizeof(double));
        } else {
            if (bg_breakdown(omega, btol)) { rc = OR_E_NOT_PD; if (fail_code) *fail_code = 2; if (fail_it) *fail_it = it; goto done; }
            double beta = rho / rho_prev * (alpha / omega);
            for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        if (M) or_ilu0_apply(M, p, ph); else memcpy(ph, p, (size_t)n * sizeof(double));
        or_gspmv(a, ph, v);
        double den = bg_dot(n, rh, v);
        if (bg_breakdown(den, btol)) { rc = OR_E_NOT_PD; if (fail_code) *fail_code = 3; if (fail_it) *fail_it = it; goto done; }
        alpha = rho / den;
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        double sn = or_norm2(n, s);
        if (sn <= thr) {
            for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
            rn = sn;
            iters = it;
            goto done;
        }
        if (M) or_ilu0_apply(M, s, sh); else memcpy(sh, s, (size_t)n * sizeof(double));
        or_gspmv(a, sh, t);
        double tt = bg_dot(n, t, t);
        if (bg_breakdown(tt, btol)) { rc = OR_E_NOT_PD; if (fail_code) *fail_code = 4; if (fail_it) *fail_it = it; goto done; }
        omega = bg_dot(n, t, s) / tt;
        if (bg_breakdown(omega, btol)) { rc = OR_E_NOT_PD; if (fail_code) *fail_code = 5; if (fail_it) *fail_it = it; goto done; }
        for (int i = 0; i < n; i++) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = s[i] - omega * t[i];
        }
        rn = or_norm2(n, r);
        if (rn <= thr) { iters = it; goto done; }
        rho_prev = rho;
    }
done:
    if (rc == OR_OK) {
        res->iterations = iters;
        res->residual_norm = rn;
        res->relative_residual = rn / scale;
        res->converged = res->relative_residual <= tol;
    }
    or_ilu0_free(owned);
    free(r); free(rh); free(p); free(ph); free(v); free(s); free(sh); free(t);
    return rc;
}

/* ======================================================================
 * Threaded factorization: the reference's schedule
 * (SupernodalMultifrontalFactorizer.Factorize, UtduNumericFactorization.cs:281-403,
 * Schedule :943-1023, ApplyPanelUpdate / UpdateTrailingColumns / UpdateColumnTile
 * :624-852).  T workers take ready supernodes from a LIFO stack (a supernode is
 * ready once all its children are done, :313-324, :1005-1018); a worker alone at the
 * top of the tree splits large trailing updates (work > parallel threshold, :660-673)
 * into min(trailing, degree * 8) column chunks that idle workers help with, degree =
 * T - active workers + 1 (:341).  The trailing update is blocked like the reference:
 * the panel times D laid out in blocks of 128 rows (:637-658), four columns at a time
 * with the rows vectorised (4 doubles, Vector<double>.Count on AVX2), each element's
 * sum over the panel taken in ascending k from 0.0 and subtracted once -- the same
 * operations, in the same order, as or_factorize, so the factor is bit-identical to
 * it for any T.  CPU timing baseline only.
 * ====================================================================== */
#include <pthread.h>

#define OR_ROW_BLOCK 128
#define OR_COL_TILE 4
#define OR_PARALLEL_THRESHOLD (1LL << 20) /* UtduSolverOptions.ParallelUpdateThreshold default */
typedef double or_v4 __attribute__((vector_size(32)));

typedef struct {
    double *front;
    const int *offs;
    const double *buf; /* buf[blk * RB * np + kk * RB + r] = L[pe + blk RB + r][ps + kk] * d */
    int f, ps, pe, np;
} mt_update;

static inline or_v4 v4_load(const double *p) { or_v4 v; memcpy(&v, p, sizeof(v)); return v; }
static inline void v4_sub_store(double *p, or_v4 a) { or_v4 v = v4_load(p); v = v - a; memcpy(p, &v, sizeof(v)); }

/* UpdateTrailingColumns (:693-762) for target columns [from, to) */
static void mt_update_columns(const mt_update *u, int from, int to)
{
    const int np = u->np, f = u->f, pe = u->pe, ps = u->ps;
    const int *offs = u->offs;
    double *front = u->front;
    double tile[128 * OR_COL_TILE];
    const int nblk = (f - pe + OR_ROW_BLOCK - 1) / OR_ROW_BLOCK;
    for (int blk = 0; blk < nblk; blk++) {
        const double *bb = u->buf + (size_t)blk * OR_ROW_BLOCK * np;
        const int r0 = pe + blk * OR_ROW_BLOCK;
        const int r1 = r0 + OR_ROW_BLOCK < f ? r0 + OR_ROW_BLOCK : f;
        const int jlim = to < r1 ? to : r1;
        int j = from;
        for (; j + OR_COL_TILE - 1 < jlim; j += OR_COL_TILE) {
            for (int k = 0; k < np; k++) /* GatherPanelTile (:771-784) */
                for (int t = 0; t < OR_COL_TILE; t++) tile[k * OR_COL_TILE + t] = front[offs[ps + k] - (ps + k) + j + t];
            const int full = (j + OR_COL_TILE - 1 > r0) ? j + OR_COL_TILE - 1 : r0;
            double *c0 = front + offs[j] - j, *c1 = front + offs[j + 1] - (j + 1);
            double *c2 = front + offs[j + 2] - (j + 2), *c3 = front + offs[j + 3] - (j + 3);
            int i = full;
            for (; i + 4 <= r1; i += 4) { /* UpdateColumnTile (:799-852), vector part */
                const double *pr = bb + (i - r0);
                or_v4 a0 = {0, 0, 0, 0}, a1 = a0, a2 = a0, a3 = a0;
                for (int k = 0; k < np; k++) {
                    const or_v4 p = v4_load(pr + (size_t)k * OR_ROW_BLOCK);
                    const double *sc = tile + k * OR_COL_TILE;
                    a0 += p * sc[0];
                    a1 += p * sc[1];
                    a2 += p * sc[2];
                    a3 += p * sc[3];
                }
                v4_sub_store(c0 + i, a0);
                v4_sub_store(c1 + i, a1);
                v4_sub_store(c2 + i, a2);
                v4_sub_store(c3 + i, a3);
            }
            for (; i < r1; i++) { /* scalar tail */
                const double *pr = bb + (i - r0);
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                for (int k = 0; k < np; k++) {
                    const double p = pr[(size_t)k * OR_ROW_BLOCK];
                    const double *sc = tile + k * OR_COL_TILE;
                    s0 += p * sc[0];
                    s1 += p * sc[1];
                    s2 += p * sc[2];
                    s3 += p * sc[3];
                }
                c0[i] -= s0; c1[i] -= s1; c2[i] -= s2; c3[i] -= s3;
            }
            const int lo = j > r0 ? j : r0; /* the corner rows (PanelDot, :787-795) */
            for (int ii = lo; ii < full; ii++)
                for (int t = 0; t <= ii - j; t++) {
                    double s = 0.0;
                    const double *pr = bb + (ii - r0);
                    for (int k = 0; k < np; k++) s += pr[(size_t)k * OR_ROW_BLOCK] * tile[k * OR_COL_TILE + t];
                    front[offs[j + t] + (ii - j - t)] -= s;
                }
        }
        for (; j < jlim; j++) { /* single columns */
            for (int k = 0; k < np; k++) tile[k * OR_COL_TILE] = front[offs[ps + k] - (ps + k) + j];
            double *cj = front + offs[j] - j;
            for (int i = (j > r0 ? j : r0); i < r1; i++) {
                const double *pr = bb + (i - r0);
                double s = 0.0;
                for (int k = 0; k < np; k++) s += pr[(size_t)k * OR_ROW_BLOCK] * tile[k * OR_COL_TILE];
                cj[i] -= s;
            }
        }
    }
}

/* shared state of one threaded factorization */
typedef struct mt_job {
    const mt_update *u;
    int chunks, trailing;
    int next, done; /* under the scheduler's lock */
    struct mt_job *link;
} mt_job;

typedef struct {
    const or_symbolic *s;
    const or_csr *a;
    const or_options *o;
    or_numeric *F;
    int T;
    double scale, thr, regv;
    int pw;
    double **updates;
    pthread_mutex_t mu;
    pthread_cond_t cv;
    int *stack, top, *remaining, completed, active, failed;
    mt_job *jobs; /* inner chunk jobs with chunks left */
    int regularized, negative;
    double smallest, largest;
} mt_state;

static void mt_run_chunk(const mt_job *jb, int c)
{
    const mt_update *u = jb->u;
    const int from = u->pe + (int)((int64_t)jb->trailing * c / jb->chunks);
    const int to = u->pe + (int)((int64_t)jb->trailing * (c + 1) / jb->chunks);
    mt_update_columns(u, from, to);
}

/* helps with one chunk of a posted job (lock held on entry and exit); 0 if none */
static int mt_help_locked(mt_state *S)
{
    mt_job *jb = S->jobs;
    if (!jb) return 0;
    const int c = jb->next++;
    if (jb->next >= jb->chunks) S->jobs = jb->link; /* no chunks left to hand out */
    pthread_mutex_unlock(&S->mu);
    mt_run_chunk(jb, c);
    pthread_mutex_lock(&S->mu);
    jb->done++;
    if (jb->done == jb->chunks) pthread_cond_broadcast(&S->cv);
    return 1;
}

/* Parallel.For over the chunks (:660-673): posts the job, runs chunks itself, waits for the rest */
static void mt_parallel_update(mt_state *S, const mt_update *u, int chunks, int trailing)
{
    mt_job jb = {u, chunks, trailing, 0, 0, NULL};
    pthread_mutex_lock(&S->mu);
    jb.link = S->jobs;
    S->jobs = &jb;
    pthread_cond_broadcast(&S->cv);
    while (jb.next < jb.chunks) {
        const int c = jb.next++;
        if (jb.next >= jb.chunks) { /* unlink: the last chunk is handed out */
            mt_job **pp = &S->jobs;
            while (*pp && *pp != &jb) pp = &(*pp)->link;
            if (*pp) *pp = jb.link;
        }
        pthread_mutex_unlock(&S->mu);
        mt_run_chunk(&jb, c);
        pthread_mutex_lock(&S->mu);
        jb.done++;
    }
    while (jb.done < jb.chunks) pthread_cond_wait(&S->cv, &S->mu);
    pthread_mutex_unlock(&S->mu);
}

typedef struct {
    mt_state *S;
    double *front, *buf;
    int *offs, *rowmap, *rel;
    int regularized, negative;
    double smallest, largest;
    int rc;
} mt_worker;

/* FactorizeSupernode (:420-531) of supernode q; the operations are or_factorize's */
static int mt_factor_supernode(mt_worker *W, int q, int degree)
{
    mt_state *S = W->S;
    const or_symbolic *s = S->s;
    const or_csr *a = S->a;
    or_numeric *F = S->F;
    double *front = W->front;
    int *offs = W->offs, *rowmap = W->rowmap, *rel = W->rel;
    const int first = s->first[q], w = s->first[q + 1] - first, f = s->pat_ptr[q + 1] - s->pat_ptr[q];
    const int *prow = s->pat_rows + s->pat_ptr[q];
    const int flen = (int)((int64_t)f * (f + 1) / 2);
    offs[0] = 0;
    for (int j = 0; j < f; j++) offs[j + 1] = offs[j] + (f - j);
    memset(front, 0, (size_t)flen * sizeof(double));
    for (int t = 0; t < f; t++) rowmap[prow[t]] = t;
    for (int loc = 0; loc < w; loc++) {
        const int col = first + loc, di = s->dvidx[col];
        if (di >= 0) front[offs[loc]] += a->values[di];
        const int lb = s->lvptr[col], st = s->lower_start[col], en = s->pptr[col + 1], cb = offs[loc] - loc;
        for (int p = st; p < en; p++) front[cb + rowmap[s->prows[p]]] += a->values[s->lvidx[lb + (p - st)]];
    }
    for (int c = s->child_ptr[q]; c < s->child_ptr[q + 1]; c++) {
        const int ch = s->children[c];
        double *upd = S->updates[ch];
        if (!upd) continue;
        const int cw = s->first[ch + 1] - s->first[ch], ct = (s->pat_ptr[ch + 1] - s->pat_ptr[ch]) - cw;
        const int *crow = s->pat_rows + s->pat_ptr[ch] + cw;
        for (int x = 0; x < ct; x++) rel[x] = rowmap[crow[x]];
        int src = 0;
        for (int x = 0; x < ct; x++) {
            const int tc = rel[x], tb = offs[tc] - tc;
            for (int y = x; y < ct; y++) front[tb + rel[y]] += upd[src + (y - x)];
            src += ct - x;
        }
        free(upd);
        S->updates[ch] = NULL;
    }
    const int pw = S->pw;
    for (int ps = 0; ps < w; ps += pw) {
        const int pe = ps + pw < w ? ps + pw : w;
        for (int k = ps; k < pe; k++) {
            const int co = offs[k], cl = f - k;
            double pivot = front[co];
            if (!(pivot > S->thr)) {
                if (!S->o->allow_regularization) return OR_E_NOT_PD;
                if (pivot < 0.0) W->negative++;
                W->regularized++;
                pivot = S->regv;
            }
            F->diag[first + k] = pivot;
            const double mag = fabs(pivot);
            if (mag < W->smallest) W->smallest = mag;
            if (mag > W->largest) W->largest = mag;
            const double inv = 1.0 / pivot;
            for (int t = 1; t < cl; t++) front[co + t] *= inv;
            for (int j = k + 1; j < pe; j++) {
                const double factor = pivot * front[co + (j - k)];
                if (factor == 0.0) continue;
                double *tgt = front + offs[j];
                const double *src = front + co + (j - k);
                for (int t = 0; t < f - j; t++) tgt[t] -= src[t] * factor;
            }
            front[co] = 1.0;
        }
        if (pe >= f) continue;
        const int np = pe - ps, trailing = f - pe, nblk = (trailing + OR_ROW_BLOCK - 1) / OR_ROW_BLOCK;
        double *buf = W->buf;
        for (int blk = 0; blk < nblk; blk++) {
            const int r0 = pe + blk * OR_ROW_BLOCK, nr = (f - r0) < OR_ROW_BLOCK ? (f - r0) : OR_ROW_BLOCK;
            for (int kk = 0; kk < np; kk++) {
                const int col = ps + kk;
                const double d = F->diag[first + col];
                const double *L = front + offs[col] - col;
                double *dst = buf + (size_t)blk * OR_ROW_BLOCK * np + (size_t)kk * OR_ROW_BLOCK;
                for (int r = 0; r < nr; r++) dst[r] = L[r0 + r] * d;
                for (int r = nr; r < OR_ROW_BLOCK; r++) dst[r] = 0.0;
            }
        }
        const mt_update u = {front, offs, buf, f, ps, pe, np};
        const int64_t work = (int64_t)trailing * trailing * np;
        if (degree > 1 && work > OR_PARALLEL_THRESHOLD) {
            const int chunks = trailing < degree * 8 ? trailing : degree * 8;
            mt_parallel_update(S, &u, chunks, trailing);
        } else {
            mt_update_columns(&u, pe, f);
        }
    }
    const int blen = offs[w];
    memcpy(F->factor + s->block_off[q], front, (size_t)blen * sizeof(double));
    if (f - w > 0) {
        const int ulen = flen - blen;
        double *upd = (double *)xmalloc((size_t)ulen * sizeof(double));
        memcpy(upd, front + blen, (size_t)ulen * sizeof(double));
        S->updates[q] = upd;
    }
    return OR_OK;
}

static void *mt_worker_main(void *arg)
{
    mt_worker *W = (mt_worker *)arg;
    mt_state *S = W->S;
    const or_symbolic *s = S->s;
    pthread_mutex_lock(&S->mu);
    for (;;) {
        if (S->failed || S->completed == s->nsuper) break;
        if (S->top > 0) {
            const int q = S->stack[--S->top];
            S->active++;
            const int degree = S->T - S->active + 1 > 1 ? S->T - S->active + 1 : 1;
            pthread_mutex_unlock(&S->mu);
            const int rc = mt_factor_supernode(W, q, degree);
            pthread_mutex_lock(&S->mu);
            S->active--;
            if (rc != OR_OK) {
                W->rc = rc;
                S->failed = 1;
                pthread_cond_broadcast(&S->cv);
                break;
            }
            S->completed++;
            const int p = s->sparent[q];
            if (p >= 0 && --S->remaining[p] == 0) S->stack[S->top++] = p;
            pthread_cond_broadcast(&S->cv);
            continue;
        }
        if (mt_help_locked(S)) continue;
        pthread_cond_wait(&S->cv, &S->mu);
    }
    pthread_mutex_unlock(&S->mu);
    return NULL;
}

int or_factorize_mt(const or_symbolic *s, const or_csr *a, const or_options *o, int threads, or_numeric **out)
{
    if (!s || !a || !o || !out) return OR_E_NULL;
    if (a->n != s->n || a->nnz != s->matrix_nnz) return OR_E_ARG;
    const int n = s->n, ns = s->nsuper;
    or_numeric *F = (or_numeric *)xcalloc(1, sizeof(or_numeric));
    F->sym = s;
    F->diag = (double *)xcalloc((size_t)n, sizeof(double));
    F->factor = (double *)xmalloc((size_t)s->factor_entries * sizeof(double));
    if (ns == 0) { *out = F; return OR_OK; }
    const int64_t maxlen = (int64_t)s->max_front * (s->max_front + 1) / 2;
    if (maxlen > 2147483647LL) { or_numeric_free(F); return OR_E_TOO_DENSE; }
    mt_state S;
    memset(&S, 0, sizeof(S));
    S.s = s; S.a = a; S.o = o; S.F = F;
    S.T = threads < 1 ? 1 : (threads < ns ? threads : ns); /* Math.Min(workers, supernodes) :309 */
    double scale = 0.0;
    for (int i = 0; i < n; i++) {
        const double d = a->diag[i] >= 0 ? a->values[a->diag[i]] : 0.0, m = fabs(d);
        if (isnan(m) || isnan(scale)) scale = NAN;
        else if (m > scale) scale = m;
    }
    if (!(scale > 0.0)) scale = 1.0;
    S.scale = scale;
    S.thr = o->pivot_tolerance * scale;
    S.regv = o->regularization * scale;
    S.pw = o->panel_width < 4 ? 4 : (o->panel_width > 128 ? 128 : o->panel_width);
    S.updates = (double **)xcalloc((size_t)ns, sizeof(double *));
    S.stack = (int *)xmalloc((size_t)ns * sizeof(int));
    S.remaining = (int *)xmalloc((size_t)ns * sizeof(int));
    for (int q = 0; q < ns; q++) {
        S.remaining[q] = s->child_ptr[q + 1] - s->child_ptr[q];
        if (S.remaining[q] == 0) S.stack[S.top++] = q;
    }
    pthread_mutex_init(&S.mu, NULL);
    pthread_cond_init(&S.cv, NULL);
    mt_worker *W = (mt_worker *)xcalloc((size_t)S.T, sizeof(mt_worker));
    pthread_t *tid = (pthread_t *)xcalloc((size_t)S.T, sizeof(pthread_t));
    const size_t bufn = (((size_t)s->max_front + OR_ROW_BLOCK - 1) / OR_ROW_BLOCK) * OR_ROW_BLOCK * (size_t)S.pw;
    for (int t = 0; t < S.T; t++) {
        W[t].S = &S;
        W[t].front = (double *)xmalloc((size_t)maxlen * sizeof(double));
        W[t].buf = (double *)xmalloc(bufn * sizeof(double));
        W[t].offs = (int *)xmalloc(((size_t)s->max_front + 1) * sizeof(int));
        W[t].rowmap = (int *)xmalloc((size_t)n * sizeof(int));
        W[t].rel = (int *)xmalloc((size_t)s->max_front * sizeof(int) + sizeof(int));
        W[t].smallest = DBL_MAX;
    }
    for (int t = 1; t < S.T; t++) pthread_create(&tid[t], NULL, mt_worker_main, &W[t]);
    mt_worker_main(&W[0]);
    for (int t = 1; t < S.T; t++) pthread_join(tid[t], NULL);
    int rc = OR_OK;
    double smallest = DBL_MAX, largest = 0.0;
    for (int t = 0; t < S.T; t++) {
        if (W[t].rc) rc = W[t].rc;
        F->regularized += W[t].regularized;
        F->negative += W[t].negative;
        if (W[t].smallest < smallest) smallest = W[t].smallest;
        if (W[t].largest > largest) largest = W[t].largest;
        free(W[t].front); free(W[t].buf); free(W[t].offs); free(W[t].rowmap); free(W[t].rel);
    }
    for (int q = 0; q < ns; q++) free(S.updates[q]);
    free(S.updates); free(S.stack); free(S.remaining); free(W); free(tid);
    pthread_mutex_destroy(&S.mu);
    pthread_cond_destroy(&S.cv);
    if (rc) { or_numeric_free(F); return rc; }
    F->smallest = smallest == DBL_MAX ? 0.0 : smallest;
    F->largest = largest;
    F->supernodes_done = ns;
    int64_t ops = 0;
    for (int q = 0; q < ns; q++) {
        const int64_t w = s->first[q + 1] - s->first[q], f = s->pat_ptr[q + 1] - s->pat_ptr[q];
        for (int64_t k = 0; k < w; k++) ops += (f - k) * (f - k);
    }
    F->ops_done = ops;
    *out = F;
    return OR_OK;
}

/* Solve (:90-112) of nrhs independent right-hand sides (column-major, leading dimension n)
 * over `threads` threads; each column is or_solve's serial solve (the reference solves
 * different right-hand sides with one factorization in parallel, SymmetricUtduSolver.cs:39-43). */
typedef struct { const or_numeric *F; const double *b; double *x; int n, nrhs, next; pthread_mutex_t *mu; } mt_solve_arg;

static void *mt_solve_main(void *p)
{
    mt_solve_arg *A = (mt_solve_arg *)p;
    for (;;) {
        pthread_mutex_lock(A->mu);
        const int c = A->next++;
        pthread_mutex_unlock(A->mu);
        if (c >= A->nrhs) break;
        or_solve(A->F, A->b + (size_t)c * A->n, A->x + (size_t)c * A->n);
    }
    return NULL;
}

int or_solve_mt(const or_numeric *F, int nrhs, const double *b, double *x, int threads)
{
    if (!F || (nrhs > 0 && (!b || !x))) return OR_E_NULL;
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    mt_solve_arg A = {F, b, x, F->sym->n, nrhs, 0, &mu};
    const int T = threads < 1 ? 1 : (threads < nrhs ? threads : (nrhs > 0 ? nrhs : 1));
    pthread_t *tid = (pthread_t *)xcalloc((size_t)T, sizeof(pthread_t));
    for (int t = 1; t < T; t++) pthread_create(&tid[t], NULL, mt_solve_main, &A);
    mt_solve_main(&A);
    for (int t = 1; t < T; t++) pthread_join(tid[t], NULL);
    free(tid);
    pthread_mutex_destroy(&mu);
    return OR_OK;
}
