"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong index or sign, transposed operand) fails: golden vectors of the
reference's tests and workbook, closed forms, dense numpy LDL^T / solves,
brute-force symbolic elimination, and hand traces of the reference's AMD.
"""
import numpy as np
import pytest

import oracle as O
from conftest import dense_coo, grid_matrix, random_matrix, tridiagonal, vector


def csr(n, r, c, v):
    return O.Csr.from_coo(n, r, c, v)


# ---- SymmetricCSRMatrix semantics (Tests/SymmetricSparseMatrixCSRTests.cs) ----

def test_coo_duplicates_sum_and_mirror_counts_twice():
    a = csr(3, [0, 0, 1, 2, 0], [0, 1, 0, 2, 0], [1.0, 2.0, 3.0, 4.0, 5.0])
    rp, ci, v = a.arrays()
    assert list(rp) == [0, 2, 2, 3] and list(ci) == [0, 1, 2]
    assert list(v) == [6.0, 5.0, 4.0]  # (0,1) and (1,0) both land on (0,1): 2 + 3


def test_coo_cancellation_is_dropped():
    a = csr(2, [0, 0, 0, 1], [1, 1, 0, 1], [1.0, -1.0, 2.0, 3.0])
    rp, ci, _ = a.arrays()
    assert list(ci) == [0, 1] and list(rp) == [0, 1, 2]
    b = csr(2, [0, 1, 0], [0, 1, 1], [1.0, 1.0, 1e-15])  # |v| <= 1e-15 dropped
    assert b.nnz == 2
    c = csr(2, [0, 1, 0], [0, 1, 1], [1.0, 1.0, 1.1e-15])
    assert c.nnz == 3


def test_coo_out_of_range_rejected():
    with pytest.raises(O.OracleError):
        csr(2, [0, 2], [0, 0], [1.0, 1.0])


def test_spmv_closed_form_and_dense():
    # tridiagonal 4/-1: A*ones = [3,2,...,2,3]
    a = csr(*tridiagonal(6, 4.0, -1.0))
    np.testing.assert_array_equal(a.spmv(np.ones(6)), [3, 2, 2, 2, 2, 3])
    n, r, c, v = random_matrix(60, 0.1, 3)
    a = csr(n, r, c, v)
    x = vector(n, 1)
    np.testing.assert_allclose(a.spmv(x), a.dense() @ x, rtol=1e-14, atol=1e-13)


def test_linecross_closed_form():
    # 3x3: [[2,-1,0],[-1,2,-1],[0,-1,2]], fix x1 = 5
    a = csr(3, [0, 1, 2, 0, 1], [0, 1, 2, 1, 2], [2.0, 2.0, 2.0, -1.0, -1.0])
    rhs = np.array([1.0, 1.0, 1.0])
    a.linecross(rhs, 5.0, 1)
    np.testing.assert_array_equal(rhs, [6.0, 5.0, 6.0])
    d = a.dense()
    np.testing.assert_array_equal(d, [[2, 0, 0], [0, 1, 0], [0, 0, 2]])
    assert a.nnz == 5  # structure kept


# ---- golden 8x8 system (CgComparisonData.cs + workbook) ----

def test_golden_natural_ldlt_matches_workbook(golden):
    A = np.array(golden["matrix"])
    a = csr(8, *dense_coo(A))
    s = O.Symbolic(a, O.options(ordering=0))
    # hand-derived from EliminationTree.cs:39-110 with the Natural ordering
    assert list(s.perm) == [0, 1, 3, 2, 4, 5, 6, 7]
    assert list(s.etree) == [-1, -1, -1, 4, 5, 6, 7, -1]
    f = O.Numeric(s, a)
    np.testing.assert_allclose(f.diag, golden["ldlt_diagonal"], rtol=1e-14)
    # L in natural numbering: elimination position p holds original column perm[p]
    L = np.eye(8)
    for q in range(s.nsuper):
        fs = s.pat_ptr[q + 1] - s.pat_ptr[q]
        w = s.first[q + 1] - s.first[q]
        rows = s.perm[s.pat_rows[s.pat_ptr[q]:s.pat_ptr[q + 1]]]
        off = s.block_off[q]
        for k in range(w):
            col = f.factor[off:off + fs - k]
            off += fs - k
            L[rows[k:], rows[k]] = col
    for i, j, val in golden["ldlt_lower"]:
        assert L[i, j] == pytest.approx(val, rel=1e-13)
    np.testing.assert_allclose(L @ np.diag(f.diag[np.argsort(s.perm)]) @ L.T, A, rtol=1e-13, atol=1e-12)
    x = f.solve(np.array(golden["rhs"]))
    np.testing.assert_allclose(x, golden["exact_solution"], atol=1e-14)


def test_golden_amd_and_inverse(golden):
    A = np.array(golden["matrix"])
    a = csr(8, *dense_coo(A))
    s = O.Symbolic(a)
    f = O.Numeric(s, a)
    x = f.solve(np.array(golden["rhs"]))
    np.testing.assert_allclose(x, golden["exact_solution"], atol=1e-14)
    ainv = np.column_stack([f.solve(e) for e in np.eye(8)])
    assert np.isclose(np.diag(ainv), golden["ainv_3_3"], rtol=1e-12).any()
    # det(A) = prod D for any ordering
    assert np.prod(f.diag) == pytest.approx(np.linalg.det(A), rel=1e-12)


def test_golden_cg_trace(golden):
    A = np.array(golden["matrix"])
    a = csr(8, *dense_coo(A))
    b = np.array(golden["rhs"])
    assert b @ b == golden["cg_r0_dot"]
    assert b @ (A @ b) == pytest.approx(golden["cg_p0_A_p0"], rel=1e-14)
    nz = golden["cg_nonzero_components"]
    for jacobi in (True, False):
        for k, key in ((3, "cg_x_after_3"), (4, "cg_x_after_4")):
            x, res = O.pcg(a, b, tol=1e-15, max_iterations=k, use_jacobi=jacobi)
            assert res["iterations"] == k and not res["converged"]
            np.testing.assert_allclose(x[nz], golden[key], rtol=1e-12)
        x1, res1 = O.pcg(a, b, tol=1e-15, max_iterations=1, use_jacobi=jacobi)
        assert res1["residual_norm"] == pytest.approx(golden["cg_r1_norm"], rel=1e-12)
        np.testing.assert_allclose(x1, golden["cg_alpha0"] * b, rtol=1e-12)


# ---- closed forms and dense references ----

@pytest.mark.parametrize("diag", [[2.0, 4.0, 5.0], [4.0], [1e-6, 1e6]])
def test_diagonal_identity_factor(diag):
    n = len(diag)
    a = csr(n, np.arange(n), np.arange(n), np.array(diag))
    s = O.Symbolic(a)
    assert list(s.perm) == list(range(n))  # isolated vertices emitted in index order
    f = O.Numeric(s, a)
    np.testing.assert_array_equal(f.diag, diag)
    np.testing.assert_array_equal(f.factor, np.ones(n))


@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("n,density,seed", [(40, 0.02, 1), (55, 0.15, 2), (45, 0.9, 3), (200, 0.03, 4)])
def test_factor_equals_dense_ldlt(n, density, seed, ordering):
    a = csr(*random_matrix(n, density, seed))
    s = O.Symbolic(a, O.options(ordering=ordering))
    f = O.Numeric(s, a)
    Ad = a.dense()
    P = Ad[np.ix_(s.perm, s.perm)]
    # dense LDL^T without pivoting of the permuted matrix
    Lr = np.linalg.cholesky(P)
    dref = np.diag(Lr) ** 2
    np.testing.assert_allclose(f.diag, dref, rtol=1e-12)
    b = vector(n, 7)
    x = f.solve(b)
    np.testing.assert_allclose(x, np.linalg.solve(Ad, b), rtol=1e-10, atol=1e-12)


def test_pivot_handling_exact_counts():
    # [[1,1],[1,1]] (+) 2 (+) 3: one regularized pivot, none negative; off -> error
    a = csr(4, [0, 1, 0, 2, 3], [0, 1, 1, 2, 3], [1.0, 1.0, 1.0, 2.0, 3.0])
    f = O.Numeric(O.Symbolic(a), a)
    assert (f.regularized, f.negative) == (1, 0)
    assert 1e-10 * 3.0 in list(f.diag)
    with pytest.raises(O.OracleError):
        O.Numeric(O.Symbolic(a, O.options(allow_regularization=False)), a)
    # [[1,3],[3,1]] (+) 1: last pivot -8 -> regularized and negative
    b = csr(3, [0, 1, 2, 0], [0, 1, 2, 1], [1.0, 1.0, 1.0, 3.0])
    g = O.Numeric(O.Symbolic(b), b)
    assert (g.regularized, g.negative) == (1, 1)
    assert g.smallest == pytest.approx(1e-10) and g.largest == 1.0


def test_zero_rhs_exact_and_ordering_invariance():
    n, r, c, v = grid_matrix(5, 5, 5, 2)
    a = csr(n, r, c, v)
    f = O.Numeric(O.Symbolic(a), a)
    assert np.all(f.solve(np.zeros(n)) == 0.0)
    b = vector(n, 23)
    xn = O.Numeric(O.Symbolic(a, O.options(ordering=0)), a).solve(b)
    np.testing.assert_allclose(f.solve(b), xn, rtol=1e-11, atol=1e-13)


# ---- symbolic phase against brute force ----

def brute_symbolic(Ad_perm):
    """Symbolic Cholesky by set unions: column structures of L and the etree."""
    n = Ad_perm.shape[0]
    struct = [set(np.nonzero(Ad_perm[j + 1:, j])[0] + j + 1) for j in range(n)]
    parent = [-1] * n
    for j in range(n):
        if struct[j]:
            p = min(struct[j])
            parent[j] = p
            struct[p] |= {i for i in struct[j] if i > p}
    counts = [len(st) + 1 for st in struct]
    return parent, counts, struct


@pytest.mark.parametrize("args", [random_matrix(120, 0.03, 5), grid_matrix(4, 4, 3, 2), tridiagonal(50, 4, -1),
                                  random_matrix(60, 0.2, 6)])
@pytest.mark.parametrize("ordering", [0, 1])
def test_etree_colcounts_patterns_brute_force(args, ordering):
    a = csr(*args)
    s = O.Symbolic(a, O.options(ordering=ordering))
    Ad = a.dense() != 0
    parent, counts, struct = brute_symbolic(Ad[np.ix_(s.perm, s.perm)].astype(float))
    assert list(s.etree) == parent
    assert list(s.colcount) == counts
    assert s.strict_nonzeros == sum(counts)
    # postorder property: children precede parents, subtrees contiguous
    assert all(p == -1 or p > j for j, p in enumerate(parent))
    # supernode patterns = union of own columns and the structures of their columns
    for q in range(s.nsuper):
        cols = range(s.first[q], s.first[q + 1])
        exp = set(cols)
        for j in cols:
            exp |= struct[j]
        got = s.pat_rows[s.pat_ptr[q]:s.pat_ptr[q + 1]]
        assert list(got) == sorted(exp)
    # cost formulas (SupernodalStructure.cs:370-418)
    f = np.diff(s.pat_ptr).astype(np.int64)
    w = np.diff(s.first).astype(np.int64)
    assert s.factor_entries == int(np.sum(f * w - w * (w - 1) // 2))
    assert s.factor_ops == int(sum(((f[q] - np.arange(w[q])) ** 2).sum() for q in range(s.nsuper)))


def graph_csr(n, edges):
    """Upper CSR of a graph (unit off-diagonals, dominant diagonal): only the pattern matters."""
    r = np.array([min(e) for e in edges] + list(range(n)), dtype=np.int64)
    c = np.array([max(e) for e in edges] + list(range(n)), dtype=np.int64)
    v = np.concatenate([np.full(len(edges), -1.0), np.full(n, 10.0)])
    return csr(n, r, c, v)


def amd_of(a):
    rp, ci, _ = a.arrays()
    return [int(x) for x in O.amd(a.n, rp, ci)]


# AMD orders traced by hand through ApproximateMinimumDegreeOrdering.Run
# (AMD/ApproximateMinimumDegreeOrdering.cs:157-224) with the reference's data structures:
# neighbour lists ascending (SymmetricPatternGraph.FromMatrix, AMD/SymmetricPatternGraph.cs:71-121),
# degree buckets pushed and popped at the head (AMD/DegreeBuckets.cs:57-69,96-110), so the last
# vertex inserted with the minimum degree is the next pivot.
AMD_KNOWN = {
    # 4 (last of bucket 1) first; each pivot's element absorbs the previous one (:384-389); pivot 1
    # mass-eliminates 0, emitted right after it (:494-507, MemberLists.Append)
    "path5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)], [4, 3, 2, 1, 0]),
    # leaves 7, 6, ... 2 pivot in turn (each older element aggressively absorbed, :466-472); the
    # centre's degree falls to 1 and, inserted after leaf 1, it pops first and mass-eliminates 1
    "star8": (8, [(0, k) for k in range(1, 8)], [7, 6, 5, 4, 3, 2, 0, 1]),
    # components {0, 2} and {1, 3, 4}: 4 pivots, then 3 (degree 1 after 4) mass-eliminates 1;
    # 2 pops before 0 (LIFO) and mass-eliminates it
    "two_components": (5, [(0, 2), (1, 3), (3, 4)], [4, 3, 1, 2, 0]),
    # the 8x8 CgComparisonData pattern (Tests/BiCGSTAB/CgComparisonData.cs:8-18): isolated 0, 1, 3
    # are emitted at once (:243-253); pivot 6 makes 4 and 5 indistinguishable (same lists after
    # the element, :531-592), supervariable 5 = {5, 4} pops next and mass-eliminates 7 and 2
    "cg8": (8, [(2, 4), (2, 5), (2, 7), (4, 5), (4, 6), (4, 7), (5, 6), (5, 7)], [0, 1, 3, 6, 5, 4, 7, 2]),
}


@pytest.mark.parametrize("name", sorted(AMD_KNOWN))
def test_amd_hand_traces(name):
    n, edges, order = AMD_KNOWN[name]
    assert amd_of(graph_csr(n, edges)) == order


def test_amd_hand_traces_golden_pattern_matches_fixture(golden):
    A = np.array(golden["matrix"])
    r, c = np.nonzero(np.triu(A, 1))
    assert sorted(zip(r.tolist(), c.tolist())) == sorted(AMD_KNOWN["cg8"][1])


def test_amd_isolated_and_dense_rows():
    # isolated vertices first, then pivot 1 mass-eliminates 0 (singular test matrix, natural ids)
    a = csr(4, [0, 1, 0, 2, 3], [0, 1, 1, 2, 3], [1.0, 1.0, 1.0, 2.0, 3.0])
    assert amd_of(a) == [2, 3, 1, 0]
    # star with a dense centre (degree 299 > max(16, 10 sqrt(300)), :233-261): the centre is left
    # out and ordered last; every leaf then has no live neighbour, an empty element, and the
    # leaves pop from the degree-1 bucket last-inserted first: 299, 298, ..., 1, then 0
    edges = [(0, k) for k in range(1, 300)]
    assert amd_of(graph_csr(300, edges)) == list(range(299, 0, -1)) + [0]


def band_csr(n, b):
    r, c = [], []
    for i in range(n):
        for j in range(i, min(n, i + b + 1)):
            r.append(i)
            c.append(j)
    r, c = np.array(r), np.array(c)
    return csr(n, r, c, np.where(r == c, 4.0 * b, -1.0))


# Supernode boundaries derived by hand from SupernodalStructure.Amalgamate / IsMergeWorthwhile
# (UCHOL/SupernodalStructure.cs:175-286), natural order.  In a band of half-width b every column
# before the last b + 1 is its own fundamental supernode (column counts b + 1, :175-193) and the
# dense tail of b + 1 columns is one.  Merging k single columns stores k(k+b) - k(k-1)/2 entries
# for k(b+1) nonzeros: added zeros k(k-1)/2, a fraction (k-1)/(k+2b+1).  Tridiagonal (b = 1):
# 0.2, 0.33, 0.43 for 2, 3, 4 columns (<= 0.8: merged), 0.5 for 5 (> 0.1: not).  b = 100: 16
# columns accepted (0.069 <= 0.1), 17 rejected (0.073 > 0.05).  b = 500: 48 accepted
# (0.045 <= 0.05), 49 rejected (more than 48 columns with any added zero).  The amalgamation
# limit (64 or 32 columns) caps each merge, and a run before the tail stops at the tail.
AMALGAMATION_KNOWN = [
    ("tridiagonal10", lambda: band_csr(10, 1), 64, [0, 4, 8, 10]),
    ("band300_100", lambda: band_csr(300, 100), 64, list(range(0, 193, 16)) + [199, 300]),
    ("band1200_500", lambda: band_csr(1200, 500), 64, list(range(0, 673, 48)) + [699, 1200]),
    ("band1200_500_limit32", lambda: band_csr(1200, 500), 32, list(range(0, 673, 32)) + [699, 1200]),
]


@pytest.mark.parametrize("name,make,limit,first", AMALGAMATION_KNOWN, ids=[k[0] for k in AMALGAMATION_KNOWN])
def test_amalgamation_known_answers(name, make, limit, first):
    s = O.Symbolic(make(), O.options(ordering=0, amalgamation_limit=limit))
    assert [int(x) for x in s.first] == first


def test_amd_reduces_fill_on_grid():
    # SymmetricUtduSolverTests.cs:438-462
    a = csr(*grid_matrix(12, 12, 12, 1))
    nat = O.Symbolic(a, O.options(ordering=0))
    amd = O.Symbolic(a)
    assert amd.strict_nonzeros * 2 < nat.strict_nonzeros
    assert amd.factor_entries < amd.strict_nonzeros * 1.2


def test_pcg_exact_counts_on_diagonal():
    a = csr(3, [0, 1, 2], [0, 1, 2], [2.0, 4.0, 5.0])
    b = a.spmv(np.array([3.0, 1.0, 2.0]))
    x, res = O.pcg(a, b)
    assert res["iterations"] == 1 and res["converged"]
    x, res = O.pcg(a, np.zeros(3))
    assert res["iterations"] == 0
    a2 = csr(2, [0, 1], [0, 1], [2.0, 3.0])
    sol = np.array([5.0, -2.0])
    x, res = O.pcg(a2, a2.spmv(sol), x0=sol)
    assert res["iterations"] == 0


@pytest.mark.parametrize("threads", [2, 5, 8])
def test_threaded_factorization_is_bit_identical(threads):
    """The CPU baseline (the reference's tree schedule, parallel trailing-update chunks and blocked
    update, or_factorize_mt) computes every element with the serial restatement's operations in
    the same order: factor, D and pivot statistics equal bit for bit, as do its solves."""
    from magla_amd import generators as G
    for args in (grid_matrix(9, 9, 9, 2), random_matrix(300, 1.0, 8), G.elasticity_3d(6, 2)):
        a = csr(*args)
        s = O.Symbolic(a)
        f1, ft = O.Numeric(s, a), O.Numeric(s, a, threads=threads)
        assert np.array_equal(f1.factor, ft.factor) and np.array_equal(f1.diag, ft.diag)
        assert (f1.regularized, f1.negative, f1.smallest, f1.largest) == \
            (ft.regularized, ft.negative, ft.smallest, ft.largest)
        B = np.column_stack([vector(a.n, 3 + j) for j in range(3)])
        X = ft.solve_many(B, threads=threads)
        for j in range(3):
            np.testing.assert_array_equal(X[:, j], f1.solve(B[:, j]))
