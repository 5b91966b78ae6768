"""Host-side product code (no GPU): the C ABI library loads and exports every
declared symbol; the symbolic phase equals the oracle's independent
restatement bit for bit; COO assembly and LineCross follow the reference."""
import os
import re

import numpy as np
import pytest

import magla_amd as M
import oracle as O
from conftest import ROOT, grid_matrix, random_matrix, tridiagonal
from magla_amd import generators as G


def declared_symbols():
    names = set()
    for h in ("magla.h", "magla_testing.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(mg_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = M.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_options_defaults_mirror_reference():
    o = M.UtduSolverOptions()._c()
    assert (o.ordering, o.supernode_amalgamation_limit, o.panel_width, o.allow_diagonal_regularization) == (1, 64, 32, 1)
    assert (o.pivot_tolerance, o.diagonal_regularization, o.parallel_update_threshold) == (1e-13, 1e-10, 1 << 20)


CASES = [grid_matrix(7, 6, 5, 2), random_matrix(250, 0.02, 99), random_matrix(45, 0.9, 3), tridiagonal(400, 4, -1),
         G.laplacian_2d(64), G.laplacian_2d(33, nine_point=True), G.diffusion_3d_lognormal(14, 1),
         G.elasticity_3d(4, 2)]


@pytest.mark.parametrize("idx", range(len(CASES)))
@pytest.mark.parametrize("ordering", [M.NATURAL, M.AMD])
def test_symbolic_equals_oracle(idx, ordering):
    n, r, c, v = CASES[idx]
    pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    oc = O.Csr.from_coo(n, r, c, v)
    rp, ci, va = oc.arrays()
    np.testing.assert_array_equal(pm.row_pointers, rp)
    np.testing.assert_array_equal(pm.column_indices, ci)
    np.testing.assert_array_equal(pm.values, va)
    sym = M.SymmetricUtduSolver(M.UtduSolverOptions(ordering=ordering)).analyze(pm)
    st = sym.structure()
    os_ = O.Symbolic(oc, O.options(ordering=ordering))
    np.testing.assert_array_equal(st["fill_order"], os_.amd)
    np.testing.assert_array_equal(sym.permutation, os_.perm)
    np.testing.assert_array_equal(sym.inverse_permutation, os_.iperm)
    np.testing.assert_array_equal(st["etree"], os_.etree)
    np.testing.assert_array_equal(st["colcount"], os_.colcount)
    np.testing.assert_array_equal(st["first"], os_.first)
    np.testing.assert_array_equal(st["sparent"], os_.sparent)
    np.testing.assert_array_equal(st["pat_ptr"], os_.pat_ptr)
    np.testing.assert_array_equal(st["pat_rows"], os_.pat_rows)
    inf = sym.info()
    assert inf["factor_entry_count"] == os_.factor_entries
    assert inf["factor_operation_count"] == os_.factor_ops
    assert inf["strict_factor_nonzero_count"] == os_.strict_nonzeros
    assert inf["max_front_size"] == os_.max_front
    assert inf["peak_contribution_bytes"] == 8 * os_.peak_contribution


def test_symbolic_on_random_graphs_many_seeds():
    rng = np.random.default_rng(0)
    for trial in range(150):
        n = int(rng.integers(1, 120))
        dens = float(rng.uniform(0.0, 0.3))
        args = random_matrix(n, dens, int(rng.integers(1 << 30)))
        pm = M.SymmetricCSRMatrix.from_coo(*args)
        oc = O.Csr.from_coo(*args)
        sym = M.SymmetricUtduSolver().analyze(pm)
        os_ = O.Symbolic(oc)
        assert np.array_equal(sym.permutation, os_.perm), trial
        st = sym.structure()
        assert np.array_equal(st["pat_rows"], os_.pat_rows), trial


def test_cost_estimates_consistent():
    # SymmetricUtduSolverTests.cs:662-686
    pm = M.SymmetricCSRMatrix.from_coo(*grid_matrix(10, 10, 10, 2))
    sym = M.SymmetricUtduSolver().analyze(pm)
    inf = sym.info()
    assert inf["size"] == pm.size and inf["supernode_count"] > 0
    assert inf["strict_factor_nonzero_count"] >= pm.nonzero_count
    assert inf["factor_entry_count"] >= inf["strict_factor_nonzero_count"]
    assert inf["estimated_factor_bytes"] == 8 * inf["factor_entry_count"]
    assert sym.estimate_peak_bytes(4) > inf["estimated_factor_bytes"]
    assert inf["factor_operation_count"] > inf["strict_factor_nonzero_count"]
    assert 0 < inf["max_front_size"] <= pm.size
    assert sorted(sym.permutation) == list(range(pm.size))
    assert np.array_equal(sym.inverse_permutation[sym.permutation], np.arange(pm.size))


@pytest.mark.parametrize("scenario", range(4))
def test_amd_valid_permutation(scenario):
    # SymmetricUtduSolverTests.cs:469-505
    if scenario == 0:
        args = grid_matrix(7, 6, 5, 2)
    elif scenario == 1:
        args = random_matrix(250, 0.02, 99)
    elif scenario == 2:
        args = (60, np.arange(60), np.arange(60), np.arange(1, 61, dtype=float))
    else:
        r = np.concatenate([np.arange(99), np.arange(100, 199), np.arange(200)])
        c = np.concatenate([np.arange(1, 100), np.arange(101, 200), np.arange(200)])
        v = np.concatenate([np.full(198, -1.0), np.full(200, 4.0)])
        args = (200, r, c, v)
    pm = M.SymmetricCSRMatrix.from_coo(*args)
    sym = M.SymmetricUtduSolver().analyze(pm)
    assert sorted(sym.permutation) == list(range(pm.size))


def test_coo_semantics_equal_oracle_randomized():
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = int(rng.integers(1, 40))
        cnt = int(rng.integers(0, 300))
        r = rng.integers(0, n, cnt)
        c = rng.integers(0, n, cnt)
        v = rng.choice([1.0, -1.0, 0.5, 1e-16, 2.0], cnt)
        pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
        rp, ci, va = O.Csr.from_coo(n, r, c, v).arrays()
        assert np.array_equal(pm.row_pointers, rp) and np.array_equal(pm.column_indices, ci)
        assert np.array_equal(pm.values, va)


def test_coo_out_of_range_raises():
    with pytest.raises(M.ArgumentOutOfRangeError):
        M.SymmetricCSRMatrix.from_coo(2, [0, 2], [0, 0], [1.0, 1.0])


def test_line_cross_equals_oracle_sequence():
    n, r, c, v = grid_matrix(4, 4, 3, 2)
    pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    oc = O.Csr.from_coo(n, r, c, v)
    rhs = np.linspace(-1, 1, n)
    rhs2 = rhs.copy()
    idx = np.array([5, 0, 17, 30, 2], dtype=np.int32)
    pv = np.array([1.5, -2.0, 0.25, 3.0, 0.0])
    pm.line_cross(rhs, pv, idx)
    for i, p in zip(idx, pv):
        oc.linecross(rhs2, p, int(i))
    np.testing.assert_array_equal(rhs, rhs2)
    np.testing.assert_array_equal(pm.values, oc.arrays()[2])


def test_elasticity_generator_shape():
    n, r, c, v = G.elasticity_3d(3, 1)
    assert n == 3 * 4 ** 3
    pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    oc = O.Csr.from_coo(n, r, c, v)
    Ad = oc.dense()
    # rigid translations are in the null space of the unconstrained stiffness
    for d in range(3):
        t = np.zeros(n)
        t[d::3] = 1.0
        assert np.abs(Ad @ t).max() < 1e-12 * np.abs(Ad).max()
    assert np.all(np.linalg.eigvalsh(Ad) > -1e-10)
    assert pm.nonzero_count == oc.nnz


def test_batch_layout_and_costs_are_sums_of_systems():
    sysl = G.config_c4(16, seed=3, side_lo=5, side_hi=20, max_rhs=3)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    batch = M.UtduBatch(mats)
    info = batch.info()
    each = [M.SymmetricUtduSolver().analyze(m).info() for m in mats]
    assert info["size"] == sum(e["size"] for e in each)
    assert info["supernode_count"] == sum(e["supernode_count"] for e in each)
    assert info["factor_entry_count"] == sum(e["factor_entry_count"] for e in each)
    assert info["factor_operation_count"] == sum(e["factor_operation_count"] for e in each)
    assert info["max_front_size"] == max(e["max_front_size"] for e in each)
    np.testing.assert_array_equal(batch.row_begin, np.concatenate([[0], np.cumsum([m.size for m in mats])]))
    np.testing.assert_array_equal(batch.nnz_begin, np.concatenate([[0], np.cumsum([m.nonzero_count for m in mats])]))


# ---- general CSR assembly (host code of the library, no GPU) ----

def test_general_coo_semantics_equal_oracle_randomized():
    g = np.random.default_rng(17)
    for trial in range(20):
        rows, cols = int(g.integers(1, 30)), int(g.integers(1, 30))
        k = int(g.integers(0, 200))
        ri = g.integers(0, rows, k)
        cj = g.integers(0, cols, k)
        v = g.choice([1.0, -1.0, 0.5, 1e-16, 2.0], k)
        a = M.CSRMatrix.from_coo(rows, cols, ri, cj, v)
        o = O.GCsr.from_coo(rows, cols, ri, cj, v)
        rp, ci, vv = o.arrays()
        np.testing.assert_array_equal(a.row_pointers, rp)
        np.testing.assert_array_equal(a.column_indices, ci)
        np.testing.assert_array_equal(a.values, vv)


def test_general_coo_out_of_range_raises():
    with pytest.raises(M.ArgumentOutOfRangeError):
        M.CSRMatrix.from_coo(2, 3, [0, 2], [0, 0], [1.0, 1.0])


def test_general_line_cross_equals_oracle():
    g = np.random.default_rng(5)
    A = (g.random((12, 12)) < 0.4) * g.uniform(-1, 1, (12, 12))
    np.fill_diagonal(A, 3.0)
    a = M.CSRMatrix.from_dense(A)
    o = O.GCsr.from_dense(A)
    r1 = g.uniform(-1, 1, 12)
    r2 = r1.copy()
    for idx, pv in [(3, 1.5), (7, -2.0), (3, 0.25)]:
        a.line_cross(r1, pv, idx)
        o.linecross(r2, pv, idx)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(a.values, o.arrays()[2])
    b = M.CSRMatrix.from_dense(np.array([[1.0, 2.0], [3.0, 0.0]]))
    with pytest.raises(M.InvalidOperationError):
        b.line_cross(np.zeros(2), 1.0, 1)  # structurally zero diagonal


# ---- known answers (hand-traced from the reference) and full-size workloads ----

def _product_symbolic(a_oracle, ordering=M.AMD, limit=64):
    rp, ci, va = a_oracle.arrays()
    pm = M.SymmetricCSRMatrix(a_oracle.n, rp.copy(), ci.copy(), va.copy())
    return M.SymmetricUtduSolver(M.UtduSolverOptions(ordering=ordering, supernode_amalgamation_limit=limit)).analyze(pm)


def test_product_amd_matches_hand_traces():
    from test_oracle import AMD_KNOWN, graph_csr
    for name, (n, edges, order) in AMD_KNOWN.items():
        sym = _product_symbolic(graph_csr(n, edges))
        assert [int(x) for x in sym.structure()["fill_order"]] == order, name
    sym = _product_symbolic(graph_csr(300, [(0, k) for k in range(1, 300)]))
    assert [int(x) for x in sym.structure()["fill_order"]] == list(range(299, 0, -1)) + [0]


def test_product_amalgamation_matches_known_answers():
    from test_oracle import AMALGAMATION_KNOWN
    for name, make, limit, first in AMALGAMATION_KNOWN:
        sym = _product_symbolic(make(), ordering=M.NATURAL, limit=limit)
        assert [int(x) for x in sym.structure()["first"]] == first, name


def _assert_symbolic_equal(pm, oc, label):
    sym = M.SymmetricUtduSolver().analyze(pm)
    st = sym.structure()
    os_ = O.Symbolic(oc)
    for key, mine, ref in [("fill_order", st["fill_order"], os_.amd), ("perm", sym.permutation, os_.perm),
                           ("iperm", sym.inverse_permutation, os_.iperm), ("etree", st["etree"], os_.etree),
                           ("colcount", st["colcount"], os_.colcount), ("first", st["first"], os_.first),
                           ("sparent", st["sparent"], os_.sparent), ("pat_ptr", st["pat_ptr"], os_.pat_ptr),
                           ("pat_rows", st["pat_rows"], os_.pat_rows)]:
        assert np.array_equal(mine, ref), (label, key)
    inf = sym.info()
    assert (inf["factor_entry_count"], inf["factor_operation_count"], inf["peak_contribution_bytes"]) == \
        (os_.factor_entries, os_.factor_ops, 8 * os_.peak_contribution), label


@pytest.mark.parametrize("config", ["c1", "c2", "c3", "c5"])
def test_symbolic_equals_oracle_full_size(config):
    """The benchmark workloads at full size: permutation, elimination tree, column counts,
    supernodes, patterns and costs bit for bit against the oracle."""
    if config == "c3":
        n, r, c, v, clamped, _ = G.config_c3()
        pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
        pm.line_cross(np.zeros(n), 0.0, clamped)
    else:
        n, r, c, v, _ = {"c1": G.config_c1, "c2": G.config_c2, "c5": G.config_c5}[config]()
        pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    oc = O.Csr.from_csr(n, pm.row_pointers, pm.column_indices, pm.values)
    _assert_symbolic_equal(pm, oc, config)


def test_symbolic_equals_oracle_all_batched_systems():
    """All 1,024 systems of the batched workload, each analysed by the product and the oracle;
    the batch's merged analysis keeps each system's permutation."""
    sysl = G.config_c4()
    mats = []
    for t, (n, r, c, v, _) in enumerate(sysl):
        pm = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
        _assert_symbolic_equal(pm, O.Csr.from_coo(n, r, c, v), t)
        mats.append(pm)
    batch = M.UtduBatch(mats)
    assert batch.count == 1024 and batch.info()["size"] == sum(m.size for m in mats)


def _contrib_slots(t):
    """Slots of a t x t contribution block in the pool: (r, c) with r >= 64 floor(c / 64), counted one by one."""
    return sum(t - 64 * (c // 64) for c in range(t))


@pytest.mark.parametrize("f,w", [(1, 1), (2, 1), (5, 5), (70, 3), (64 + 3, 3), (133, 5), (200, 70), (300, 300),
                                 (257, 1), (64 * 3 + 1, 64), (130, 2)])
def test_front_entries_map_one_to_one_into_their_regions(f, w):
    """Every lower entry (i >= j) of a front gets its own arena slot; pivot columns are the f x w
    column-major factor block at off, the trailing block fills [coff, coff + slots) with no gap."""
    off, coff = 1000, 10 ** 7
    i, j = np.tril_indices(f)
    addr = M.front_addr(f, w, off, coff, i, j)
    assert len(np.unique(addr)) == len(addr)
    fac = j < w
    np.testing.assert_array_equal(addr[fac], off + j[fac].astype(np.int64) * f + i[fac])
    t = f - w
    con = addr[~fac]
    assert len(con) == t * (t + 1) // 2
    if t > 0:
        assert con.min() == coff and con.max() == coff + _contrib_slots(t) - 1
        # a column's entries are consecutive rows
        for c in (0, t // 2, t - 1):
            col = addr[(j == w + c)]
            np.testing.assert_array_equal(np.diff(col), 1)


LAYOUT_CASES = [G.diffusion_3d_lognormal(12, 3), G.elasticity_3d(5, 1), G.laplacian_2d(40, nine_point=True),
                random_matrix(300, 0.02, 5)]


@pytest.mark.parametrize("idx", range(len(LAYOUT_CASES)))
def test_plan_layout_regions_never_overlap_while_live(idx):
    """Factor blocks tile the factor store; a contribution block lives from its front's level to
    its parent's (both included), and blocks live at the same level never share pool slots."""
    n, r, c, v = LAYOUT_CASES[idx]
    if idx == 3:  # two disconnected copies: a forest with two roots
        n, r, c, v = 2 * n, np.concatenate([r, r + n]), np.concatenate([c, c + n]), np.concatenate([v, v])
    sym = M.SymmetricUtduSolver().analyze(M.SymmetricCSRMatrix.from_coo(n, r, c, v))
    st = M.plan_stats(sym)
    lay = M.plan_layout(sym)
    f, w, lev, par, off, coff = (lay[:, k] for k in range(6))
    assert st["factor_elems"] == int((f * w).sum())
    order = np.argsort(off)
    assert off[order[0]] == 0
    np.testing.assert_array_equal(off[order[1:]], (off + f * w)[order[:-1]])
    assert (lev[par >= 0] < lev[par[par >= 0]]).all()
    assert (par >= 0).sum() < len(par)  # roots exist
    t = f - w
    slots = np.array([_contrib_slots(int(x)) for x in t])
    pool0 = st["factor_elems"]
    end = np.where(par >= 0, lev[np.maximum(par, 0)], lev)
    assert (coff[slots > 0] >= pool0).all()
    assert ((coff + slots)[slots > 0] <= st["arena_elems"]).all()
    assert (coff + slots).max() - coff[slots > 0].min() <= st["pool_elems"]
    peak = 0
    for L in range(st["levels"]):
        live = np.nonzero((slots > 0) & (lev <= L) & (end >= L))[0]
        iv = sorted(zip(coff[live], coff[live] + slots[live]))
        for (a0, a1), (b0, b1) in zip(iv, iv[1:]):
            assert a1 <= b0, (L, a0, a1, b0, b1)
        peak = max(peak, int(slots[live].sum()))
    assert peak <= st["pool_live_elems"] <= st["pool_elems"]


def test_plan_layout_c3_within_the_reference_memory_estimate():
    """3D elasticity 40^3: factor store plus contribution pool in FP64 stay within 1.5x the
    reference's estimate (factor entries + peak contribution entries)."""
    n, r, c, v, clamped, b = G.config_c3()
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    a.line_cross(np.zeros(n), 0.0, clamped)
    sym = M.SymmetricUtduSolver().analyze(a)
    st = M.plan_stats(sym)
    inf = sym.info()
    ref = inf["estimated_factor_bytes"] + inf["peak_contribution_bytes"]
    assert 8 * st["arena_elems"] <= 1.5 * ref, (8 * st["arena_elems"], ref)
