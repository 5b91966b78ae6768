import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmagla.so on the device)")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "cg_comparison.json")) as f:
        return json.load(f)


def dense_coo(a):
    r, c = np.nonzero(np.triu(np.asarray(a)))
    return r, c, np.asarray(a)[r, c]


# Matrix builders restating the reference tests' fixtures
# (Tests/UCHOL/SymmetricUtduSolverTests.cs:32-143).
def grid_matrix(nx, ny, nz, dof):
    """BuildGridMatrix (:61-102): dof x dof coupling blocks, dominant diagonal."""
    nodes = nx * ny * nz
    rows, cols, vals = [], [], []
    nb = np.zeros(nodes, dtype=np.int64)

    def index(x, y, z):
        return (z * ny + y) * nx + x

    def couple(a, b):
        for d in range(dof):
            for e in range(dof):
                rows.append(a * dof + d)
                cols.append(b * dof + e)
                vals.append(-1.0 if d == e else -0.0625)
        nb[a] += 1
        nb[b] += 1

    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                node = index(x, y, z)
                if x + 1 < nx:
                    couple(node, index(x + 1, y, z))
                if y + 1 < ny:
                    couple(node, index(x, y + 1, z))
                if z + 1 < nz:
                    couple(node, index(x, y, z + 1))
    for node in range(nodes):
        for d in range(dof):
            for e in range(d + 1, dof):
                rows.append(node * dof + d)
                cols.append(node * dof + e)
                vals.append(-0.25)
            rows.append(node * dof + d)
            cols.append(node * dof + d)
            vals.append(nb[node] * dof + 0.25 * (dof - 1) + 1.0)
    return nodes * dof, np.array(rows), np.array(cols), np.array(vals)


def random_matrix(n, density, seed):
    """BuildRandomMatrix (:111-133) with numpy's generator: random upper pattern, values U[-1,1],
    diagonal = sum |offdiag| + 1 (strictly dominant, SPD)."""
    g = np.random.default_rng(seed)
    mask = np.triu(g.random((n, n)) < density, 1)
    r, c = np.nonzero(mask)
    v = g.uniform(-1.0, 1.0, len(r))
    off = np.zeros(n)
    np.add.at(off, r, np.abs(v))
    np.add.at(off, c, np.abs(v))
    return n, np.concatenate([r, np.arange(n)]), np.concatenate([c, np.arange(n)]), np.concatenate([v, off + 1.0])


def tridiagonal(n, d, o):
    r = np.concatenate([np.arange(n), np.arange(n - 1)])
    c = np.concatenate([np.arange(n), np.arange(1, n)])
    v = np.concatenate([np.full(n, d), np.full(n - 1, o)])
    return n, r, c, v


def vector(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)
