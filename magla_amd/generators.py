"""Seeded synthetic systems for the benchmark configurations.

Each generator returns (n, rows, cols, vals) coordinate entries; the matrix is
assembled by SymmetricCSRMatrix.from_coo (the reference's COO semantics:
normalise, sum duplicates, drop |v| <= 1e-15, sort).  Random numbers come from
numpy's Philox generator seeded with `seed`; uniform[-1,1] is 2u-1 with a
53-bit u, as in UtduSolverBenchmark.cs:209 (NextDouble()*2-1).
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 20260812  # UtduSolverBenchmark.cs:206


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def uniform_rhs(n, k, seed):
    """k right-hand sides, entries i.i.d. uniform[-1,1] (column-major (n, k) or (n,) for k=1)."""
    g = rng(seed)
    b = 2.0 * g.random((k, n)) - 1.0
    return b[0].copy() if k == 1 else np.asfortranarray(b.T)


def laplacian_2d(m, shift=0.0, nine_point=False):
    """2D 5-point (diag 4, -1 to in-grid neighbours) or 9-point (diag 8, -1 to all 8 neighbours)
    Laplacian on an m x m grid of interior nodes, i = y*m + x, Dirichlet neighbours dropped."""
    n = m * m
    idx = np.arange(n).reshape(m, m)
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    vals = [np.full(n, (8.0 if nine_point else 4.0) + shift)]
    offs = [(0, 1), (1, 0)] + ([(1, 1), (1, -1)] if nine_point else [])
    for dy, dx in offs:
        y0, y1 = 0, m - dy
        x0, x1 = max(0, -dx), m - max(0, dx)
        a = idx[y0:y1, x0:x1].ravel()
        b = idx[y0 + dy:y1 + dy, x0 + dx:x1 + dx].ravel()
        rows.append(a)
        cols.append(b)
        vals.append(np.full(a.size, -1.0))
    return n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def diffusion_3d_faces(m, kx, ky, kz, kb):
    """3D 7-point finite-volume diffusion on an m^3 grid from face conductivities:
    kx: (m, m, m-1) faces between (z,y,x) and (z,y,x+1), likewise ky (m, m-1, m), kz (m-1, m, m);
    kb: (m, m, m) sum of the Dirichlet boundary-face conductivities of each node.
    A_ii = sum of its 6 face conductivities, A_ij = -k_face.  Node index i = (z*m + y)*m + x."""
    n = m ** 3
    idx = np.arange(n).reshape(m, m, m)
    diag = kb.copy()
    diag[:, :, :-1] += kx
    diag[:, :, 1:] += kx
    diag[:, :-1, :] += ky
    diag[:, 1:, :] += ky
    diag[:-1, :, :] += kz
    diag[1:, :, :] += kz
    rows = [idx.ravel(), idx[:, :, :-1].ravel(), idx[:, :-1, :].ravel(), idx[:-1, :, :].ravel()]
    cols = [idx.ravel(), idx[:, :, 1:].ravel(), idx[:, 1:, :].ravel(), idx[1:, :, :].ravel()]
    vals = [diag.ravel(), -kx.ravel(), -ky.ravel(), -kz.ravel()]
    return n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def diffusion_3d_lognormal(m, seed, sigma=1.0):
    """Config C2: face conductivities i.i.d. lognormal(0, sigma) on interior and Dirichlet boundary faces."""
    g = rng(seed)
    kx = g.lognormal(0.0, sigma, (m, m, m - 1))
    ky = g.lognormal(0.0, sigma, (m, m - 1, m))
    kz = g.lognormal(0.0, sigma, (m - 1, m, m))
    kb = np.zeros((m, m, m))
    # 6 boundary face families, each m^2 faces
    for axis in range(3):
        for side in (0, -1):
            sl = [slice(None)] * 3
            sl[axis] = side
            kb[tuple(sl)] += g.lognormal(0.0, sigma, (m, m))
    return diffusion_3d_faces(m, kx, ky, kz, kb)


def config_c1():
    """2D 5-point Laplacian, 64x64 (n = 4096), 1 RHS."""
    n, r, c, v = laplacian_2d(64)
    return n, r, c, v, uniform_rhs(n, 1, SEED_BASE + 1)


def config_c2(m=48):
    """3D 7-point variable-coefficient diffusion, 48^3 (n = 110,592), 4 RHS."""
    n, r, c, v = diffusion_3d_lognormal(m, SEED_BASE + 2)
    return n, r, c, v, uniform_rhs(n, 4, SEED_BASE + 2 + 1000)
