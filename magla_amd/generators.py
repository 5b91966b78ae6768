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


def hex8_stiffness(h, nu):
    """24x24 stiffness of a cube trilinear hexahedron of edge h for E = 1 (2x2x2 Gauss).
    Local node a = (ax, ay, az) in {0,1}^3 ordered a = ax + 2 ay + 4 az; dof 3a + d."""
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1.0 / (2 * (1 + nu))
    Cm = np.zeros((6, 6))
    Cm[:3, :3] = lam
    Cm[np.arange(3), np.arange(3)] += 2 * mu
    Cm[np.arange(3, 6), np.arange(3, 6)] = mu
    corners = np.array([(a & 1, (a >> 1) & 1, (a >> 2) & 1) for a in range(8)], dtype=float) * 2 - 1
    g = 1.0 / np.sqrt(3.0)
    K = np.zeros((24, 24))
    for gx in (-g, g):
        for gy in (-g, g):
            for gz in (-g, g):
                # dN/dxi, scaled by 2/h (reference cube [-1,1]^3 -> [0,h]^3)
                dN = np.zeros((8, 3))
                for a in range(8):
                    sx, sy, sz = corners[a]
                    dN[a, 0] = sx * (1 + sy * gy) * (1 + sz * gz) / 8
                    dN[a, 1] = sy * (1 + sx * gx) * (1 + sz * gz) / 8
                    dN[a, 2] = sz * (1 + sx * gx) * (1 + sy * gy) / 8
                dN *= 2.0 / h
                B = np.zeros((6, 24))
                for a in range(8):
                    B[0, 3 * a] = dN[a, 0]
                    B[1, 3 * a + 1] = dN[a, 1]
                    B[2, 3 * a + 2] = dN[a, 2]
                    B[3, 3 * a] = dN[a, 1]
                    B[3, 3 * a + 1] = dN[a, 0]
                    B[4, 3 * a + 1] = dN[a, 2]
                    B[4, 3 * a + 2] = dN[a, 1]
                    B[5, 3 * a] = dN[a, 2]
                    B[5, 3 * a + 2] = dN[a, 0]
                K += B.T @ Cm @ B * (h / 2) ** 3
    return 0.5 * (K + K.T)


def elasticity_3d(m, seed, nu=0.3, sigma=0.5):
    """Config C3 matrix: m^3 trilinear hex elements on the unit cube, (m+1)^3 nodes x 3 dof
    (node (i,j,k) -> (k(m+1)+j)(m+1)+i, dof 3*node+d), Young's modulus lognormal(0, sigma) per
    element, Poisson ratio nu.  Returns COO of the upper triangle (each element contributes its
    upper entries once; duplicates are summed by the CSR builder) and the element moduli."""
    g = rng(seed)
    E = g.lognormal(0.0, sigma, m ** 3)
    K0 = hex8_stiffness(1.0 / m, nu)
    p = m + 1
    ei, ej, ek = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    ei, ej, ek = ek.ravel(), ej.ravel(), ei.ravel()  # element e = (k m + j) m + i
    base = (ek * p + ej) * p + ei
    local = np.array([(a & 1) + ((a >> 1) & 1) * p + ((a >> 2) & 1) * p * p for a in range(8)])
    nodes = base[:, None] + local[None, :]  # (ne, 8)
    dofs = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)  # (ne, 24)
    iu, ju = np.triu_indices(24)
    r = dofs[:, iu]
    c = dofs[:, ju]
    v = E[:, None] * K0[iu, ju][None, :]
    n = 3 * p ** 3
    return n, r.ravel().astype(np.int32), c.ravel().astype(np.int32), v.ravel()


def clamped_dofs_x0(m):
    """DOF of the nodes on the face x = 0 (i = 0)."""
    p = m + 1
    kk, jj = np.meshgrid(np.arange(p), np.arange(p), indexing="ij")
    nodes = ((kk * p + jj) * p).ravel()
    return (3 * nodes[:, None] + np.arange(3)[None, :]).ravel().astype(np.int32)


def config_c3(m=40, k=8):
    """3D elasticity, m^3 hex elements, x=0 face clamped through LineCross, k nodal-load RHS N(0,1)
    on the free DOF.  Returns (SymmetricCSRMatrix-ready arrays via the caller's builder)."""
    n, r, c, v = elasticity_3d(m, SEED_BASE + 3)
    clamped = clamped_dofs_x0(m)
    g = rng(SEED_BASE + 3 + 1000)
    b = g.standard_normal((k, n))
    b[:, clamped] = 0.0
    return n, r, c, v, clamped, np.asfortranarray(b.T)


def gaussian_random_field(m, seed, length=0.1, variance=2.0, embed=2):
    """Stationary Gaussian random field on the m^3 interior nodes of the unit cube (spacing
    h = 1/(m+1), node (i,j,k) at ((i+1)h, (j+1)h, (k+1)h)) with covariance
    variance * exp(-|x-y|^2 / (2 length^2)), by circulant embedding on an (embed*m)^3 periodic
    grid of the same spacing: g = IFFT(sqrt(lambda) * FFT(w)) with w white N(0,1) and lambda the
    FFT of the periodic covariance row (negative eigenvalues clipped to 0); the m^3 window is kept.
    Returned as (m, m, m) indexed [z, y, x]."""
    h = 1.0 / (m + 1)
    M = embed * m
    d = np.minimum(np.arange(M), M - np.arange(M)) * h
    r2 = d[:, None, None] ** 2 + d[None, :, None] ** 2 + d[None, None, :] ** 2
    row = variance * np.exp(-r2 / (2.0 * length * length))
    lam = np.maximum(np.fft.fftn(row).real, 0.0)
    w = rng(seed).standard_normal((M, M, M))
    g = np.fft.ifftn(np.sqrt(lam) * np.fft.fftn(w)).real
    return g[:m, :m, :m].copy()


def diffusion_3d_grf(m, seed, length=0.1, variance=2.0):
    """Config C5 matrix: 3D 7-point diffusion on m^3 interior nodes with node coefficients
    kappa = exp(GRF); interior face conductivity = harmonic mean of its two node kappas; a
    Dirichlet boundary face takes the kappa of its node."""
    kap = np.exp(gaussian_random_field(m, seed, length, variance))
    hm = lambda a, b: 2.0 * a * b / (a + b)
    kx = hm(kap[:, :, :-1], kap[:, :, 1:])
    ky = hm(kap[:, :-1, :], kap[:, 1:, :])
    kz = hm(kap[:-1, :, :], kap[1:, :, :])
    kb = np.zeros((m, m, m))
    for axis in range(3):
        for side in (0, -1):
            sl = [slice(None)] * 3
            sl[axis] = side
            kb[tuple(sl)] += kap[tuple(sl)]
    return diffusion_3d_faces(m, kx, ky, kz, kb)


def config_c5(m=64):
    """Learned-preconditioner config: 3D 7-point diffusion, m^3 (64^3: n = 262,144), kappa = exp(GRF,
    length 0.1, variance 2); test right-hand side uniform[-1,1]."""
    n, r, c, v = diffusion_3d_grf(m, SEED_BASE + 5)
    return n, r, c, v, uniform_rhs(n, 1, SEED_BASE + 5 + 1000)


def config_c4(count=1024, seed=SEED_BASE + 4, side_lo=16, side_hi=128, max_rhs=32):
    """Batched config: `count` independent SPD systems, each a 2D 5-point or 9-point Laplacian (p = 1/2)
    on an s x s grid, s ~ U{side_lo..side_hi}, plus a diagonal shift ~ U[0, 0.1]; 1..max_rhs
    right-hand sides each, uniform[-1,1].  Returns a list of (n, rows, cols, vals, b (n, k))."""
    g = rng(seed)
    out = []
    for t in range(count):
        s = int(g.integers(side_lo, side_hi + 1))
        nine = bool(g.random() < 0.5)
        shift = float(0.1 * g.random())
        k = int(g.integers(1, max_rhs + 1))
        n, r, c, v = laplacian_2d(s, shift, nine)
        b = 2.0 * g.random((k, n)) - 1.0
        out.append((n, r, c, v, np.asfortranarray(b.T)))
    return out


def convection_diffusion_2d(m, peclet=2.0, seed=SEED_BASE + 6):
    """Nonsymmetric test system for BiCGSTAB: upwind 5-point convection-diffusion on an m x m
    grid in natural order (i = y*m + x), Dirichlet neighbours dropped; per node the upwind
    weights wx = peclet*(1 + U/2), wy = peclet*(1/2 + U) with U ~ U[0,1); diagonal 4 + wx + wy,
    west -1 - wx, south -1 - wy, east and north -1 (an M-matrix, diagonally dominant).
    Returns (n, rows, cols, vals) with duplicate-free COO entries."""
    g = rng(seed)
    n = m * m
    wx = peclet * (1.0 + 0.5 * g.random(n))
    wy = peclet * (0.5 + g.random(n))
    idx = np.arange(n).reshape(m, m)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [4.0 + wx + wy]
    for dy, dx, w in ((0, -1, -1.0 - wx), (0, 1, -np.ones(n)), (-1, 0, -1.0 - wy), (1, 0, -np.ones(n))):
        src = idx[max(0, -dy):m - max(0, dy), max(0, -dx):m - max(0, dx)].ravel()
        dst = idx[max(0, dy):m - max(0, -dy), max(0, dx):m - max(0, -dx)].ravel()
        rows.append(src)
        cols.append(dst)
        vals.append(w[src])
    return n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
