"""Shared test helpers: seeded smooth MAC fields and comparison metrics."""
from __future__ import annotations

import numpy as np


def face_coords(res, axis, dx):
    """Coordinates of the faces of `axis`: (d, *face_shape), face a at I_a*dx,
    other axes at (I+0.5)*dx (3D/taichi_utils.py:33-46)."""
    d = len(res)
    shp = list(res)
    shp[axis] += 1
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shp], indexing="ij")
    out = []
    for c in range(d):
        g = grids[c] if c == axis else grids[c] + 0.5
        out.append(g * dx)
    return np.stack(out)


def smooth_mac_field(res, seed=0, modes=4, amp=1.0):
    """A seeded smooth MAC velocity (list of float32 face arrays)."""
    rng = np.random.default_rng(seed)
    d = len(res)
    dx = 1.0 / res[1]
    u = []
    for a in range(d):
        X = face_coords(res, a, dx)
        f = np.zeros(X.shape[1:], np.float64)
        for _ in range(modes):
            k = rng.uniform(0.5, 3.0, size=d) * 2 * np.pi
            ph = rng.uniform(0, 2 * np.pi)
            c = rng.normal() * amp
            f += c * np.sin(np.tensordot(k, X, axes=1) + ph)
        u.append(f.astype(np.float32))
    return u


def affine_mac_field(res, a0, B):
    """u_a(x) = a0[a] + B[a] . x sampled on the faces of axis a."""
    d = len(res)
    dx = 1.0 / res[1]
    u = []
    for a in range(d):
        X = face_coords(res, a, dx)
        f = a0[a] + np.tensordot(np.asarray(B[a], np.float64), X, axes=1)
        u.append(f.astype(np.float32))
    return u


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def neumann_laplacian(shape):
    """Graph Laplacian of a box of cells with all-Neumann walls (scipy sparse),
    assembled from 1D difference operators: sum_a D_a^T D_a."""
    import scipy.sparse as sp

    d = len(shape)
    A = None
    for a in range(d):
        n = shape[a]
        D = sp.diags([-np.ones(n - 1), np.ones(n - 1)], [0, 1], shape=(n - 1, n))
        L1 = (D.T @ D).tocsr()
        mats = [sp.identity(shape[b], format="csr") if b != a else L1 for b in range(d)]
        K = mats[0]
        for m in mats[1:]:
            K = sp.kron(K, m, format="csr")
        A = K if A is None else A + K
    return A.tocsr()


def polyP(Z):
    """RK4 stability polynomial of a matrix: I + Z + Z^2/2 + Z^3/6 + Z^4/24."""
    I = np.eye(Z.shape[0])
    Z2 = Z @ Z
    Z3 = Z2 @ Z
    return I + Z + Z2 / 2 + Z3 / 6 + Z2 @ Z2 / 24


def fluid_laplacian(solid):
    """Graph Laplacian of the fluid cells of a box with all-Neumann walls and
    solid cells (scipy sparse, on all cells; solid rows and columns are empty):
    every pair of face-adjacent fluid cells is an edge."""
    import scipy.sparse as sp

    shape = solid.shape
    n = solid.size
    idx = np.arange(n).reshape(shape)
    rows, cols = [], []
    for a in range(len(shape)):
        lo = [slice(None)] * len(shape)
        hi = [slice(None)] * len(shape)
        lo[a] = slice(0, -1)
        hi[a] = slice(1, None)
        ok = ~solid[tuple(lo)] & ~solid[tuple(hi)]
        rows.append(idx[tuple(lo)][ok])
        cols.append(idx[tuple(hi)][ok])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    W = sp.coo_matrix((np.ones(r.size), (r, c)), shape=(n, n))
    W = (W + W.T).tocsr()
    deg = np.asarray(W.sum(axis=1)).ravel()
    return (sp.diags(deg) - W).tocsr()


def box_solid(shape, lo=0.4, hi=0.6):
    """Cells whose centre lies in the box [lo, hi]^d of the unit-height domain
    (dx = 1/shape[1])."""
    dx = 1.0 / shape[1]
    axes = [(np.arange(n) + 0.5) * dx for n in shape]
    grids = np.meshgrid(*axes, indexing="ij")
    m = np.ones(shape, bool)
    for g in grids:
        m &= (g >= lo) & (g <= hi)
    return m
