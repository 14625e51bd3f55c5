"""Mesh and coefficient generators (host, numpy).

* ``rectangle`` restates the reference's structured generator
  ``Mesh(m, n, h; version)`` (src/MFEM/MFEM.jl:134-171).
* ``jittered_square`` builds the synthetic workloads: the unit square split
  into 2*m*m counter-clockwise triangles, interior nodes jittered by
  U[-j*h, j*h] in each coordinate, node ids randomly permuted.
* ``grf`` is a Gaussian random field made of random Fourier features
  (Rahimi & Recht 2007): g(x) = sigma*sqrt(2/K) * sum_k cos(w_k.x + phi_k),
  w_k ~ N(0, l^-2 I), phi_k ~ U[0, 2 pi].
"""
import numpy as np


def rectangle(m, n, h, version=1, seed=None):
    """Nodes ((m+1)(n+1) x 2, x fastest) and 0-based triangles (2mn x 3) of
    src/MFEM/MFEM.jl:134-171.  Version 1 splits each quad into
    [BL,BR,TL],[BR,TL,TR]; version 2 into [BL,TR,TL],[BL,BR,TR]; version 3
    picks one of the two per quad at random (the reference calls Julia's
    rand(); here a numpy Generator seeded by ``seed``)."""
    j, i = np.meshgrid(np.arange(m + 1), np.arange(n + 1))
    coords = np.stack([j.ravel() * h, i.ravel() * h], axis=1).astype(np.float64)
    ii, jj = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    bl = (ii * (m + 1) + jj).ravel()
    br, tl, tr = bl + 1, bl + (m + 1), bl + (m + 2)
    v1a = np.stack([bl, br, tl], 1)
    v1b = np.stack([br, tl, tr], 1)
    v2a = np.stack([bl, tr, tl], 1)
    v2b = np.stack([bl, br, tr], 1)
    if version == 1:
        a, b = v1a, v1b
    elif version == 2:
        a, b = v2a, v2b
    elif version == 3:
        pick = np.random.default_rng(seed).random(bl.size) > 0.5
        a = np.where(pick[:, None], v1a, v2a)
        b = np.where(pick[:, None], v1b, v2b)
    else:
        raise ValueError("version must be 1, 2 or 3")
    elems = np.empty((2 * bl.size, 3), dtype=np.int64)
    elems[0::2] = a
    elems[1::2] = b
    return coords, elems


def jittered_square(m, seed, jitter=0.2, permute=True):
    """Unit square, m x m quads, each split into the CCW triangles
    [BL,BR,TL] and [BR,TR,TL]; interior nodes moved by U[-jitter*h,
    jitter*h]^2 (jitter <= 0.25 keeps every triangle positively oriented);
    node ids permuted at random when ``permute``.  Returns (coords, elems
    0-based)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / m
    j, i = np.meshgrid(np.arange(m + 1), np.arange(m + 1))
    x = (j.ravel() * h).astype(np.float64)
    y = (i.ravel() * h).astype(np.float64)
    interior = (j.ravel() > 0) & (j.ravel() < m) & (i.ravel() > 0) & (i.ravel() < m)
    d = rng.uniform(-jitter * h, jitter * h, size=(x.size, 2))
    x = np.where(interior, x + d[:, 0], x)
    y = np.where(interior, y + d[:, 1], y)
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    bl = (ii * (m + 1) + jj).ravel()
    br, tl, tr = bl + 1, bl + (m + 1), bl + (m + 2)
    elems = np.empty((2 * bl.size, 3), dtype=np.int64)
    elems[0::2] = np.stack([bl, br, tl], 1)
    elems[1::2] = np.stack([br, tr, tl], 1)
    coords = np.stack([x, y], 1)
    if permute:
        p = rng.permutation(x.size)  # old id -> new id
        newc = np.empty_like(coords)
        newc[p] = coords
        coords = newc
        elems = p[elems]
    return coords, elems


def grf(points, seed, nmodes=32, sigma=1.0, ell=0.1):
    """Random-Fourier-feature Gaussian random field at ``points`` (n x 2)."""
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, 1.0 / ell, size=(nmodes, 2))
    phi = rng.uniform(0.0, 2 * np.pi, size=nmodes)
    pts = np.asarray(points, dtype=np.float64)
    g = np.zeros(pts.shape[0])
    for k in range(nmodes):  # chunk-free: nmodes is small
        g += np.cos(pts @ w[k] + phi[k])
    return sigma * np.sqrt(2.0 / nmodes) * g


def gauss_points(coords, elems_oriented, order=2):
    """Gauss points of oriented triangles (nq*nelem x 2, element-major), the
    layout of gauss_nodes (deps/MFEM/Common.cpp:110-111)."""
    if order in (-1, 2):
        a, b = 1.0 / 6.0, 2.0 / 3.0
        q = np.array([[a, a], [a, b], [b, a]])
    elif order in (0, 1):
        q = np.array([[1.0 / 3.0, 1.0 / 3.0]])
    else:
        raise ValueError("order must be 1 or 2")
    c = np.asarray(coords)[np.asarray(elems_oriented)]  # nelem x 3 x 2
    hs = np.stack([1 - q[:, 0] - q[:, 1], q[:, 0], q[:, 1]], 1)  # nq x 3
    g = np.einsum("qk,ekd->eqd", hs, c)
    return g.reshape(-1, 2)


def orient(coords, elems):
    """Swap the first two vertices of clockwise triangles (the orientation fix
    of deps/MFEM/Common.cpp:40)."""
    e = np.array(elems, dtype=np.int64, copy=True)
    c = np.asarray(coords)
    p0, p1, p2 = c[e[:, 0]], c[e[:, 1]], c[e[:, 2]]
    det = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    neg = det < 0
    e[neg, 0], e[neg, 1] = e[neg, 1].copy(), e[neg, 0].copy()
    return e
