"""The committed golden fixtures (tests/golden/, written by make_golden.py
from the oracle) against the oracle as it is now and against closed forms
that do not depend on the oracle: the post-Dirichlet pattern size of the
square mesh, the dense block formula of ImposeDirichlet, the solve residual,
and the symmetry of the Dirichlet-eliminated matrix."""
import os

import numpy as np
import scipy.sparse as sp

from admfem import meshgen
from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def nnz_square(m):
    """Entries of the post-BC matrix of the m x m square split into 2m^2
    triangles: interior diagonal, interior-interior edge pairs, boundary
    diagonals."""
    return (m - 1) ** 2 + 2 * (2 * (m - 1) * (m - 2) + (m - 2) ** 2) + 4 * m


def readme(m):
    c, e = meshgen.rectangle(m, m, 1.0 / m)
    om = O.OracleMesh(c, e)
    return c, e, om, np.sin(om.gx) * (1 + om.gy ** 2) + 1.0, 1e5 * (om.gx + om.gy)


def test_fixture_patterns_have_the_closed_form_size():
    for name in ("c1_m8.npz", "c1_m64.npz"):
        d = load(name)
        m = int(d["m"])
        assert d["indptr"][-1] == d["indices"].size == nnz_square(m)
    assert nnz_square(64) == 27537  # the size table of the 64x64 case


def test_c1_m8_fixture_reproduced_by_oracle():
    d = load("c1_m8.npz")
    c, e, om, kappa, f = readme(8)
    np.testing.assert_array_equal(d["coords"], c)
    np.testing.assert_array_equal(d["elems_oriented"], om.elems)
    np.testing.assert_allclose(d["kappa"], kappa, rtol=1e-15)
    Ki, Kv = om.laplace_coo(kappa)
    np.testing.assert_array_equal(d["K_indices"], Ki)
    np.testing.assert_allclose(d["K_values"], Kv, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(d["F"], om.source(f), rtol=1e-14)
    oi, ov, orhs = O.dirichlet(Ki, Kv, d["bd"], om.source(f), np.zeros(d["bd"].size))
    np.testing.assert_array_equal(d["D_indices"], oi)
    np.testing.assert_allclose(d["D_values"], ov, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(d["D_rhs"], orhs, rtol=1e-14, atol=1e-10)


def test_c1_m8_fixture_dirichlet_block_form_and_solve():
    d = load("c1_m8.npz")
    N = d["coords"].shape[0]
    K = sp.coo_matrix((d["K_values"], (d["K_indices"][:, 0], d["K_indices"][:, 1])),
                      shape=(N, N)).toarray()
    D = sp.coo_matrix((d["D_values"], (d["D_indices"][:, 0], d["D_indices"][:, 1])),
                      shape=(N, N)).toarray()
    b = d["bd"] - 1
    keep = np.ones(N, bool)
    keep[b] = False
    want = np.zeros_like(K)
    want[np.ix_(keep, keep)] = K[np.ix_(keep, keep)]
    want[b, b] = 1.0
    np.testing.assert_allclose(D, want, rtol=0, atol=1e-13 * np.abs(K).max())
    np.testing.assert_allclose(D, D.T, atol=1e-13 * np.abs(K).max())
    r = D @ d["u"] - d["D_rhs"]
    assert np.abs(r).max() < 1e-9 * np.abs(d["D_rhs"]).max()
    # adjoint: D^T lam = dL/du with L = sum_obs u^2
    g = np.zeros(N)
    g[d["obs"]] = 2 * d["u"][d["obs"]]
    assert np.abs(D.T @ d["lam"] - g).max() < 1e-9 * np.abs(g).max()


def test_c1_m64_and_heat_fixtures_reproduced_by_oracle():
    d = load("c1_m64.npz")
    c, e, om, kappa, f = readme(64)
    bd = om.bcnode()
    np.testing.assert_array_equal(np.sort(d["bd"]), np.sort(bd))

    def dl(u):
        g = np.zeros_like(u)
        g[d["obs"]] = 2 * u[d["obs"]]
        return g
    r = O.forward_adjoint(om, kappa, f, d["bd"], np.zeros(d["bd"].size), dl)
    np.testing.assert_allclose(r["u"], d["u"], rtol=1e-11, atol=1e-11 * np.abs(d["u"]).max())
    g0 = d["grad_kappa"]
    np.testing.assert_allclose(r["grad_kappa"], g0, rtol=1e-9, atol=1e-11 * np.abs(g0).max())
    A = r["A_bc"].tocsr()
    A.sort_indices()
    np.testing.assert_array_equal(A.indptr, d["indptr"])
    np.testing.assert_array_equal(A.indices, d["indices"])
    h = load("heat_m6.npz")
    om6 = O.OracleMesh(h["coords"], h["elems"])

    def dl6(u):
        g = np.zeros_like(u)
        g[h["obs"]] = 2 * u[h["obs"]]
        return g
    r6 = O.forward_adjoint(om6, h["kappa"], h["f"], h["bd"], np.zeros(h["bd"].size), dl6,
                           dt=float(h["dt"]), u_prev=h["u_prev"])
    np.testing.assert_allclose(r6["u"], h["u"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r6["grad_kappa"], h["grad_kappa"], rtol=1e-10, atol=1e-14)
