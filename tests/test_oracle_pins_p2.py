"""Pins the oracle's P2 restatement (degree 2: MFEM's order-2 H1 triangle,
deps/MFEM/Common.cpp:20-142 with degree 2) and its 6-point order-4 rule.

No P2 matrix from the reference is available here (the FEniCS files its
ftest.jl scripts read are not in the tree), so the pins are the reference's
own identity FemSource1/ftest.jl checks for degrees 1 and 2 (the load vector
equals the mass matrix times ones) and mathematics a plausible restatement
mistake breaks: polynomial exactness of the rule, interpolation of quadratics
by the nodal basis (which also fixes which edge each edge dof sits on),
exact stiffness / mass integrals of quadratics, third-order convergence of
the manufactured solution, and the adjoint identities of the backward ops.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spl

from admfem import meshgen
from oracle import oracle as O


def rule(order):
    xi, eta, w = np.zeros(8), np.zeros(8), np.zeros(8)
    n = O.lib().orc_tri_rule(order, xi, eta, w)
    return xi[:n], eta[:n], w[:n]


def p2mesh(m, seed=None, order=-1):
    if seed is None:
        c, e = meshgen.rectangle(m, m, 1.0 / m)
    else:
        c, e = meshgen.jittered_square(m, seed=seed)
    return O.OracleMesh(c, e, order=order, degree=2)


def dof_coords(M):
    """Vertex coordinates, then edge midpoints, in dof order."""
    ed = M.edges() - 1
    mid = 0.5 * (M.coords[ed[:, 0]] + M.coords[ed[:, 1]])
    return np.concatenate([M.coords, mid])


QUAD = [
    (lambda x, y: 1 + 2 * x - y + 3 * x * x - x * y + 0.5 * y * y,
     lambda x, y: 2 + 6 * x - y, lambda x, y: -1 - x + y),
    (lambda x, y: x * y, lambda x, y: y, lambda x, y: x),
]


# ---------------------------------------------------------------- rule
@pytest.mark.parametrize("order", [3, 4])
def test_six_point_rule_exact_to_degree_four(order):
    """The order-4 rule integrates xi^a eta^b, a+b <= 4, exactly over the
    reference triangle (a! b! / (a+b+2)!), and not degree 6 (it is the
    6-point degree-4 rule, not a higher one)."""
    xi, eta, w = rule(order)
    assert xi.size == 6
    for a in range(5):
        for b in range(5 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs((w * xi ** a * eta ** b).sum() - exact) < 1e-16
    assert abs((w * xi ** 6).sum() - 720 / math.factorial(8)) > 1e-6
    # two symmetric orbits (a,a),(a,1-2a),(1-2a,a), as the 3-point rule lists
    for o in range(2):
        a = xi[3 * o]
        np.testing.assert_allclose([xi[3 * o:3 * o + 3], eta[3 * o:3 * o + 3]],
                                   [[a, a, 1 - 2 * a], [a, 1 - 2 * a, a]], rtol=0, atol=1e-16)


# ---------------------------------------------------------------- basis
def test_p2_dofs_counts_and_bcnode():
    M = p2mesh(6)
    assert M.nd == 6 and M.nq == 6 and M.ngauss == 6 * M.nelem
    assert M.ndof == M.nnode + M.nedge == 13 * 13
    np.testing.assert_array_equal(M.dofs[:, :3], M.elems)
    assert (M.dofs[:, 3:] >= M.nnode).all()
    bd = M.bcnode()
    assert bd.size == 4 * 6 + 4 * 6           # perimeter vertices + perimeter edges
    x = dof_coords(M)[bd - 1]
    on = (np.abs(x[:, 0]) < 1e-15) | (np.abs(x[:, 0] - 1) < 1e-15) | \
         (np.abs(x[:, 1]) < 1e-15) | (np.abs(x[:, 1] - 1) < 1e-15)
    assert on.all()


@pytest.mark.parametrize("seed", [None, 5])
def test_p2_basis_interpolates_quadratics_and_their_gradients(seed):
    M = p2mesh(5, seed=seed)
    X = dof_coords(M)
    h, hx, hy = (a.reshape(-1, 6) for a in (M.h, M.hx, M.hy))
    np.testing.assert_allclose(h.sum(1), 1.0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(hx.sum(1), 0.0, rtol=0, atol=1e-11)
    np.testing.assert_allclose(hy.sum(1), 0.0, rtol=0, atol=1e-11)
    d = np.repeat(M.dofs, M.nq, axis=0)       # dofs of each Gauss point's element
    for q, qx, qy in QUAD:
        u = q(X[:, 0], X[:, 1])
        np.testing.assert_allclose((u[d] * h).sum(1), q(M.gx, M.gy), rtol=0, atol=1e-13)
        np.testing.assert_allclose((u[d] * hx).sum(1), qx(M.gx, M.gy), rtol=0, atol=1e-11)
        np.testing.assert_allclose((u[d] * hy).sum(1), qy(M.gx, M.gy), rtol=0, atol=1e-11)
        # the op forms: DofToGaussPoints and FemGrad
        np.testing.assert_allclose(M.dof_to_gauss(u), q(M.gx, M.gy), rtol=0, atol=1e-13)
        g = M.fem_grad(u)
        np.testing.assert_allclose(g[:, 0], qx(M.gx, M.gy), rtol=0, atol=1e-11)
        np.testing.assert_allclose(g[:, 1], qy(M.gx, M.gy), rtol=0, atol=1e-11)
    # FemToGaussPoints stays linear on the vertices (hs), whatever the degree
    u = 1 + 2 * M.coords[:, 0] - 3 * M.coords[:, 1]
    np.testing.assert_allclose(M.fem_to_gauss(u), 1 + 2 * M.gx - 3 * M.gy, rtol=0, atol=1e-14)


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("seed", [None, 8])
def test_p2_stiffness_and_mass_integrate_quadratics_exactly(seed):
    """u'Kv = int grad u . grad v and u'Mv = int u v for quadratic u, v
    (degree 2 and 4 integrands: exact under the order-4 rule)."""
    M = p2mesh(4, seed=seed)
    X = dof_coords(M)
    x, y = X[:, 0], X[:, 1]
    K = O.csr_matrix(*M.laplace_coo(np.ones(M.ngauss)), M.ndof)
    Mm = O.csr_matrix(*M.mass_coo(np.ones(M.ngauss)), M.ndof)
    u, v = x * x, x * y
    assert abs(u @ (K @ v) - 0.5) < 1e-12        # int 2x*y over the unit square
    r = x * x + y * y
    assert abs(r @ (K @ r) - 8.0 / 3.0) < 1e-12   # int 4x^2 + 4y^2
    assert abs(np.ones(M.ndof) @ (Mm @ np.ones(M.ndof)) - 1.0) < 1e-13
    assert abs(u @ (Mm @ (y * y)) - 1.0 / 9.0) < 1e-13
    assert abs(v @ (Mm @ v) - 1.0 / 9.0) < 1e-13
    np.testing.assert_allclose(K @ np.ones(M.ndof), 0.0, atol=1e-12)
    assert abs(K - K.T).max() < 1e-13 and abs(Mm - Mm.T).max() < 1e-15


def test_p2_source_equals_mass_times_ones():
    """deps/MFEM/FemSource1/ftest.jl (degrees 1 and 2): the load vector of c
    equals compute_fem_mass_matrix1(c) * ones(ndof)."""
    M = p2mesh(8)
    c = np.random.default_rng(0).random(M.ngauss)
    F = M.source(c)
    Mc = O.csr_matrix(*M.mass_coo(c), M.ndof)
    np.testing.assert_allclose(F, Mc @ np.ones(M.ndof), rtol=1e-13, atol=1e-16)


def test_p2_laplace_term_equals_stiffness_times_u():
    M = p2mesh(5, seed=2)
    rng = np.random.default_rng(1)
    nu, u = rng.random(M.ngauss) + 0.5, rng.standard_normal(M.ndof)
    K = O.csr_matrix(*M.laplace_coo(nu), M.ndof)
    np.testing.assert_allclose(M.laplace_term(u, nu), K @ u, rtol=1e-12, atol=1e-13)


def test_p2_backward_ops_are_adjoints():
    M = p2mesh(4, seed=3)
    rng = np.random.default_rng(2)
    dk, gvv = rng.standard_normal(M.ngauss), rng.standard_normal(36 * M.ngauss)
    for fwd, bwd in ((M.laplace_coo, M.laplace_backward), (M.mass_coo, M.mass_backward)):
        assert abs(gvv @ fwd(dk)[1] - bwd(gvv) @ dk) < 1e-11 * np.abs(gvv).sum()
    gr = rng.standard_normal(M.ndof)
    assert abs(gr @ M.source(dk) - M.source_backward(gr) @ dk) < 1e-12
    u, go = rng.standard_normal(M.ndof), rng.standard_normal(M.ngauss)
    assert abs(go @ M.dof_to_gauss(u) - M.dof_to_gauss_backward(go) @ u) < 1e-12
    g2 = rng.standard_normal((M.ngauss, 2))
    assert abs((g2 * M.fem_grad(u)).sum() - M.fem_grad_backward(g2) @ u) < 1e-10
    nu, gout = rng.random(M.ngauss) + 0.5, rng.standard_normal(M.ndof)
    gn, gu = M.laplace_term_backward(u, nu, gout)
    assert abs(gout @ M.laplace_term(u, nu) - gn @ nu) < 1e-10
    assert abs(gout @ M.laplace_term(u, nu) - gu @ u) < 1e-10


def test_p2_manufactured_solution_converges_third_order():
    """-lap u = 2 pi^2 sin(pi x) sin(pi y), u = 0 on the boundary
    (docs/src/mfem_tutorial.md:29-60 with degree 2): the vertex error falls
    by ~8x per halving of h (4x would mean a first-order basis slipped in)."""
    errs = []
    for m in (4, 8, 16):
        M = p2mesh(m)
        f = 2 * np.pi ** 2 * np.sin(np.pi * M.gx) * np.sin(np.pi * M.gy)
        bd = M.bcnode()
        oi, ov, orhs = O.dirichlet(*M.laplace_coo(np.ones(M.ngauss)), bd, M.source(f),
                                   np.zeros(bd.size))
        u = spl.spsolve(O.csr_matrix(oi, ov, M.ndof).tocsc(), orhs)
        ex = np.sin(np.pi * M.coords[:, 0]) * np.sin(np.pi * M.coords[:, 1])
        errs.append(np.abs(u[:M.nnode] - ex).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert (rates > 2.7).all(), (errs, rates)
