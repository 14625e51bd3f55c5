"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check comes from the reference's own tests or from mathematics that a
plausible restatement mistake (a dropped weight factor, a wrong sign, a
transposed index, a wrong Dirichlet ordering) would violate.  The reference
cannot be built here (MFEM, Eigen, TensorFlow, Julia are absent) and ships no
golden vectors, so these are the pins.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from admfem import meshgen
from oracle import oracle as O


def om(m, n=None, h=None, version=1):
    n = m if n is None else n
    h = 1.0 / m if h is None else h
    c, e = meshgen.rectangle(m, n, h, version=version)
    return O.OracleMesh(c, e)


# ---------------------------------------------------------------- mesh
def test_mesh_2x2_known_answers():
    """test/MFEM2.jl:7-27: nodes of Mesh(2,2,0.5), 24 Gauss points, area 1/8."""
    c, e = meshgen.rectangle(2, 2, 0.5)
    want = np.array([[0, 0], [.5, 0], [1, 0], [0, .5], [.5, .5], [1, .5], [0, 1], [.5, 1], [1, 1]])
    np.testing.assert_array_equal(c, want)
    m = O.OracleMesh(c, e)
    assert m.ngauss == 24
    np.testing.assert_allclose(m.area, np.full(8, 0.125), rtol=0, atol=1e-15)


def test_orientation_fix_swaps_first_two_vertices_of_cw_triangles():
    """FinalizeTriMesh(1,0,true) (deps/MFEM/Common.cpp:40): [BR,TL,TR] of the
    version-1 generator is clockwise and becomes [TL,BR,TR]."""
    m = om(2, h=0.5)
    assert m.elems[0].tolist() == [0, 1, 3]      # [1,2,4] CCW, unchanged
    assert m.elems[1].tolist() == [3, 1, 4]      # [2,4,5] CW -> [4,2,5]
    c = m.coords[m.elems]
    det = (c[:, 1, 0] - c[:, 0, 0]) * (c[:, 2, 1] - c[:, 0, 1]) - \
          (c[:, 2, 0] - c[:, 0, 0]) * (c[:, 1, 1] - c[:, 0, 1])
    assert (det > 0).all()


def test_weights_sum_to_area_and_gauss_points_inside():
    c, e = meshgen.jittered_square(9, seed=3)
    m = O.OracleMesh(c, e)
    np.testing.assert_allclose(m.w.reshape(-1, 3).sum(1), m.area, rtol=1e-14)
    # Heron area equals the determinant area
    p = m.coords[m.elems]
    det = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - \
          (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])
    np.testing.assert_allclose(m.area, det / 2, rtol=1e-12)
    # Gauss points are the convex combinations (2/3,1/6,1/6)-style of vertices
    g = np.stack([m.gx, m.gy], 1).reshape(-1, 3, 2)
    np.testing.assert_allclose(g.mean(1), p.mean(1), rtol=1e-13, atol=1e-15)
    # gradients of the barycentric functions sum to zero and reproduce x, y
    hx, hy = m.hx.reshape(-1, 3), m.hy.reshape(-1, 3)
    np.testing.assert_allclose(hx.sum(1), 0, atol=1e-9)
    xs = np.repeat(p[:, :, 0], 3, axis=0)
    np.testing.assert_allclose((hx * xs).sum(1), 1.0, rtol=1e-12)
    np.testing.assert_allclose((hy * xs).sum(1), 0.0, atol=1e-10)


def test_edges_euler_and_bcnode_perimeter():
    m = om(5)
    ed = m.edges()
    assert ed.shape[0] == m.nnode + m.nelem - 1          # V - E + F = 1
    assert len({tuple(r) for r in ed.tolist()}) == ed.shape[0]
    assert (ed[:, 0] < ed[:, 1]).all() and ed.min() == 1
    # MFEM numbering: the first element's edges (0,1),(1,2),(2,0) come first
    e0 = m.elems[0] + 1
    first = [tuple(sorted((e0[a], e0[b]))) for a, b in ((0, 1), (1, 2), (2, 0))]
    assert [tuple(r) for r in ed[:3].tolist()] == first
    bd = m.bcnode()
    c = m.coords[bd - 1]
    on = (np.isclose(c[:, 0], 0) | np.isclose(c[:, 0], 1) | np.isclose(c[:, 1], 0) |
          np.isclose(c[:, 1], 1))
    assert on.all() and bd.size == 4 * 5


# ---------------------------------------------------------------- Laplace
def test_laplace_unit_kappa_is_five_point_stencil():
    """κ≡1 on Mesh(m,m,1/m): interior rows have diagonal 4, axis neighbours
    -1, and the split diagonal's pair as an explicit zero."""
    m = 6
    M = om(m)
    I, V = M.laplace_coo(np.ones(M.ngauss))
    K = O.csr_matrix(I, V, M.nnode)
    n1 = m + 1
    for r in range(1, m):
        for c in range(1, m):
            k = r * n1 + c
            row = K.getrow(k)
            cols = dict(zip(row.indices.tolist(), row.data.tolist()))
            assert set(cols) == {k, k - 1, k + 1, k - n1, k + n1, k + m, k - m}
            assert cols[k] == pytest.approx(4.0, abs=1e-13)
            for nb in (k - 1, k + 1, k - n1, k + n1):
                assert cols[nb] == pytest.approx(-1.0, abs=1e-13)
            assert abs(cols[k + m]) < 1e-13 and abs(cols[k - m]) < 1e-13


def test_laplace_triplet_order_and_element_matrix():
    """Triplet t = 9g+3p+q of FemLaplaceScalar.h:17-23 carries (dof[p],dof[q])
    and w*kappa*grad phi_p.grad phi_q."""
    M = om(2, h=0.5)
    kap = np.arange(1, M.ngauss + 1, dtype=float)
    I, V = M.laplace_coo(kap)
    e, k = 1, 2
    g = 3 * e + k
    v = M.elems[e]
    p = M.coords[v]
    det = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1])
    G = np.array([[p[1, 1] - p[2, 1], p[2, 0] - p[1, 0]],
                  [p[2, 1] - p[0, 1], p[0, 0] - p[2, 0]],
                  [p[0, 1] - p[1, 1], p[1, 0] - p[0, 0]]]) / det
    Ke = (G @ G.T) * kap[g] * (det / 2 / 3)
    for pp in range(3):
        for qq in range(3):
            t = 9 * g + 3 * pp + qq
            assert I[t].tolist() == [v[pp], v[qq]]
            assert V[t] == pytest.approx(Ke[pp, qq], rel=1e-13, abs=1e-13)


def test_laplace_row_sums_symmetry_linearity():
    c, e = meshgen.jittered_square(7, seed=1)
    M = O.OracleMesh(c, e)
    rng = np.random.default_rng(0)
    k1, k2 = rng.random(M.ngauss) + 0.1, rng.random(M.ngauss) + 0.1
    K1 = O.csr_matrix(*M.laplace_coo(k1), M.nnode)
    K2 = O.csr_matrix(*M.laplace_coo(k2), M.nnode)
    K12 = O.csr_matrix(*M.laplace_coo(2 * k1 - 3 * k2), M.nnode)
    assert abs(K1 @ np.ones(M.nnode)).max() < 1e-12
    assert abs(K1 - K1.T).max() < 1e-13
    assert abs(K12 - (2 * K1 - 3 * K2)).max() < 1e-12


# ---------------------------------------------------------------- mass/source
def test_mass_closed_form_and_total_area():
    c, e = meshgen.jittered_square(5, seed=2)
    M = O.OracleMesh(c, e)
    I, V = M.mass_coo(np.ones(M.ngauss))
    Ve = V.reshape(-1, 3, 9).sum(1).reshape(-1, 3, 3)  # sum over Gauss points
    ref = (np.ones((3, 3)) + np.eye(3))[None] * (M.area / 12)[:, None, None]
    np.testing.assert_allclose(Ve, ref, rtol=1e-13)
    Msp = O.csr_matrix(I, V, M.nnode)
    assert Msp.sum() == pytest.approx(1.0, rel=1e-13)
    assert abs(Msp - Msp.T).max() < 1e-15


def test_source_equals_mass_times_ones_and_integrates_linear_f():
    """deps/MFEM/FemSource1/ftest.jl:4-15 (F(c) == M(c)*1), and sum F = ∫f
    for linear f (the 3-point rule is exact for quadratics)."""
    c, e = meshgen.jittered_square(6, seed=4)
    M = O.OracleMesh(c, e)
    cc = np.random.default_rng(1).random(M.ngauss)
    F = M.source(cc)
    Mc = O.csr_matrix(*M.mass_coo(cc), M.nnode)
    np.testing.assert_allclose(F, Mc @ np.ones(M.nnode), rtol=1e-13, atol=1e-16)
    f = 3 * M.gx - 2 * M.gy + 1.0
    assert M.source(f).sum() == pytest.approx(1.5 - 1.0 + 1.0, rel=1e-13)


# ---------------------------------------------------------------- Dirichlet
def test_dirichlet_matches_dense_block_formula():
    """test/mfem.jl:78-88: random 5x5, bd=[5;2;3], bdval=[1;2;3] against
    [A_II 0; 0 I], [r_I - A_IB u_B; u_B] (src/MFEM/MUtils.jl:176-183)."""
    rng = np.random.default_rng(7)
    A = rng.random((5, 5))
    rhs = rng.random(5)
    bd = np.array([5, 2, 3])
    bdval = np.array([1.0, 2.0, 3.0])
    jj, ii = np.meshgrid(np.arange(5), np.arange(5))
    I = np.stack([ii.T.ravel(), jj.T.ravel()], 1)  # column-major like sparse(A)
    V = A.T.ravel()
    oi, ov, orhs = O.dirichlet(I, V, bd, rhs, bdval)
    B = sp.coo_matrix((ov, (oi[:, 0], oi[:, 1])), shape=(5, 5)).toarray()
    idx = np.ones(5, bool)
    idx[bd - 1] = False
    A1 = np.zeros((5, 5))
    A1[np.ix_(idx, idx)] = A[np.ix_(idx, idx)]
    A1[bd - 1, bd - 1] = 1.0
    r1 = rhs.copy()
    r1[idx] = rhs[idx] - A[np.ix_(idx, bd - 1)] @ bdval
    r1[bd - 1] = bdval
    np.testing.assert_allclose(B, A1, atol=1e-15)
    np.testing.assert_allclose(orhs, r1, rtol=1e-14)
    # output order: kept triplets in input order, then ascending diagonals
    assert oi[-3:].tolist() == [[1, 1], [2, 2], [4, 4]]


def test_dirichlet_duplicate_bd_last_wins():
    I = np.array([[0, 0], [0, 1], [1, 1]])
    V = np.array([2.0, -1.0, 2.0])
    oi, ov, orhs = O.dirichlet(I, V, np.array([2, 2]), np.zeros(2), np.array([5.0, 7.0]))
    assert orhs.tolist() == [7.0, 7.0]
    assert oi.tolist() == [[0, 0], [1, 1]]


def _adjoint_check(fwd, bwd, x0, ydim, seed=0):
    """<g, J v> == <J^T g, v> for a linear map (J v by a forward difference,
    exact up to rounding for a linear map)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(x0.size)
    g = rng.standard_normal(ydim)
    Jv = fwd(x0 + v) - fwd(x0)
    lhs = g @ Jv
    rhs = bwd(g) @ v
    assert lhs == pytest.approx(rhs, rel=1e-10, abs=1e-12)


def test_backward_ops_are_adjoints_of_forward():
    c, e = meshgen.jittered_square(4, seed=5)
    M = O.OracleMesh(c, e)
    k0 = np.random.default_rng(3).random(M.ngauss)
    _adjoint_check(lambda k: M.laplace_coo(k)[1], M.laplace_backward, k0, 9 * M.ngauss)
    _adjoint_check(lambda k: M.mass_coo(k)[1], M.mass_backward, k0, 9 * M.ngauss)
    _adjoint_check(M.source, M.source_backward, k0, M.nnode)


def test_dirichlet_backward_is_adjoint():
    c, e = meshgen.jittered_square(4, seed=6)
    M = O.OracleMesh(c, e)
    I, V = M.laplace_coo(np.ones(M.ngauss))
    bd = M.bcnode()
    bd = np.concatenate([bd, bd[:3]])  # duplicates exercise the last-wins map
    rng = np.random.default_rng(2)
    rhs0, bv0 = rng.random(M.nnode), rng.random(bd.size)
    nv, nn, nb = V.size, M.nnode, bd.size
    x0 = np.concatenate([V, rhs0, bv0])

    def fwd(x):
        oi, ov, orhs = O.dirichlet(I, x[:nv], bd, x[nv:nv + nn], x[nv + nn:])
        return np.concatenate([ov, orhs])

    nout = O.dirichlet(I, V, bd, rhs0, bv0)[1].size

    def bwd(g):
        gv, gr, gb = O.dirichlet_backward(g[:nout], g[nout:], I, V, bd, bv0, nn)
        return np.concatenate([gv, gr, gb])

    # the map is bilinear in (vv, bdval) through the lifting: check the
    # adjoint at a point with bdval perturbations off
    rng2 = np.random.default_rng(9)
    v = np.concatenate([rng2.standard_normal(nv), rng2.standard_normal(nn), np.zeros(nb)])
    g = rng2.standard_normal(nout + nn)
    lhs = g @ (fwd(x0 + v) - fwd(x0))
    assert lhs == pytest.approx(bwd(g) @ v, rel=1e-10)
    v = np.concatenate([np.zeros(nv + nn), rng2.standard_normal(nb)])
    lhs = g @ (fwd(x0 + v) - fwd(x0))
    assert lhs == pytest.approx(bwd(g) @ v, rel=1e-10)


# ---------------------------------------------------------------- solve + gradient
def test_manufactured_solution_converges_second_order():
    """-Δu = 2π² sin πx sin πy, u = 0 on the boundary (docs/src/mfem_tutorial.md:29-60)."""
    errs = []
    for m in (8, 16, 32):
        M = om(m)
        f = 2 * np.pi ** 2 * np.sin(np.pi * M.gx) * np.sin(np.pi * M.gy)
        bd = M.bcnode()
        out = O.forward_adjoint(M, np.ones(M.ngauss), f, bd, np.zeros(bd.size),
                                lambda u: np.zeros_like(u))
        ue = np.sin(np.pi * M.coords[:, 0]) * np.sin(np.pi * M.coords[:, 1])
        errs.append(abs(out["u"] - ue).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 4.7 and 3.3 < r2 < 4.7, errs


def _problem(m, seed):
    c, e = meshgen.jittered_square(m, seed=seed)
    M = O.OracleMesh(c, e)
    kap = np.exp(meshgen.grf(np.stack([M.gx, M.gy], 1), seed + 1))
    f = np.random.default_rng(seed + 2).uniform(-1, 1, M.ngauss)
    bd = M.bcnode()
    rng = np.random.default_rng(seed + 3)
    interior = np.setdiff1d(np.arange(M.nnode), bd - 1)
    obs = rng.choice(interior, size=max(3, interior.size // 10), replace=False)
    uobs = rng.standard_normal(obs.size) * 1e-3
    return M, kap, f, bd, obs, uobs


def _dloss(obs, uobs):
    def g(u):
        out = np.zeros_like(u)
        out[obs] = 2 * (u[obs] - uobs)
        return out
    return g


def _loss(M, kap, f, bd, obs, uobs):
    out = O.forward_adjoint(M, kap, f, bd, np.zeros(bd.size), lambda u: np.zeros_like(u))
    return ((out["u"][obs] - uobs) ** 2).sum()


def test_gradient_chain_matches_central_differences():
    M, kap, f, bd, obs, uobs = _problem(6, 11)
    out = O.forward_adjoint(M, kap, f, bd, np.zeros(bd.size), _dloss(obs, uobs))
    rng = np.random.default_rng(5)
    eps = 1e-6
    for _ in range(5):
        d = rng.standard_normal(M.ngauss)
        fd = (_loss(M, kap + eps * d, f, bd, obs, uobs) -
              _loss(M, kap - eps * d, f, bd, obs, uobs)) / (2 * eps)
        assert out["grad_kappa"] @ d == pytest.approx(fd, rel=1e-6, abs=1e-14)


def test_gradient_chain_equals_closed_form():
    """-w_k ∇λ̃·∇u per element (λ̃: adjoint zeroed on the boundary)."""
    M, kap, f, bd, obs, uobs = _problem(7, 21)
    out = O.forward_adjoint(M, kap, f, bd, np.zeros(bd.size), _dloss(obs, uobs))
    lam = out["lam"].copy()
    lam[bd - 1] = 0.0
    u = out["u"]
    hx, hy = M.hx.reshape(-1, 3), M.hy.reshape(-1, 3)
    v = np.repeat(M.elems, 3, axis=0)
    du = np.stack([(hx * u[v]).sum(1), (hy * u[v]).sum(1)], 1)
    dl = np.stack([(hx * lam[v]).sum(1), (hy * lam[v]).sum(1)], 1)
    closed = -M.w * (du * dl).sum(1)
    np.testing.assert_allclose(out["grad_kappa"], closed, rtol=1e-9,
                               atol=1e-12 * abs(closed).max())


def test_heat_step_gradient_matches_differences():
    M, kap, f, bd, obs, uobs = _problem(5, 31)
    dt = 0.03
    up = np.sin(np.pi * M.coords[:, 0]) * np.sin(np.pi * M.coords[:, 1])
    kw = dict(dt=dt, u_prev=up)
    out = O.forward_adjoint(M, kap, f, bd, np.zeros(bd.size), _dloss(obs, uobs), **kw)

    def L(k):
        o = O.forward_adjoint(M, k, f, bd, np.zeros(bd.size), lambda u: 0 * u, **kw)
        return ((o["u"][obs] - uobs) ** 2).sum()

    d = np.random.default_rng(8).standard_normal(M.ngauss)
    eps = 1e-6
    fd = (L(kap + eps * d) - L(kap - eps * d)) / (2 * eps)
    assert out["grad_kappa"] @ d == pytest.approx(fd, rel=1e-6)


# ------------------------------------------------- fields at Gauss points
def test_fem_to_gauss_exact_for_linear_and_second_order_for_quadratic():
    """deps/MFEM/FemToGaussPoints/TestFemToGaussPoints.jl: P1 interpolation of
    nodal values to the Gauss points; exact for linear fields, O(h^2) for the
    test's 2x^2 + y."""
    for m in (8, 16):
        mesh = om(m)
        x, y = mesh.coords[:, 0], mesh.coords[:, 1]
        np.testing.assert_allclose(mesh.fem_to_gauss(1.5 + 2 * x - 3 * y),
                                   1.5 + 2 * mesh.gx - 3 * mesh.gy, rtol=1e-13, atol=1e-13)
        err = np.abs(mesh.fem_to_gauss(2 * x ** 2 + y) - (2 * mesh.gx ** 2 + mesh.gy)).max()
        assert err < 2 * (1.0 / m) ** 2
    assert err > 0.1 * 2 * (1.0 / 16) ** 2 / 4   # not exact: really interpolates


def test_fem_grad_exact_for_linear_fields():
    """deps/MFEM/FemGrad/TestFemGrad.jl: f = 2x + x + 14y has gradient (3, 14)
    at every Gauss point; layout [g][x, y]."""
    c, e = meshgen.jittered_square(7, seed=5)
    mesh = O.OracleMesh(c, e)
    g = mesh.fem_grad(3 * c[:, 0] + 14 * c[:, 1])
    assert g.shape == (mesh.ngauss, 2)
    np.testing.assert_allclose(g[:, 0], 3.0, rtol=1e-12)
    np.testing.assert_allclose(g[:, 1], 14.0, rtol=1e-12)


def test_laplace_term_equals_stiffness_times_u():
    """deps/MFEM/ComputeLaplaceTermMfem/ftest.jl: compute_fem_laplace_term1(u,
    nu) = compute_fem_laplace_matrix1(nu) * u."""
    c, e = meshgen.jittered_square(9, seed=2)
    mesh = O.OracleMesh(c, e)
    rng = np.random.default_rng(0)
    nu, u = rng.random(mesh.ngauss) + 0.5, rng.standard_normal(mesh.nnode)
    K = O.csr_matrix(*mesh.laplace_coo(nu), mesh.nnode)
    np.testing.assert_allclose(mesh.laplace_term(u, nu), K @ u, rtol=1e-12,
                               atol=1e-12 * np.abs(K @ u).max())


def test_field_op_backwards_are_adjoints():
    c, e = meshgen.jittered_square(6, seed=8)
    mesh = O.OracleMesh(c, e)
    rng = np.random.default_rng(1)
    u, gq = rng.standard_normal(mesh.nnode), rng.standard_normal(mesh.ngauss)
    assert np.isclose(gq @ mesh.fem_to_gauss(u), mesh.fem_to_gauss_backward(gq) @ u, rtol=1e-12)
    g2 = rng.standard_normal((mesh.ngauss, 2))
    assert np.isclose((g2 * mesh.fem_grad(u)).sum(), mesh.fem_grad_backward(g2) @ u, rtol=1e-12)
    nu, go = rng.random(mesh.ngauss) + 0.5, rng.standard_normal(mesh.nnode)
    gn, gu = mesh.laplace_term_backward(u, nu, go)
    out = mesh.laplace_term(u, nu)
    assert np.isclose(go @ out, gu @ u, rtol=1e-12)      # linear in u
    assert np.isclose(go @ out, gn @ nu, rtol=1e-12)     # linear in nu


# ---------------------------------------------------------------- pcl Jacobians
def test_pcl_jacobians_are_derivatives_of_the_forward_ops():
    """src/pcl.jl's two Jacobians against the (already pinned) forward ops
    they differentiate, both of which are linear in the differentiated
    input: ImposeDirichlet's output values are J^T v plus the appended
    (b,b,1) constants, and the Laplace COO values are H^T kappa."""
    mesh = om(3)
    rng = np.random.default_rng(12)
    kap = rng.random(mesh.ngauss)
    oi, vv = mesh.laplace_coo(kap)
    H = O.pcl_laplace_jacobian(mesh)
    assert H.shape == (mesh.ngauss, 9 * mesh.ngauss)
    np.testing.assert_allclose(H.T @ kap, vv, rtol=1e-13, atol=1e-15)
    bd = mesh.bcnode()
    v = rng.standard_normal(vv.size)
    pi, pv, _ = O.dirichlet(oi, v, bd, np.zeros(mesh.nnode), np.zeros(bd.size))
    J = O.pcl_impose_dirichlet(oi + 1, bd, pv.size)
    kept = pv.size - np.unique(bd).size
    np.testing.assert_array_equal(J.T @ v, np.concatenate([pv[:kept], np.zeros(pv.size - kept)]))
    assert J.sum() == kept


def test_library_pcl_impose_dirichlet_matches_oracle():
    """The library's pcl_ImposeDirichlet (host-side bookkeeping, no device
    work, so callable here) against the oracle, with the reference's 1-based
    column-major layouts (src/pcl.jl:14-21)."""
    from admfem import _lib
    mesh = om(4)
    oi, _ = mesh.laplace_coo(np.ones(mesh.ngauss))
    bd = mesh.bcnode()
    sN = oi.shape[0]
    outdof = sN  # wider than needed: the extra columns stay zero
    ind = np.ascontiguousarray((oi + 1).T.ravel())  # [sN x 2] column-major, 1-based
    J = np.zeros(sN * outdof)
    _lib.lib().pcl_ImposeDirichlet(_lib.ptr(J), _lib.ptr(ind), _lib.ptr(bd), bd.size, sN)
    np.testing.assert_array_equal(J.reshape(outdof, sN).T, O.pcl_impose_dirichlet(oi + 1, bd, outdof))
