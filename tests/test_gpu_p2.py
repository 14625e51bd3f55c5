"""P2 (degree 2) on the GPU through the C-ABI against the oracle's P2
restatement: the mesh (dofs, edges, boundary dofs, Gauss points of the
6-point rule), the COO operators and their Grad ops, the load vector, the
field maps, the matrix-free Laplace term, the reference's host-pointer
symbols on a P2 default mesh, and a Poisson solve with its gradient.

Bars as for P1: indices bit-exact; operator values <= 1e-13 relative (the
gradients of the basis take a different rounding route: barycentric
gradients on the GPU, dshape_ref J^-1 in the oracle); the load vector and
mass values follow the reference's operation order term by term; solves
and gradients <= 1e-10 against the oracle's direct solve.
"""
import ctypes

import numpy as np
import pytest
import torch

import admfem
from admfem import _lib, meshgen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a)).to(device="cuda", dtype=dtype).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def pair(m, seed=None, order=-1):
    if seed is None:
        c, e = meshgen.rectangle(m, m, 1.0 / m)
    else:
        c, e = meshgen.jittered_square(m, seed=seed)   # node ids permuted
    gm = admfem.Mesh(c, e, order, 2, zero_based=True)
    om = O.OracleMesh(c, e, order=order, degree=2)
    return gm, om


# m = 190 (36,481 nodes) splits the edge and dof set-up over host threads
@pytest.mark.parametrize("m,seed", [(2, None), (7, 3), (190, 5)])
def test_p2_mesh_matches_oracle(m, seed):
    print(f"P2 mesh m={m}", flush=True)
    gm, om = pair(m, seed)
    assert gm.elem_ndof == 6 and gm.ndof == om.ndof and gm.ngauss == om.ngauss == 6 * om.nelem
    np.testing.assert_array_equal(gm.elems - 1, om.elems)
    np.testing.assert_array_equal(gm.conn - 1, om.dofs)
    np.testing.assert_array_equal(gm.edges, om.edges())
    np.testing.assert_array_equal(admfem.bcnode(gm), om.bcnode())
    g = admfem.gauss_nodes(gm)
    np.testing.assert_allclose(g[:, 0], om.gx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g[:, 1], om.gy, rtol=0, atol=1e-15)
    np.testing.assert_allclose(admfem.gauss_weights(gm), om.w, rtol=1e-14)


def test_p2_coo_ops_and_grads_match_oracle():
    print("P2 COO ops m=6", flush=True)
    gm, om = pair(6, 4)
    L = _lib.lib()
    rng = np.random.default_rng(1)
    kap = rng.random(om.ngauss) + 0.5
    nt = 36 * om.ngauss
    ind = torch.empty((nt, 2), dtype=torch.int64, device="cuda")
    vv = torch.empty(nt, dtype=torch.float64, device="cuda")
    dk = dev(kap)
    gvv = rng.standard_normal(nt)
    dgvv = dev(gvv)
    out = torch.empty(om.ngauss, dtype=torch.float64, device="cuda")
    for name, fwd, bwd, ofwd, obwd, tol in (
            ("laplace", L.admfem_fem_laplace_scalar, L.admfem_fem_laplace_scalar_grad,
             om.laplace_coo, om.laplace_backward, 1e-13),
            ("mass", L.admfem_fem_mass_matrix, L.admfem_fem_mass_matrix_grad,
             om.mass_coo, om.mass_backward, 1e-15)):
        _lib.check(fwd(gm.handle, _lib.ptr(dk), _lib.ptr(ind), _lib.ptr(vv), _lib.stream_ptr()))
        oi, ov = ofwd(kap)
        np.testing.assert_array_equal(host(ind), oi)
        assert rel(host(vv), ov) < tol, name
        _lib.check(bwd(gm.handle, _lib.ptr(dgvv), _lib.ptr(out), _lib.stream_ptr()))
        assert rel(host(out), obwd(gvv)) < 1e-12, name


def test_p2_source_and_grad_match_oracle():
    gm, om = pair(9, 2)
    L = _lib.lib()
    rng = np.random.default_rng(2)
    f = rng.uniform(-1, 1, om.ngauss)
    rhs = torch.empty(om.ndof, dtype=torch.float64, device="cuda")
    df = dev(f)
    _lib.check(L.admfem_fem_source_scalar(gm.handle, _lib.ptr(df), _lib.ptr(rhs),
                                          _lib.stream_ptr()))
    np.testing.assert_allclose(host(rhs), om.source(f), rtol=1e-14, atol=1e-17)
    g = rng.standard_normal(om.ndof)
    gf = torch.empty(om.ngauss, dtype=torch.float64, device="cuda")
    dg = dev(g)
    _lib.check(L.admfem_fem_source_scalar_grad(gm.handle, _lib.ptr(dg), _lib.ptr(gf),
                                               _lib.stream_ptr()))
    np.testing.assert_allclose(host(gf), om.source_backward(g), rtol=1e-13, atol=1e-16)
    # FemSource1/ftest.jl (degree 2): source(c) = mass(c) * ones, on the GPU
    M = admfem.compute_fem_mass_matrix1(torch.tensor(f, device="cuda"), gm)
    Mone = host(M @ torch.ones(om.ndof, dtype=torch.float64, device="cuda"))
    np.testing.assert_allclose(host(rhs), Mone, rtol=1e-12, atol=1e-15)


def test_p2_field_ops_and_grads_match_oracle():
    print("P2 field ops m=8", flush=True)
    gm, om = pair(8, 6)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(om.ndof)
    uv = rng.standard_normal(om.nnode)
    nu = rng.random(om.ngauss) + 0.5
    assert rel(admfem.fem_to_gauss_points(uv, gm), om.fem_to_gauss(uv)) < 1e-14
    assert rel(admfem.dof_to_gauss_points(u, gm), om.dof_to_gauss(u)) < 1e-14
    assert rel(admfem.eval_grad_on_gauss_pts1(u, gm), om.fem_grad(u)) < 1e-12
    assert rel(admfem.compute_fem_laplace_term1(u, nu, gm), om.laplace_term(u, nu)) < 1e-12
    ut = torch.tensor(u, device="cuda", requires_grad=True)
    nt = torch.tensor(nu, device="cuda", requires_grad=True)
    gq = rng.standard_normal(om.ngauss)
    (admfem.dof_to_gauss_points(ut, gm) * torch.tensor(gq, device="cuda")).sum().backward()
    assert rel(ut.grad.cpu().numpy(), om.dof_to_gauss_backward(gq)) < 1e-13
    ut.grad = None
    g2 = rng.standard_normal((om.ngauss, 2))
    (admfem.eval_grad_on_gauss_pts1(ut, gm) * torch.tensor(g2, device="cuda")).sum().backward()
    assert rel(ut.grad.cpu().numpy(), om.fem_grad_backward(g2)) < 1e-12
    ut.grad = None
    go = rng.standard_normal(om.ndof)
    (admfem.compute_fem_laplace_term1(ut, nt, gm) * torch.tensor(go, device="cuda")).sum().backward()
    gn0, gu0 = om.laplace_term_backward(u, nu, go)
    assert rel(ut.grad.cpu().numpy(), gu0) < 1e-12
    assert rel(nt.grad.cpu().numpy(), gn0) < 1e-12
    vt = torch.tensor(uv, device="cuda", requires_grad=True)
    (admfem.fem_to_gauss_points(vt, gm) * torch.tensor(gq, device="cuda")).sum().backward()
    assert rel(vt.grad.cpu().numpy(), om.fem_to_gauss_backward(gq)) < 1e-13


def test_p2_poisson_solve_and_gradient_match_oracle():
    """K(kappa) u = F with homogeneous Dirichlet on the P2 boundary dofs,
    L = sum (u - u_obs)^2 over observed dofs, dL/dkappa by torch autograd
    through the GPU ops (COO -> ImposeDirichlet -> sparse CG solve) against
    the oracle's SuperLU solve and Grad-op chain."""
    print("P2 Poisson m=12", flush=True)
    gm, om = pair(12, 9)
    rng = np.random.default_rng(4)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 2))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    obs = rng.choice(np.setdiff1d(np.arange(om.ndof), bd - 1), 40, replace=False)
    uobs = rng.standard_normal(obs.size) * 0.01
    ref = O.forward_adjoint(om, kap, f, bd, np.zeros(bd.size),
                            lambda u: np.bincount(obs, 2 * (u[obs] - uobs), minlength=om.ndof))
    kt = torch.tensor(kap, device="cuda", requires_grad=True)
    K = admfem.compute_fem_laplace_matrix1(kt, gm)
    F = admfem.compute_fem_source_term1(torch.tensor(f, device="cuda"), gm)
    Kb, Fb = admfem.impose_Dirichlet_boundary_conditions(K, F, bd, torch.zeros(bd.size,
                                                                                 device="cuda",
                                                                                 dtype=torch.float64))
    u = admfem.solve(Kb, Fb, rtol=1e-13)
    assert rel(host(u.detach()), ref["u"]) < 1e-10
    loss = ((u[torch.tensor(obs, device="cuda")] - torch.tensor(uobs, device="cuda")) ** 2).sum()
    loss.backward()
    assert rel(host(kt.grad), ref["grad_kappa"]) < 1e-9


def test_p2_legacy_symbols_on_default_mesh():
    """init_nnfem_mesh(..., degree = 2) and the getters / _Julia ops the
    reference's Julia binds (src/MFEM/MFEM.jl:93-106, MCore.jl:80,117)."""
    L = _lib.lib()
    c, e = meshgen.rectangle(3, 3, 1.0 / 3)
    om = O.OracleMesh(c, e, degree=2)
    verts = np.ascontiguousarray(np.concatenate([c, np.zeros((c.shape[0], 1))], 1).ravel())
    el = np.ascontiguousarray(e.astype(np.int32).ravel())
    ne = np.zeros(1, dtype=np.int64)
    p = L.init_nnfem_mesh(_lib.ptr(verts), c.shape[0], _lib.ptr(el), e.shape[0], -1, 6, 2,
                          _lib.ptr(ne))
    assert bool(p)
    ctypes.CDLL(None).free(ctypes.cast(p, ctypes.c_void_p))
    assert int(ne[0]) == om.nedge
    assert L.mfem_get_elem_ndof() == 6 and L.mfem_get_ndof() == om.ndof
    assert L.mfem_get_ngauss() == om.ngauss
    conn = np.zeros(6 * om.nelem, dtype=np.int64)
    L.mfem_get_connectivity(_lib.ptr(conn))
    np.testing.assert_array_equal(conn.reshape(-1, 6), om.dofs + 1)
    kap = np.random.default_rng(0).random(om.ngauss)
    ind = np.zeros(2 * 36 * om.ngauss, dtype=np.int64)
    vv = np.zeros(36 * om.ngauss)
    L.FemLaplaceScalar_forward_Julia(_lib.ptr(ind), _lib.ptr(vv), _lib.ptr(kap))
    oi, ov = om.laplace_coo(kap)
    np.testing.assert_array_equal(ind.reshape(-1, 2), oi)
    assert rel(vv, ov) < 1e-13
    rhs = np.ones(om.ndof)  # accumulates
    L.FemSourceScalar_forward_Julia(_lib.ptr(rhs), _lib.ptr(kap))
    np.testing.assert_allclose(rhs, 1.0 + om.source(kap), rtol=1e-14)
    u = np.random.default_rng(1).standard_normal(om.ndof)
    out = np.zeros(om.ngauss)
    L.DofToGaussPointsMfem_forward_Julia(_lib.ptr(out), _lib.ptr(u))
    assert rel(out, om.dof_to_gauss(u)) < 1e-14
    lt = np.zeros(om.ndof)
    L.ComputeLaplaceTermMfem_forward_Julia(_lib.ptr(lt), _lib.ptr(kap), _lib.ptr(u))
    assert rel(lt, om.laplace_term(u, kap)) < 1e-12


def test_p1_order_four_fused_system_matches_oracle():
    """The 6-point rule on a P1 mesh through the fused system path (order
    keyword of Mesh, src/MFEM/MFEM.jl:60)."""
    c, e = meshgen.jittered_square(10, seed=11)
    gm = admfem.Mesh(c, e, 4, zero_based=True)
    om = O.OracleMesh(c, e, order=4)
    assert gm.ngauss == om.ngauss == 6 * om.nelem
    rng = np.random.default_rng(6)
    kap, f = rng.random(om.ngauss) + 0.5, rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    S = admfem.DirichletSystem(gm, bd)
    S.assemble(kappa=kap, f=f)
    A, r = S.csr()
    oi, ov, orhs = O.dirichlet(*om.laplace_coo(kap), bd, om.source(f), np.zeros(bd.size))
    A0 = O.csr_matrix(oi, ov, om.nnode)
    A0.sort_indices()
    np.testing.assert_array_equal(A.indptr, A0.indptr)
    np.testing.assert_array_equal(A.indices, A0.indices)
    assert rel(A.data, A0.data) < 1e-13 and rel(r, orhs) < 1e-13
