"""Fields at Gauss points and the matrix-free Laplace term on the GPU
(FemToGaussPoints, DofToGaussPoints, FemGrad, ComputeLaplaceTerm and their
Grad ops) against the oracle's restatement of the reference loops, through
the C-ABI (handle ops, torch autograd, and the reference's host-pointer
symbols on the default mesh)."""
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


def pair(m, seed):
    c, e = meshgen.jittered_square(m, seed=seed)   # node ids randomly permuted
    return admfem.Mesh(c, e, zero_based=True), O.OracleMesh(c, e)


def test_field_ops_and_grads_match_oracle():
    print("field ops m=13", flush=True)
    gm, om = pair(13, 4)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(om.nnode)
    nu = rng.random(om.ngauss) + 0.5
    assert rel(admfem.fem_to_gauss_points(u, gm), om.fem_to_gauss(u)) < 1e-14
    assert rel(admfem.dof_to_gauss_points(u, gm), om.fem_to_gauss(u)) < 1e-14
    assert rel(admfem.eval_grad_on_gauss_pts1(u, gm), om.fem_grad(u)) < 1e-12
    assert rel(admfem.compute_fem_laplace_term1(u, nu, gm), om.laplace_term(u, nu)) < 1e-12
    # gradients through torch autograd, against the oracle's Grad loops
    ut = torch.tensor(u, device="cuda", requires_grad=True)
    nt = torch.tensor(nu, device="cuda", requires_grad=True)
    gq = rng.standard_normal(om.ngauss)
    (admfem.fem_to_gauss_points(ut, gm) * torch.tensor(gq, device="cuda")).sum().backward()
    assert rel(ut.grad.cpu().numpy(), om.fem_to_gauss_backward(gq)) < 1e-13
    ut.grad = None
    g2 = rng.standard_normal((om.ngauss, 2))
    (admfem.eval_grad_on_gauss_pts1(ut, gm) * torch.tensor(g2, device="cuda")).sum().backward()
    assert rel(ut.grad.cpu().numpy(), om.fem_grad_backward(g2)) < 1e-12
    ut.grad = None
    go = rng.standard_normal(om.nnode)
    (admfem.compute_fem_laplace_term1(ut, nt, gm) * torch.tensor(go, device="cuda")).sum().backward()
    gn0, gu0 = om.laplace_term_backward(u, nu, go)
    assert rel(ut.grad.cpu().numpy(), gu0) < 1e-12
    assert rel(nt.grad.cpu().numpy(), gn0) < 1e-12


def test_laplace_term_equals_assembled_matrix_product():
    """ComputeLaplaceTermMfem/ftest.jl: K(nu) u matrix-free = the assembled
    K times u, both on the GPU."""
    gm, om = pair(20, 7)
    rng = np.random.default_rng(2)
    u, nu = rng.standard_normal(om.nnode), rng.random(om.ngauss) + 0.1
    K = admfem.compute_fem_laplace_matrix1(nu, gm)     # Array path: SciPy
    assert rel(admfem.compute_fem_laplace_term1(u, nu, gm), K @ u) < 1e-12


def test_field_legacy_symbols_on_default_mesh():
    """FemToGaussPointsMfem_Julia, DofToGaussPointsMfem_forward_Julia and
    ComputeLaplaceTermMfem_forward_Julia with host pointers on the mesh of
    init_nnfem_mesh, as src/MFEM/MUtils.jl:248,268 and MCore.jl:744 call
    them."""
    L = _lib.lib()
    c, e = meshgen.rectangle(5, 4, 0.2)
    om = O.OracleMesh(c, e)
    verts = np.ascontiguousarray(np.concatenate([c, np.zeros((c.shape[0], 1))], 1).ravel())
    el = np.ascontiguousarray(e.astype(np.int32).ravel())
    ne = np.zeros(1, dtype=np.int64)
    p = L.init_nnfem_mesh(_lib.ptr(verts), c.shape[0], _lib.ptr(el), e.shape[0], 2, 6, 1,
                          _lib.ptr(ne))
    assert bool(p)
    ctypes.CDLL(None).free(ctypes.cast(p, ctypes.c_void_p))
    rng = np.random.default_rng(3)
    u, nu = rng.standard_normal(om.nnode), rng.random(om.ngauss)
    out = np.zeros(om.ngauss)
    L.FemToGaussPointsMfem_Julia(_lib.ptr(out), _lib.ptr(u))
    assert rel(out, om.fem_to_gauss(u)) < 1e-14
    out[:] = 0
    L.DofToGaussPointsMfem_forward_Julia(_lib.ptr(out), _lib.ptr(u))
    assert rel(out, om.fem_to_gauss(u)) < 1e-14
    acc = np.ones(om.nnode)  # accumulates
    L.ComputeLaplaceTermMfem_forward_Julia(_lib.ptr(acc), _lib.ptr(nu), _lib.ptr(u))
    assert rel(acc - 1.0, om.laplace_term(u, nu)) < 1e-12
