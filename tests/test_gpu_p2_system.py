"""The fused P2 system (admfem_system_create on a degree-2 mesh) on the GPU
against the oracle's P2 chain: P2 COO operators (FemLaplaceScalar /
ComputeFemMassMatrixMfem / FemSourceScalar, P2 branch) -> ImposeDirichlet ->
direct solve -> the Grad-op chain for dL/dkappa.

Bars as for the P1 fused system: the post-Dirichlet pattern exact (explicit
zeros kept), values and rhs <= 1e-13 relative; solves and gradients
<= 1e-10 (FP64 CG to 1e-13) or <= 1e-8 (mixed CG to 1e-11) against the
oracle's direct solve.
"""
import ctypes

import numpy as np
import pytest
import torch

import admfem
from admfem import _lib, meshgen
from admfem.server import Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def pair(m, seed):
    if seed is None:
        c, e = meshgen.rectangle(m, m, 1.0 / m)   # right triangles: exact zeros in K
    else:
        c, e = meshgen.jittered_square(m, seed=seed)   # node ids permuted
    return admfem.Mesh(c, e, -1, 2, zero_based=True), O.OracleMesh(c, e, degree=2)


def oracle_system(om, kap, f, bd, bdval, dt=None, u_prev=None, rho=None):
    I, V = om.laplace_coo(kap)
    F = om.source(f)
    if dt is None:
        AI, AV, rhs = I, V, F
    else:
        MI, MV = om.mass_coo(rho if rho is not None else np.ones(om.ngauss))
        AI = np.concatenate([MI, I])
        AV = np.concatenate([MV, dt * V])
        rhs = O.csr_matrix(MI, MV, om.ndof) @ u_prev + dt * F
    oi, ov, orhs = O.dirichlet(AI, AV, bd, rhs, bdval)
    A = O.csr_matrix(oi, ov, om.ndof).tocsr()
    A.sort_indices()
    return A, orhs


def inputs(om, seed):
    rng = np.random.default_rng(seed)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), seed))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    return rng, kap, f, bd


# heat: False (Poisson), "rho" (a mass density field), "unit" (rho = 1)
@pytest.mark.parametrize("m,seed,heat", [(9, 3, False), (9, 3, "rho"), (9, 3, "unit"),
                                         (8, None, False), (8, None, "rho"), (70, 5, False)])
def test_p2_system_assembly_matches_oracle(m, seed, heat):
    print(f"P2 system assembly m={m} seed={seed} heat={heat}", flush=True)
    gm, om = pair(m, seed)
    rng, kap, f, bd = inputs(om, 5)
    bdval = rng.standard_normal(bd.size)
    bdval[::3] = 0.0
    S = admfem.DirichletSystem(gm, bd)
    assert S.n == om.ndof
    if heat:
        dt, up = 0.05, rng.standard_normal(om.ndof)
        rho = rng.random(om.ngauss) + 0.5 if heat == "rho" else None
        S.assemble(kappa=kap, f=f, bdval=bdval, rho=rho, u_prev=up, lap_coef=dt, mass_coef=1.0,
                   src_coef=dt)
        A0, r0 = oracle_system(om, kap, f, bd, bdval, dt=dt, u_prev=up, rho=rho)
    else:
        S.assemble(kappa=kap, f=f, bdval=bdval)
        A0, r0 = oracle_system(om, kap, f, bd, bdval)
    A, r = S.csr()
    assert S.nnz == A0.nnz
    np.testing.assert_array_equal(A.indptr, A0.indptr)    # pattern exact
    np.testing.assert_array_equal(A.indices, A0.indices)
    assert rel(A.data, A0.data) < 1e-13
    assert rel(r, r0) < 1e-13
    if seed is None and not heat:   # the structured mesh's stiffness has exact zeros
        assert (A0.data == 0).any()


@pytest.mark.parametrize("m,prec,heat", [(12, "fp64", False), (12, "mixed", False),
                                         (12, "fp64", True), (150, "fp64", False)])
def test_p2_system_forward_adjoint_matches_oracle(m, prec, heat):
    """u = A^-1 rhs, L = sum (u[obs] - obs_val)^2, lam = A^-1 dL/du and
    dL/dkappa through admfem_system_forward_adjoint, against the oracle's
    direct solve and Grad-op chain (oracle.forward_adjoint)."""
    print(f"P2 forward+adjoint m={m} {prec} heat={heat}", flush=True)
    gm, om = pair(m, 4)
    rng, kap, f, bd = inputs(om, 6)
    bdval = rng.standard_normal(bd.size) * 0.1
    obs = rng.choice(np.setdiff1d(np.arange(om.ndof), bd - 1), min(200, om.ndof // 4),
                     replace=False)
    uobs = rng.standard_normal(obs.size) * 0.01
    kw, okw = {}, {}
    if heat:
        dt, up, rho = 0.05, rng.standard_normal(om.ndof), rng.random(om.ngauss) + 0.5
        kw = dict(u_prev=up, rho=rho, lap_coef=dt, mass_coef=1.0, src_coef=dt)
        okw = dict(dt=dt, u_prev=up, rho=rho)
    ref = O.forward_adjoint(om, kap, f, bd, bdval,
                            lambda u: np.bincount(obs, 2 * (u[obs] - uobs), minlength=om.ndof),
                            **okw)
    S = admfem.DirichletSystem(gm, bd)
    tol = 1e-10 if prec == "fp64" else 1e-8
    out = S.forward_adjoint(obs, uobs, precision=prec, rtol=1e-13 if prec == "fp64" else 1e-11,
                            kappa=kap, f=f, bdval=bdval, **kw)
    print(out["stats"], flush=True)
    assert all(s["status"] == 0 for s in out["stats"])
    u = host(out["u"])
    assert rel(u, ref["u"]) < tol
    lam = host(out["lam"])
    interior = np.setdiff1d(np.arange(om.ndof), bd - 1)
    assert rel(lam[interior], ref["lam"][interior]) < tol
    assert abs(out["loss"] - ((ref["u"][obs] - uobs) ** 2).sum()) <= tol * out["loss"] + 1e-300
    assert rel(host(out["grad_kappa"]), ref["grad_kappa"]) < 10 * tol


def test_p2_system_gradient_entry_points():
    """admfem_system_kappa_gradient and its allocation-free variant on a P2
    system: equal to each other bit for bit and to the forward_adjoint
    gradient from the same u, lam."""
    print("P2 gradient entry points m=20", flush=True)
    gm, om = pair(20, 8)
    rng, kap, f, bd = inputs(om, 7)
    obs = rng.choice(np.setdiff1d(np.arange(om.ndof), bd - 1), 50, replace=False)
    S = admfem.DirichletSystem(gm, bd)
    out = S.forward_adjoint(obs, np.zeros(obs.size), precision="fp64", rtol=1e-12, kappa=kap,
                            f=f)
    g_fa = host(out["grad_kappa"])
    u, lam = out["u"], out["lam"]
    L = _lib.lib()
    got = []
    for fn in (L.admfem_system_kappa_gradient, L.admfem_system_kappa_gradient_noalloc):
        g = torch.full((om.ngauss,), np.nan, dtype=torch.float64, device="cuda")
        _lib.check(fn(S._h, _lib.ptr(u), _lib.ptr(lam), ctypes.c_double(1.0), _lib.ptr(g),
                      _lib.stream_ptr()))
        got.append(host(g))
    np.testing.assert_array_equal(got[0], got[1])
    assert rel(got[0], g_fa) < 1e-14
    # the oracle's Grad op of FemLaplaceScalar on the same u, lam~ (boundary rows zeroed)
    lt = host(lam).copy()
    lt[bd - 1] = 0.0
    uh = host(u)
    I, V = om.laplace_coo(np.ones(om.ngauss))
    gvv = -lt[I[:, 0]] * uh[I[:, 1]]
    assert rel(got[0], om.laplace_backward(gvv)) < 1e-12


def test_p2_system_refused_by_server():
    gm, om = pair(6, 2)
    S = admfem.DirichletSystem(gm, om.bcnode())
    srv = Server(1 << 28)
    with pytest.raises(Exception) as ei:
        srv.add_topology(S)
    assert "P1" in str(ei.value)
    srv.close()
