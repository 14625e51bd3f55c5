"""The fused P2 system (admfem_system_create on a degree-2 mesh) and P2
requests in the serving engine, on the GPU, against the oracle's P2 chain: P2 COO operators (FemLaplaceScalar /
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


def p2_topology(m):
    c, e = meshgen.jittered_square(m, seed=0, jitter=0.0, permute=False)
    gm = admfem.Mesh(c, e, -1, 2, zero_based=True)
    return c, e, admfem.DirichletSystem(gm, admfem.bcnode(gm))


def p2_request(c, e, seed, heat):
    """Host inputs of one P2 request: the topology's interior vertices
    jittered, kappa = exp(GRF) and f ~ U[-1, 1] at the Gauss points, 5% of
    the interior dofs observed, and for a heat step u_prev at the dofs."""
    rng = np.random.default_rng(seed)
    m = int(round(np.sqrt(c.shape[0]))) - 1
    cc = c.copy()
    inner = (c[:, 0] > 0) & (c[:, 0] < 1) & (c[:, 1] > 0) & (c[:, 1] < 1)
    cc[inner] += rng.uniform(-0.2 / m, 0.2 / m, size=(inner.sum(), 2))
    om = O.OracleMesh(cc, e, degree=2)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), seed + 1))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    interior = np.setdiff1d(np.arange(om.ndof), bd - 1)
    obs = rng.choice(interior, size=interior.size // 20, replace=False).astype(np.int32)
    uobs = rng.standard_normal(obs.size) * 1e-3
    up = rng.standard_normal(om.ndof) if heat else None
    return dict(coords=cc, om=om, kappa=kap, f=f, bd=bd, obs=obs, uobs=uobs, u_prev=up,
                dt=0.02 if heat else None)


def p2_oracle(P):
    kw = {} if P["dt"] is None else dict(dt=P["dt"], u_prev=P["u_prev"])
    return O.forward_adjoint(
        P["om"], P["kappa"], P["f"], P["bd"], np.zeros(P["bd"].size),
        lambda u: np.bincount(P["obs"], 2 * (u[P["obs"]] - P["uobs"]), minlength=u.size), **kw)


def p2_submit(sv, tid, P):
    return sv.submit(tid, kind="heat" if P["dt"] else "poisson", dt=P["dt"] or 0.0,
                     coords=P["coords"], kappa=P["kappa"], f=P["f"], u_prev=P["u_prev"],
                     obs_idx=P["obs"], obs_val=P["uobs"])


@pytest.mark.parametrize("mode", ["batched", "seq_fp64"])
def test_p2_serving_matches_oracle(mode):
    """P2 topologies in the serving engine beside a P1 one: Poisson and heat
    requests with host inputs against the oracle's P2 chain."""
    print(f"P2 serving vs oracle, mode={mode}", flush=True)
    sv = Server(2 << 30, mode=mode, rtol=1e-11)
    cases = []
    for m, heat, seed in ((8, False, 1), (11, True, 2), (6, False, 3)):
        c, e, S = p2_topology(m)
        tid = sv.add_topology(S, jitter=0.2 / m)
        P = p2_request(c, e, seed, heat)
        cases.append((p2_submit(sv, tid, P), P))
    c1, e1 = meshgen.jittered_square(16, seed=0, jitter=0.0, permute=False)
    g1 = admfem.Mesh(c1, e1, zero_based=True)
    t1 = sv.add_topology(admfem.DirichletSystem(g1, admfem.bcnode(g1)), 0.2 / 16)
    others = [sv.submit(t1, seed=s) for s in range(3)]
    sv.run()
    for rid, P in cases:
        r = sv.result(rid)
        assert r["done"] and r["status"] == 0, r
        ref = p2_oracle(P)
        assert rel(r["u"], ref["u"]) < 1e-8
        assert rel(r["grad"], ref["grad_kappa"]) < 1e-8
        assert r["loss"] == pytest.approx(((ref["u"][P["obs"]] - P["uobs"]) ** 2).sum(), rel=1e-8)
    assert all(sv.result(i)["status"] == 0 for i in others)


def test_p2_serving_same_alone_and_in_mix():
    """A P2 request's result is bit for bit the same alone and among P1 and
    synthetic P2 requests (heat and Poisson)."""
    print("P2 serving alone vs mix", flush=True)
    c, e, S = p2_topology(10)
    P = p2_request(c, e, 9, heat=True)
    alone = Server(1 << 30, rtol=1e-10)
    ra = p2_submit(alone, alone.add_topology(S, 0.02), P)
    alone.run()
    mix = Server(1 << 30, rtol=1e-10)
    tp2 = mix.add_topology(S, 0.02)
    c1, e1 = meshgen.jittered_square(30, seed=0, jitter=0.0, permute=False)
    g1 = admfem.Mesh(c1, e1, zero_based=True)
    tp1 = mix.add_topology(admfem.DirichletSystem(g1, admfem.bcnode(g1)), 0.2 / 30)
    ids = [mix.submit(tp1 if s % 2 else tp2, kind="heat" if s % 3 == 0 else "poisson",
                      dt=0.05, seed=50 + s) for s in range(6)]
    rm = p2_submit(mix, tp2, P)
    mix.run()
    a, b = alone.result(ra), mix.result(rm)
    assert a["status"] == 0 and b["status"] == 0
    assert mix.stats()["peak_active"] > 1
    np.testing.assert_array_equal(a["u"], b["u"])
    np.testing.assert_array_equal(a["grad"], b["grad"])
    for i in ids:
        r = mix.result(i)
        assert r["status"] == 0 and np.isfinite(r["grad"]).all()
