"""The benchmark configurations as parity cases on the GPU.

* 64x64 README case (Mesh(64,64,1/64), kappa = sin(x)(1+y^2)+1,
  f = 1e5(x+y)) and its 8x8 / heat-step companions: the HIP path against the
  committed golden fixtures (tests/golden/).
* 1024x1024 jittered, node-permuted mesh, plain FP64 path: u and dL/dkappa
  within 1e-10 of the oracle's direct solve.
* 512x512 differentiation form: the mixed-precision dL/dkappa against central
  finite differences (step 1e-6) in 20 seeded directions and against the
  oracle's gradient chain.
* 2048x2048 precision form (8.4M triangles, 4.2M nodes): mixed forward +
  adjoint at rtol 1e-10 checked through size-independent properties -- the
  closed-form post-BC nnz, symmetry, the true FP64 residuals of u and lambda
  recomputed on the host with SciPy from the assembled matrix, agreement
  with the plain FP64 path, and central differences of the loss.
"""
import os

import numpy as np
import pytest
import torch

import admfem
from admfem import meshgen
from oracle import oracle as O

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def readme(m):
    c, e = meshgen.rectangle(m, m, 1.0 / m)
    gm = admfem.Mesh(c, e, zero_based=True)
    gx, gy = admfem.gauss_nodes(gm).T
    return gm, np.sin(gx) * (1 + gy ** 2) + 1.0, 1e5 * (gx + gy)


def test_c1_m8_ops_chain_matches_golden():
    print("golden c1_m8 through the op chain", flush=True)
    d = load("c1_m8.npz")
    gm = admfem.Mesh(d["coords"], d["elems"], zero_based=True)
    np.testing.assert_array_equal(gm.elems - 1, d["elems_oriented"])
    gx, gy = admfem.gauss_nodes(gm).T
    np.testing.assert_allclose(gx, d["gx"], rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(admfem.gauss_weights(gm), d["w"], rtol=1e-15)
    k = torch.tensor(d["kappa"], device="cuda", requires_grad=True)
    K = admfem.compute_fem_laplace_matrix1(k, gm)
    np.testing.assert_array_equal(host(K.indices), d["K_indices"])
    np.testing.assert_allclose(host(K.values.detach()), d["K_values"], rtol=1e-13, atol=1e-13)
    F = admfem.compute_fem_source_term1(torch.tensor(d["f"], device="cuda"), gm)
    np.testing.assert_allclose(host(F), d["F"], rtol=1e-14)
    B, rhs = admfem.impose_Dirichlet_boundary_conditions(K, F, d["bd"], np.zeros(d["bd"].size))
    np.testing.assert_array_equal(host(B.indices), d["D_indices"])
    u = admfem.solve(B, rhs, rtol=1e-14)
    assert rel(host(u.detach()), d["u"]) < 1e-10
    o = torch.as_tensor(d["obs"], device="cuda")
    (u[o] ** 2).sum().backward()
    assert rel(host(k.grad), d["grad_kappa"]) < 1e-10


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-8)])
def test_c1_m64_readme_case_matches_golden(prec, tol):
    print(f"golden c1_m64 {prec}", flush=True)
    d = load("c1_m64.npz")
    gm, kappa, f = readme(64)
    S = admfem.DirichletSystem(gm, d["bd"])
    out = S.forward_adjoint(d["obs"], np.zeros(d["obs"].size), precision=prec,
                            rtol=1e-13 if prec == "fp64" else 1e-11, kappa=kappa, f=f)
    A, _ = S.csr()
    np.testing.assert_array_equal(A.indptr, d["indptr"])   # pattern exact
    np.testing.assert_array_equal(A.indices, d["indices"])
    assert rel(host(out["u"]), d["u"]) < tol
    assert rel(host(out["grad_kappa"]), d["grad_kappa"]) < tol


def test_heat_step_matches_golden():
    h = load("heat_m6.npz")
    gm = admfem.Mesh(h["coords"], h["elems"], zero_based=True)
    S = admfem.DirichletSystem(gm, h["bd"])
    dt = float(h["dt"])
    out = S.forward_adjoint(h["obs"], np.zeros(h["obs"].size), precision="fp64", rtol=1e-14,
                            kappa=h["kappa"], f=h["f"], u_prev=h["u_prev"], lap_coef=dt,
                            mass_coef=1.0, src_coef=dt)
    _, r = S.csr()
    np.testing.assert_allclose(r, h["A_rhs"], rtol=1e-13, atol=1e-15)
    assert rel(host(out["u"]), h["u"]) < 1e-10
    assert rel(host(out["grad_kappa"]), h["grad_kappa"]) < 1e-10


def jittered_problem(m, seeds):
    c, e = meshgen.jittered_square(m, seed=seeds[0])
    gm = admfem.Mesh(c, e, zero_based=True)
    g = admfem.gauss_nodes(gm)
    kappa = np.exp(meshgen.grf(g, seeds[1]))
    f = np.random.default_rng(seeds[2]).uniform(-1, 1, gm.ngauss)
    bd = admfem.bcnode(gm)
    interior = np.setdiff1d(np.arange(gm.nnode), bd - 1)
    obs = np.sort(np.random.default_rng(seeds[3]).choice(interior, size=gm.nnode // 100,
                                                         replace=False))
    return c, e, gm, kappa, f, bd, obs


def test_c2_plain_fp64_m1024_matches_direct_solve():
    print("1024x1024 plain FP64 vs the oracle's direct solve", flush=True)
    c, e, gm, kappa, f, bd, obs = jittered_problem(1024, (21, 22, 23, 24))
    uobs = np.random.default_rng(25).standard_normal(obs.size) * 1e-3
    S = admfem.DirichletSystem(gm, bd)
    out = S.forward_adjoint(obs, uobs, precision="fp64", rtol=1e-12, kappa=kappa, f=f)
    print(out["stats"], flush=True)
    om = O.OracleMesh(c, e)

    def dl(u):
        g = np.zeros_like(u)
        g[obs] = 2 * (u[obs] - uobs)
        return g
    ref = O.forward_adjoint(om, kappa, f, bd, np.zeros(bd.size), dl)
    eu, eg = rel(host(out["u"]), ref["u"]), rel(host(out["grad_kappa"]), ref["grad_kappa"])
    print(f"rel err u {eu:.2e} grad {eg:.2e}", flush=True)
    assert eu < 1e-10 and eg < 1e-10


def test_c4_gradient_m512_finite_differences_and_oracle():
    print("512x512 gradient: central differences and oracle chain", flush=True)
    c, e, gm, kappa, f, bd, obs = jittered_problem(512, (31, 32, 33, 34))
    S = admfem.DirichletSystem(gm, bd)
    g2 = admfem.gauss_nodes(gm)
    k2 = np.exp(meshgen.grf(g2, 35))           # u_obs from a second kappa
    S.assemble(kappa=k2, f=f)
    u2, _ = S.solve(precision="fp64", rtol=1e-13)
    uobs = host(u2)[obs]
    out = S.forward_adjoint(obs, uobs, precision="mixed", rtol=1e-11, kappa=kappa, f=f,
                            want_lam=False)
    grad = host(out["grad_kappa"])
    assert grad.size == 3 * 2 * 512 * 512

    def loss(k):
        r = S.forward_adjoint(obs, uobs, precision="fp64", rtol=1e-14, kappa=k, f=f,
                              want_lam=False)
        return r["loss"]
    # Central differences in 20 seeded directions.  Their error is the
    # O(eps^2) truncation plus the loss's rounding divided by eps: with L
    # known to ~1e-12 relative (FP64 solves at rtol 1e-14) and a directional
    # derivative ~100x smaller than L, step 1e-6 resolves ~1e-4 and step 1e-4
    # (truncation ~1e-8) ~1e-6.
    rng = np.random.default_rng(36)
    worst = {1e-6: 0.0, 1e-4: 0.0}
    for _ in range(20):
        d = rng.standard_normal(kappa.size) * kappa
        an = float(grad @ d)
        for eps in worst:
            fd = (loss(kappa + eps * d) - loss(kappa - eps * d)) / (2 * eps)
            worst[eps] = max(worst[eps], abs(fd - an) / abs(an))
    print(f"worst relative FD mismatch {worst}", flush=True)
    assert worst[1e-6] < 3e-4 and worst[1e-4] < 3e-6
    # Taylor remainder (the reference's gradtest.jl pattern): second order
    d = rng.standard_normal(kappa.size) * kappa
    L0, an = loss(kappa), float(grad @ d)
    rem = [abs(loss(kappa + g * d) - L0 - g * an) for g in (1e-1, 1e-2, 1e-3)]
    print(f"Taylor remainders {rem}", flush=True)
    assert rem[1] < rem[0] / 50 and rem[2] < rem[1] / 50
    om = O.OracleMesh(c, e)

    def dl(u):
        g = np.zeros_like(u)
        g[obs] = 2 * (u[obs] - uobs)
        return g
    ref = O.forward_adjoint(om, kappa, f, bd, np.zeros(bd.size), dl)
    eg = rel(grad, ref["grad_kappa"])
    print(f"rel err vs oracle chain {eg:.2e}", flush=True)
    assert eg < 1e-8


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-8)])
def test_twoholes_poisson_example_matches_golden(prec, tol, tmp_path):
    """docs/src/codes/Poisson/inverse.jl's forward problem on the reference's
    twoholes_large.stl (an unstructured mesh with two holes), read through
    Mesh(filename) from an STL written from the fixture."""
    d = load("twoholes_large.npz")
    c, e = d["coords"], d["elems"]
    p = tmp_path / "twoholes.stl"
    with open(p, "w") as fh:
        fh.write("solid twoholes\n")
        for t in e:
            fh.write("facet normal 0 0 1\n outer loop\n")
            fh.write("".join(f"  vertex {float(c[v, 0])!r} {float(c[v, 1])!r} 0\n" for v in t))
            fh.write(" endloop\nendfacet\n")
        fh.write("endsolid twoholes\n")
    gm = admfem.Mesh(str(p))
    np.testing.assert_array_equal(gm.nodes, c)
    gx, gy = admfem.gauss_nodes(gm).T
    kappa, f = np.sin(gx) * (1 + gy ** 2) + 1.0, 1e5 * (gx + gy)
    bd = admfem.bcnode(gm)
    np.testing.assert_array_equal(np.sort(bd), np.sort(d["bd"]))
    S = admfem.DirichletSystem(gm, bd)
    out = S.forward_adjoint(np.arange(gm.nnode), np.zeros(gm.nnode), precision=prec,
                            rtol=1e-13 if prec == "fp64" else 1e-11, kappa=kappa, f=f)
    assert rel(host(out["u"]), d["u"]) < tol
    assert rel(host(out["grad_kappa"]), d["grad_kappa"]) < tol


def grf_device(points, seed, nmodes=32, sigma=1.0, ell=0.1):
    """meshgen.grf's field (same modes from the same seed) evaluated with
    torch on the GPU: 25M Gauss points x 32 modes is slow in numpy."""
    rng = np.random.default_rng(seed)
    w = torch.as_tensor(rng.normal(0.0, 1.0 / ell, size=(nmodes, 2)), device="cuda")
    phi = torch.as_tensor(rng.uniform(0.0, 2 * np.pi, size=nmodes), device="cuda")
    pts = torch.as_tensor(points, device="cuda")
    g = torch.zeros(pts.shape[0], dtype=torch.float64, device="cuda")
    for k in range(nmodes):
        g += torch.cos(pts @ w[k] + phi[k])
    return sigma * np.sqrt(2.0 / nmodes) * g


def post_bc_nnz(m):
    """Entries of the post-Dirichlet matrix of the m x m square (interior
    pairs, explicit zeros kept, one diagonal per boundary node)."""
    return (m - 1) ** 2 + 2 * (2 * (m - 1) * (m - 2) + (m - 2) ** 2) + 4 * m


def test_c3_precision_form_m2048():
    import time
    t0 = time.perf_counter()
    m = 2048
    print(f"{m}x{m} precision form: set-up", flush=True)
    c, e = meshgen.jittered_square(m, seed=41)
    gm = admfem.Mesh(c, e, zero_based=True)
    g = admfem.gauss_nodes(gm)
    kappa = torch.exp(grf_device(g, 42))
    kappa_obs = torch.exp(grf_device(g, 43))
    f = torch.as_tensor(np.random.default_rng(44).uniform(-1, 1, gm.ngauss), device="cuda")
    bd = admfem.bcnode(gm)
    assert bd.size == 4 * m
    interior = np.setdiff1d(np.arange(gm.nnode), bd - 1)
    obs = np.sort(np.random.default_rng(45).choice(interior, size=gm.nnode // 100,
                                                   replace=False))
    S = admfem.DirichletSystem(gm, bd)
    assert S.n == (m + 1) ** 2 and S.nnz == post_bc_nnz(m) == 29_323_281
    S.assemble(kappa=kappa_obs, f=f)
    u2, _ = S.solve(precision="fp64", rtol=1e-12)
    uobs = host(u2)[obs]
    print(f"set-up {time.perf_counter() - t0:.1f}s", flush=True)
    out = {}
    for prec in ("mixed", "fp64"):
        out[prec] = S.forward_adjoint(obs, uobs, precision=prec, rtol=1e-10, kappa=kappa, f=f)
        print(prec, out[prec]["stats"], flush=True)
    # the assembled matrix (user order), its symmetry, and the true FP64
    # residuals of the mixed solution and adjoint
    A, rhs = S.csr()
    assert A.nnz == S.nnz
    asym = abs(A - A.T).max()
    assert asym <= 1e-15 * abs(A).max(), asym
    u, lam = host(out["mixed"]["u"]), host(out["mixed"]["lam"])
    absA = abs(A)

    def residual(x, b):
        # ||b - A x|| / ||b|| and the rounding bound of evaluating it in FP64
        # (8 eps |A||x| per row: at most 7 entries per row plus the subtraction)
        r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        return r, 8 * 2.0 ** -53 * np.linalg.norm(absA @ abs(x)) / np.linalg.norm(b)
    ru, nu = residual(u, rhs)
    gu = np.zeros(S.n)
    gu[obs] = 2 * (u[obs] - uobs)
    rl, nl = residual(lam, gu)
    print(f"true residuals: u {ru:.2e} (rounding {nu:.1e}), lambda {rl:.2e} "
          f"(rounding {nl:.1e})", flush=True)
    assert ru <= 1e-10 + nu and rl <= 1e-10 + nl
    assert abs(out["mixed"]["loss"] - float(((u[obs] - uobs) ** 2).sum())) <= \
        1e-12 * out["mixed"]["loss"]
    # the mixed and the plain FP64 path agree
    eu = rel(u, host(out["fp64"]["u"]))
    grad = host(out["mixed"]["grad_kappa"])
    eg = rel(grad, host(out["fp64"]["grad_kappa"]))
    print(f"mixed vs fp64: u {eu:.2e}, grad {eg:.2e}", flush=True)
    assert eu <= 1e-8 and eg <= 1e-8

    def loss(k):
        return S.forward_adjoint(obs, uobs, precision="fp64", rtol=1e-13, kappa=k, f=f,
                                 want_lam=False)["loss"]
    # central differences of L in 3 seeded directions (relative perturbations
    # of kappa at every Gauss point), steps 1e-3 and 1e-4
    rng = np.random.default_rng(46)
    worst = {1e-3: 0.0, 1e-4: 0.0}
    for _ in range(3):
        d = torch.as_tensor(rng.standard_normal(gm.ngauss), device="cuda") * kappa
        an = float(torch.dot(out["mixed"]["grad_kappa"], d))
        for eps in worst:
            fd = (loss(kappa + eps * d) - loss(kappa - eps * d)) / (2 * eps)
            worst[eps] = max(worst[eps], abs(fd - an) / abs(an))
    print(f"worst relative FD mismatch {worst}; total {time.perf_counter() - t0:.1f}s",
          flush=True)
    assert worst[1e-3] < 1e-5 and worst[1e-4] < 1e-5
