"""The reference's operator chain through torch autograd on the GPU:
compute_fem_laplace_matrix1 -> compute_fem_source_term1 ->
impose_Dirichlet_boundary_conditions -> A \\ F -> loss.backward(), against the
oracle's composition of the reference's forward and Grad ops."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import torch

import admfem
from admfem import _lib, meshgen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def setup(m, seed):
    c, e = meshgen.jittered_square(m, seed=seed)
    gm = admfem.Mesh(c, e, zero_based=True)
    om = O.OracleMesh(c, e)
    rng = np.random.default_rng(seed)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), seed + 1))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    interior = np.setdiff1d(np.arange(om.nnode), bd - 1)
    obs = rng.choice(interior, size=max(3, interior.size // 10), replace=False)
    uobs = rng.standard_normal(obs.size) * 1e-2
    return gm, om, kap, f, bd, obs, uobs


def dloss(obs, uobs):
    def g(u):
        out = np.zeros_like(u)
        out[obs] = 2 * (u[obs] - uobs)
        return out
    return g


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_poisson_chain_autograd_matches_oracle():
    gm, om, kap, f, bd, obs, uobs = setup(14, 3)
    k = torch.tensor(kap, device="cuda", requires_grad=True)
    fk = torch.tensor(f, device="cuda")
    K = admfem.compute_fem_laplace_matrix1(k, gm)
    F = admfem.compute_fem_source_term1(fk, gm)
    Kb, Fb = admfem.impose_Dirichlet_boundary_conditions(K, F, bd, np.zeros(bd.size))
    u = admfem.solve(Kb, Fb, rtol=1e-13)
    o = torch.as_tensor(obs, device="cuda")
    loss = ((u[o] - torch.as_tensor(uobs, device="cuda")) ** 2).sum()
    loss.backward()
    ref = O.forward_adjoint(om, kap, f, bd, np.zeros(bd.size), dloss(obs, uobs))
    assert rel(u.detach().cpu().numpy(), ref["u"]) < 1e-10
    assert rel(k.grad.cpu().numpy(), ref["grad_kappa"]) < 1e-9


def test_heat_chain_autograd_matches_oracle():
    gm, om, kap, f, bd, obs, uobs = setup(11, 5)
    dt = 0.02
    up = np.sin(np.pi * om.coords[:, 0]) * np.sin(np.pi * om.coords[:, 1])
    k = torch.tensor(kap, device="cuda", requires_grad=True)
    M = admfem.compute_fem_mass_matrix1(gm)
    K = admfem.compute_fem_laplace_matrix1(k, gm)
    A = M + K * dt
    rhs = M @ torch.as_tensor(up, device="cuda") + dt * admfem.compute_fem_source_term1(
        torch.as_tensor(f, device="cuda"), gm)
    Ab, rb = admfem.impose_Dirichlet_boundary_conditions(A, rhs, bd, np.zeros(bd.size))
    u = admfem.solve(Ab, rb, rtol=1e-13)
    o = torch.as_tensor(obs, device="cuda")
    ((u[o] - torch.as_tensor(uobs, device="cuda")) ** 2).sum().backward()
    ref = O.forward_adjoint(om, kap, f, bd, np.zeros(bd.size), dloss(obs, uobs), dt=dt,
                            u_prev=up)
    assert rel(u.detach().cpu().numpy(), ref["u"]) < 1e-10
    assert rel(k.grad.cpu().numpy(), ref["grad_kappa"]) < 1e-9


def test_array_paths_and_unit_coefficient_helpers():
    gm, om, kap, f, bd, obs, uobs = setup(6, 1)
    K = admfem.compute_fem_laplace_matrix1(kap, gm)  # numpy -> scipy (MCore.jl:112-120)
    K0 = O.csr_matrix(*om.laplace_coo(kap), om.nnode)
    assert abs(K - K0).max() < 1e-13 * abs(K0).max()
    K1 = admfem.compute_fem_laplace_matrix1(gm)      # kappa = 1, Array path (MCore.jl:137)
    assert abs(K1 - O.csr_matrix(*om.laplace_coo(np.ones(om.ngauss)), om.nnode)).max() < 1e-13
    F = admfem.compute_fem_source_term1(f, gm)
    np.testing.assert_allclose(F, om.source(f), rtol=1e-14, atol=1e-17)
    # SparseMatrixCSC form (MUtils.jl:184-198): [A_II 0; 0 I], [r_I - A_IB u_B; u_B]
    bv = np.linspace(0.5, 1.5, bd.size)
    Bs, rs = admfem.impose_Dirichlet_boundary_conditions(K, F, bd, bv)
    keep = np.ones(om.nnode, bool)
    keep[bd - 1] = False
    Kd = K.toarray()
    Bref = np.zeros_like(Kd)
    Bref[np.ix_(keep, keep)] = Kd[np.ix_(keep, keep)]
    Bref[bd - 1, bd - 1] = 1.0
    rref = F.copy()
    rref[keep] -= Kd[np.ix_(keep, bd - 1)] @ bv
    rref[bd - 1] = bv
    assert np.abs(Bs.toarray() - Bref).max() < 1e-14 * np.abs(Kd).max()
    np.testing.assert_allclose(rs, rref, rtol=1e-13, atol=1e-13 * np.abs(rref).max())
    K1t = admfem.compute_fem_laplace_matrix1(torch.ones(om.ngauss, dtype=torch.float64), gm)
    ki, kv = K1t.indices.cpu().numpy(), K1t.values.detach().cpu().numpy()
    _, kv0 = om.laplace_coo(np.ones(om.ngauss))
    np.testing.assert_allclose(kv, kv0, rtol=1e-14, atol=1e-14 * np.abs(kv0).max())
    B = admfem.impose_Dirichlet_boundary_conditions(K1t, bd)  # homogeneous helper
    oi, ov, _ = O.dirichlet(ki, kv, bd, np.zeros(om.nnode), np.zeros(bd.size))
    np.testing.assert_array_equal(B.indices.cpu().numpy(), oi)
    np.testing.assert_array_equal(B.values.detach().cpu().numpy(), ov)


def test_sparse_solve_op_and_gradient_general_spd():
    """SparseSolver / SparseSolverGrad on a random SPD matrix with duplicate
    triplets, against SciPy and the closed-form gradient."""
    rng = np.random.default_rng(0)
    n = 60
    R = sp.random(n, n, density=0.08, random_state=1)
    A = (R @ R.T + sp.eye(n) * 2.0).tocoo()
    I = np.concatenate([A.row, A.row[:20]])
    J = np.concatenate([A.col, A.col[:20]])
    V = np.concatenate([A.data, np.zeros(20)])  # explicit duplicate zeros
    b = rng.standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    vv = torch.tensor(V, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    u = admfem.solve(admfem.SparseTensor(ind, vv, n), bt, rtol=1e-14)
    w = torch.as_tensor(rng.standard_normal(n), device="cuda")
    (u * w).sum().backward()
    u0 = spl.spsolve(A.tocsc(), b)
    lam = spl.spsolve(A.T.tocsc(), w.cpu().numpy())
    assert rel(u.detach().cpu().numpy(), u0) < 1e-10
    assert rel(bt.grad.cpu().numpy(), lam) < 1e-10
    assert rel(vv.grad.cpu().numpy(), -lam[I] * u0[J]) < 1e-10


@pytest.mark.parametrize("shuffle", [False, True])
def test_sparse_solve_keeps_banded_post_dirichlet_order(shuffle):
    """The reference's main use of the op: the Laplace COO of a row-major
    rectangle mesh after Dirichlet elimination, whose first row is a corner
    boundary node holding only its diagonal.  That order is banded, so the
    solve keeps it; the same mesh with its nodes numbered at random is
    solved in reverse Cuthill-McKee order.  Both against the oracle."""
    m = 40
    c, e = meshgen.jittered_square(m, seed=2, jitter=0.0, permute=shuffle)
    gm = admfem.Mesh(c, e, zero_based=True)
    om = O.OracleMesh(c, e)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 4))
    f = np.random.default_rng(5).uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    K = admfem.compute_fem_laplace_matrix1(torch.tensor(kap, device="cuda"), gm)
    F = admfem.compute_fem_source_term1(torch.tensor(f, device="cuda"), gm)
    Kb, Fb = admfem.impose_Dirichlet_boundary_conditions(K, F, bd, np.zeros(bd.size))
    u = admfem.solve(Kb, Fb, rtol=1e-13)
    info = admfem.ops.solve_pattern_info()
    assert info["n"] == om.nnode and info["reordered"] == shuffle
    ref = O.forward_adjoint(om, kap, f, bd, np.zeros(bd.size), lambda u: np.zeros_like(u))
    assert rel(u.cpu().numpy(), ref["u"]) < 1e-10


def test_sparse_solve_far_coupling():
    """A chain closed into a cycle by one coupling 69,999 rows away (the solver
    renumbers it by reverse Cuthill-McKee, bandwidth 2) against SciPy."""
    n = 70000
    main = np.full(n, 4.0)
    off = np.full(n - 1, -1.0)
    A = sp.diags([off, main, off], [-1, 0, 1], format="coo")
    I = np.concatenate([A.row, [0, n - 1]])
    J = np.concatenate([A.col, [n - 1, 0]])
    V = np.concatenate([A.data, [-0.5, -0.5]])
    b = np.random.default_rng(3).standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    u = admfem.solve(admfem.SparseTensor(ind, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), rtol=1e-13)
    ref = spl.spsolve(sp.csr_matrix((V, (I, J)), shape=(n, n)), b)
    assert np.abs(u.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-10


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_sparse_solve_star_wider_than_16_bit_offsets(precision):
    """A star (node 0 coupled to all 69,999 others): every numbering leaves
    column offsets beyond 16 bits, so the FP32 steps read the 32-bit
    columns, and row 0 is far wider than the straight-line widths (its
    entries past the eighth take the loop); against SciPy."""
    n = 70000
    rng = np.random.default_rng(5)
    main = np.full(n, 2.0)
    main[0] = float(n)
    w = -rng.random(n - 1) * 0.5
    I = np.concatenate([np.arange(n), np.zeros(n - 1, np.int64), np.arange(1, n)])
    J = np.concatenate([np.arange(n), np.arange(1, n), np.zeros(n - 1, np.int64)])
    V = np.concatenate([main, w, w])
    b = rng.standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    u = admfem.solve(admfem.SparseTensor(ind, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), precision=precision, rtol=1e-12)
    ref = spl.spsolve(sp.csr_matrix((V, (I, J)), shape=(n, n)), b)
    assert np.abs(u.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-9


@pytest.mark.parametrize("permute", [False, True])
def test_sparse_solve_grid_laplacian_any_numbering(permute):
    """The shifted 5-point Laplacian of a 150 x 150 grid, row-major (banded:
    solved in the caller's order) or with its unknowns numbered at random
    (solved in reverse Cuthill-McKee order, the solution scattered back),
    against SciPy, with the gradient of the op."""
    m = 150
    n = m * m
    T = sp.diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    A = (sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m)) + 0.05 * sp.eye(n)).tocoo()
    I, J, V = A.row.astype(np.int64), A.col.astype(np.int64), A.data
    rng = np.random.default_rng(11)
    if permute:
        p = rng.permutation(n)
        I, J = p[I], p[J]
    b = rng.standard_normal(n)
    w = rng.standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    vv = torch.tensor(V, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    u = admfem.solve(admfem.SparseTensor(ind, vv, n), bt, rtol=1e-13)
    (u * torch.as_tensor(w, device="cuda")).sum().backward()
    As = sp.csc_matrix((V, (I, J)), shape=(n, n))
    u0 = spl.spsolve(As, b)
    lam = spl.spsolve(As.T.tocsc(), w)
    assert rel(u.detach().cpu().numpy(), u0) < 1e-10
    assert rel(bt.grad.cpu().numpy(), lam) < 1e-10
    assert rel(vv.grad.cpu().numpy(), -lam[I] * u0[J]) < 1e-10
    assert admfem.ops.solve_pattern_info()["reordered"] == permute


def test_sparse_solve_several_components_random_numbering():
    """Two disjoint shifted grid Laplacians (80 x 80 and 60 x 60) and 50
    decoupled unknowns, numbered at random: the graph has two components
    besides the isolated nodes, so the reverse Cuthill-McKee order comes from
    the host; against SciPy."""
    blocks = []
    for m in (80, 60):
        T = sp.diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
        blocks.append(sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m)) + 0.05 * sp.eye(m * m))
    blocks.append(sp.eye(50) * 3.0)
    A = sp.block_diag(blocks).tocoo()
    n = A.shape[0]
    rng = np.random.default_rng(13)
    p = rng.permutation(n)
    I, J, V = p[A.row].astype(np.int64), p[A.col].astype(np.int64), A.data
    b = rng.standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    u = admfem.solve(admfem.SparseTensor(ind, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), rtol=1e-13)
    u0 = spl.spsolve(sp.csc_matrix((V, (I, J)), shape=(n, n)), b)
    assert rel(u.cpu().numpy(), u0) < 1e-10
    assert admfem.ops.solve_pattern_info()["reordered"]


def test_sparse_solve_one_sided_explicit_zeros_random_numbering():
    """A shifted grid Laplacian numbered at random plus explicit zeros on one
    side of the diagonal only: the pattern is not symmetric, so the device
    Cuthill-McKee pass must notice and leave the order to the host; the
    values (and the solution) are the SPD matrix's."""
    m = 90
    n = m * m
    T = sp.diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    A = (sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m)) + 0.05 * sp.eye(n)).tocoo()
    rng = np.random.default_rng(17)
    p = rng.permutation(n)
    zi = rng.integers(0, n, 300)
    zj = (zi + rng.integers(m + 2, 5 * m, 300)) % n   # no partner entry (j, i)
    I = np.concatenate([p[A.row], zi]).astype(np.int64)
    J = np.concatenate([p[A.col], zj]).astype(np.int64)
    V = np.concatenate([A.data, np.zeros(zi.size)])
    b = rng.standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    u = admfem.solve(admfem.SparseTensor(ind, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), rtol=1e-13)
    u0 = spl.spsolve(sp.csc_matrix((V, (I, J)), shape=(n, n)), b)
    assert rel(u.cpu().numpy(), u0) < 1e-10


def test_sparse_solve_pattern_kept_between_calls():
    """The op keeps the last pattern: the same indices with new values, then
    indices of the same sizes but another pattern, then an out-of-range
    index (an error that leaves the op usable) -- each against SciPy."""
    rng = np.random.default_rng(21)
    n = 300

    def spd(seed):
        R = sp.random(n, n, density=0.02, random_state=seed)
        A = (R @ R.T + sp.eye(n) * 3.0).tocoo()
        return A.row.astype(np.int64), A.col.astype(np.int64), A.data

    I, J, V = spd(1)
    b = rng.standard_normal(n)
    ind = torch.as_tensor(np.stack([I, J], 1), device="cuda")
    for scale in (1.0, 2.5):  # same pattern, new values
        u = admfem.solve(admfem.SparseTensor(ind, torch.tensor(scale * V, device="cuda"), n),
                         torch.tensor(b, device="cuda"), rtol=1e-14)
        ref = spl.spsolve(sp.csc_matrix((scale * V, (I, J)), shape=(n, n)), b)
        assert rel(u.cpu().numpy(), ref) < 1e-10
    # another pattern with the same n and nnz: a permuted copy of the matrix
    p = rng.permutation(n)
    I2, J2 = p[I], p[J]
    ind2 = torch.as_tensor(np.stack([I2, J2], 1), device="cuda")
    u = admfem.solve(admfem.SparseTensor(ind2, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), rtol=1e-14)
    ref = spl.spsolve(sp.csc_matrix((V, (I2, J2)), shape=(n, n)), b)
    assert rel(u.cpu().numpy(), ref) < 1e-10
    bad = ind2.clone()
    bad[5, 1] = n
    with pytest.raises(_lib.AdmfemError) as ei:
        admfem.solve(admfem.SparseTensor(bad, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), rtol=1e-14)
    assert ei.value.code == _lib.ERR_ARG
    u = admfem.solve(admfem.SparseTensor(ind, torch.tensor(V, device="cuda"), n),
                     torch.tensor(b, device="cuda"), rtol=1e-14)
    ref = spl.spsolve(sp.csc_matrix((V, (I, J)), shape=(n, n)), b)
    assert rel(u.cpu().numpy(), ref) < 1e-10
