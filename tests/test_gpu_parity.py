"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on
the same seeded inputs.

Bars: integer/index outputs bit-exact; FP64 operator values <= 1e-13
relative (the same formulas, FMA contraction aside); FP64 solves and
gradients <= 1e-10 relative to the oracle's direct solve; the mixed-precision
path <= 1e-8.
"""
import ctypes

import numpy as np
import pytest
import torch

import admfem
from admfem import _lib, meshgen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a)).to(device="cuda", dtype=dtype).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def pair(m, seed=0, jittered=True):
    if jittered:
        c, e = meshgen.jittered_square(m, seed=seed)
    else:
        c, e = meshgen.rectangle(m, m, 1.0 / m)
    gm = admfem.Mesh(c, e, zero_based=True)
    om = O.OracleMesh(c, e)
    return gm, om


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ---------------------------------------------------------------- mesh
# m = 260 (67,861 nodes) is split over several host threads by the set-up
@pytest.mark.parametrize("m,jit", [(2, False), (5, True), (17, True), (260, True)])
def test_mesh_getters_match_oracle(m, jit):
    print(f"mesh getters m={m} jittered={jit}", flush=True)
    gm, om = pair(m, seed=m, jittered=jit)
    np.testing.assert_array_equal(gm.elems - 1, om.elems)
    assert admfem.get_ngauss(gm) == om.ngauss
    g = admfem.gauss_nodes(gm)
    np.testing.assert_allclose(g[:, 0], om.gx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g[:, 1], om.gy, rtol=0, atol=1e-15)
    np.testing.assert_allclose(admfem.gauss_weights(gm), om.w, rtol=1e-14)
    np.testing.assert_allclose(admfem.get_area(gm), om.area, rtol=1e-14)
    np.testing.assert_array_equal(gm.edges, om.edges())
    np.testing.assert_array_equal(admfem.bcnode(gm), om.bcnode())


def test_mesh_2x2_known_answers_gpu():
    gm = admfem.Mesh(2, 2, 0.5)
    assert admfem.get_ngauss(gm) == 24
    np.testing.assert_allclose(admfem.get_area(gm), np.full(8, 0.125), atol=1e-15)


@pytest.mark.parametrize("m", [9, 260])
def test_pattern_matches_oracle_full_and_post_bc(m):
    print(f"pattern m={m}", flush=True)
    gm, om = pair(m, seed=3)
    I, V = om.laplace_coo(np.ones(om.ngauss))
    rp, col, _ = O.coo_to_csr(I, V, om.nnode)
    grp, gcol = gm.pattern()
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gcol, col)


# ---------------------------------------------------------------- COO ops
def test_laplace_and_mass_coo_match_oracle():
    gm, om = pair(11, seed=4)
    L = _lib.lib()
    rng = np.random.default_rng(1)
    kap = rng.random(om.ngauss) + 0.5
    nt = 9 * om.ngauss
    ind = torch.empty((nt, 2), dtype=torch.int64, device="cuda")
    vv = torch.empty(nt, dtype=torch.float64, device="cuda")
    dk = dev(kap)
    for fwd, ofwd in ((L.admfem_fem_laplace_scalar, om.laplace_coo),
                      (L.admfem_fem_mass_matrix, om.mass_coo)):
        _lib.check(fwd(gm.handle, _lib.ptr(dk), _lib.ptr(ind), _lib.ptr(vv), _lib.stream_ptr()))
        oi, ov = ofwd(kap)
        np.testing.assert_array_equal(host(ind), oi)
        assert relerr(host(vv), ov) < 1e-13
    # gradient ops
    gvv = rng.standard_normal(nt)
    dgvv = dev(gvv)  # keep the tensor alive while the kernel may read it
    out = torch.empty(om.ngauss, dtype=torch.float64, device="cuda")
    _lib.check(L.admfem_fem_laplace_scalar_grad(gm.handle, _lib.ptr(dgvv), _lib.ptr(out),
                                                _lib.stream_ptr()))
    assert relerr(host(out), om.laplace_backward(gvv)) < 1e-12
    _lib.check(L.admfem_fem_mass_matrix_grad(gm.handle, _lib.ptr(dgvv), _lib.ptr(out),
                                             _lib.stream_ptr()))
    assert relerr(host(out), om.mass_backward(gvv)) < 1e-12


def test_source_and_grad_match_oracle_bitwise_order():
    gm, om = pair(13, seed=5)
    L = _lib.lib()
    f = np.random.default_rng(2).uniform(-1, 1, om.ngauss)
    rhs = torch.empty(om.nnode, dtype=torch.float64, device="cuda")
    df = dev(f)
    _lib.check(L.admfem_fem_source_scalar(gm.handle, _lib.ptr(df), _lib.ptr(rhs),
                                          _lib.stream_ptr()))
    np.testing.assert_allclose(host(rhs), om.source(f), rtol=1e-14, atol=1e-17)
    g = np.random.default_rng(3).standard_normal(om.nnode)
    gf = torch.empty(om.ngauss, dtype=torch.float64, device="cuda")
    dg = dev(g)
    _lib.check(L.admfem_fem_source_scalar_grad(gm.handle, _lib.ptr(dg), _lib.ptr(gf),
                                               _lib.stream_ptr()))
    np.testing.assert_allclose(host(gf), om.source_backward(g), rtol=1e-14, atol=1e-17)


def test_legacy_reference_symbols_on_default_mesh():
    """init_nnfem_mesh / mfem_get_* / *_forward_Julia with host pointers, as
    the reference's Julia ccalls use them (src/MFEM/MFEM.jl:93-106)."""
    import ctypes
    L = _lib.lib()
    c, e = meshgen.rectangle(3, 3, 1.0 / 3)
    om = O.OracleMesh(c, e)
    verts = np.ascontiguousarray(np.concatenate([c, np.zeros((c.shape[0], 1))], 1).ravel())
    el = np.ascontiguousarray(e.astype(np.int32).ravel())
    ne = np.zeros(1, dtype=np.int64)
    p = L.init_nnfem_mesh(_lib.ptr(verts), c.shape[0], _lib.ptr(el), e.shape[0], 2, 6, 1,
                          _lib.ptr(ne))
    assert bool(p)
    edges = np.ctypeslib.as_array(p, shape=(2 * int(ne[0]),)).copy()
    ctypes.CDLL(None).free(ctypes.cast(p, ctypes.c_void_p))
    np.testing.assert_array_equal(edges.reshape(2, -1).T, om.edges())
    assert L.mfem_get_ngauss() == om.ngauss and L.mfem_get_ndof() == om.nnode
    assert L.mfem_get_elem_ndof() == 3
    x, y, w = np.zeros(om.ngauss), np.zeros(om.ngauss), np.zeros(om.ngauss)
    L.mfem_get_gauss(_lib.ptr(x), _lib.ptr(y))
    L.mfem_get_gauss_weights(_lib.ptr(w))
    np.testing.assert_allclose(x, om.gx, atol=1e-15)
    np.testing.assert_allclose(w, om.w, rtol=1e-14)
    conn = np.zeros(3 * om.nelem, dtype=np.int64)
    L.mfem_get_connectivity(_lib.ptr(conn))
    np.testing.assert_array_equal(conn.reshape(-1, 3), om.elems + 1)
    ev = np.zeros(3 * om.nelem, dtype=np.int64)
    L.mfem_get_element_to_vertices(_lib.ptr(ev))
    np.testing.assert_array_equal(ev.reshape(3, -1).T, om.elems + 1)
    kap = np.random.default_rng(0).random(om.ngauss)
    ind = np.zeros(2 * 9 * om.ngauss, dtype=np.int64)
    vv = np.zeros(9 * om.ngauss)
    L.FemLaplaceScalar_forward_Julia(_lib.ptr(ind), _lib.ptr(vv), _lib.ptr(kap))
    oi, ov = om.laplace_coo(kap)
    np.testing.assert_array_equal(ind.reshape(-1, 2), oi)
    assert relerr(vv, ov) < 1e-13
    rhs = np.ones(om.nnode)  # accumulates
    L.FemSourceScalar_forward_Julia(_lib.ptr(rhs), _lib.ptr(kap))
    np.testing.assert_allclose(rhs, 1.0 + om.source(kap), rtol=1e-14)
    # src/pcl.jl:35-39: zeros(ngauss, 9 ngauss), column-major, filled in place
    H = np.zeros(om.ngauss * 9 * om.ngauss)
    L.pcl_FemLaplaceScalar_Jacobian(_lib.ptr(H))
    Href = O.pcl_laplace_jacobian(om)
    np.testing.assert_allclose(H.reshape(9 * om.ngauss, om.ngauss).T, Href, rtol=1e-13,
                               atol=1e-16)


# ---------------------------------------------------------------- Dirichlet COO
@pytest.mark.parametrize("dup", [False, True])
def test_impose_dirichlet_coo_matches_oracle(dup):
    gm, om = pair(8, seed=6)
    L = _lib.lib()
    rng = np.random.default_rng(4)
    I, V = om.laplace_coo(rng.random(om.ngauss) + 0.2)
    bd = om.bcnode()
    if dup:
        bd = np.concatenate([bd, bd[::3]])
    bdval = rng.standard_normal(bd.size)
    rhs = rng.standard_normal(om.nnode)
    oi, ov, orhs = O.dirichlet(I, V, bd, rhs, bdval)
    sN, N = V.size, om.nnode
    cap = sN + bd.size
    goi = torch.zeros((cap, 2), dtype=torch.int64, device="cuda")
    gov = torch.zeros(cap, dtype=torch.float64, device="cuda")
    gorhs = torch.zeros(N, dtype=torch.float64, device="cuda")
    cnt = np.zeros(1, dtype=np.int64)
    dI, dV, dbd, dbv, dr = dev(I, torch.int64), dev(V), dev(bd, torch.int64), dev(bdval), dev(rhs)
    _lib.check(L.admfem_impose_dirichlet(_lib.ptr(dI), _lib.ptr(dV), sN, _lib.ptr(dbd),
                                         _lib.ptr(dbv), bd.size, _lib.ptr(dr), N, _lib.ptr(goi),
                                         _lib.ptr(gov), _lib.ptr(gorhs), _lib.ptr(cnt),
                                         _lib.stream_ptr()))
    n = int(cnt[0])
    assert n == oi.shape[0]
    np.testing.assert_array_equal(host(goi)[:n], oi)
    np.testing.assert_array_equal(host(gov)[:n], ov)
    np.testing.assert_array_equal(host(gorhs), orhs)  # same operation order: bitwise
    # gradient op
    g_ov, g_or = rng.standard_normal(n), rng.standard_normal(N)
    gv = torch.zeros(sN, dtype=torch.float64, device="cuda")
    gr = torch.zeros(N, dtype=torch.float64, device="cuda")
    gb = torch.zeros(bd.size, dtype=torch.float64, device="cuda")
    dgov, dgor = dev(g_ov), dev(g_or)
    _lib.check(L.admfem_impose_dirichlet_grad(_lib.ptr(dgov), _lib.ptr(dgor),
                                              _lib.ptr(dI), _lib.ptr(dV), sN, _lib.ptr(dbd),
                                              _lib.ptr(dbv), bd.size, N, _lib.ptr(gv),
                                              _lib.ptr(gr), _lib.ptr(gb), _lib.stream_ptr()))
    ogv, ogr, ogb = O.dirichlet_backward(g_ov, g_or, I, V, bd, bdval, N)
    np.testing.assert_array_equal(host(gv), ogv)
    np.testing.assert_array_equal(host(gr), ogr)
    np.testing.assert_array_equal(host(gb), ogb)


def _dirichlet_dev(I, V, bd, bdval, rhs, N, shift_in=False, shift_out=False):
    """The op through the C-ABI; shift_in / shift_out place the index arrays
    8 bytes off a 16-byte boundary (the kernels then take 8-byte loads and
    stores instead of 16-byte ones)."""
    L = _lib.lib()
    sN = V.size
    cap = sN + bd.size

    def ind_buf(n, src=None):
        flat = torch.zeros(2 * max(n, 1) + 2, dtype=torch.int64, device="cuda")
        off = 1 if (shift_in if src is not None else shift_out) else 2
        view = flat[off:off + 2 * n].view(n, 2) if n else flat[off:off]
        if src is not None and n:
            view.copy_(torch.as_tensor(src, dtype=torch.int64))
        return flat, view

    in_keep, dI = ind_buf(sN, I)
    out_keep, goi = ind_buf(cap)
    gov = torch.zeros(max(cap, 1), dtype=torch.float64, device="cuda")
    gorhs = torch.zeros(N, dtype=torch.float64, device="cuda")
    cnt = np.zeros(1, dtype=np.int64)
    dV, dbd, dbv, dr = dev(V), dev(bd, torch.int64), dev(bdval), dev(rhs)
    assert (dI.data_ptr() % 16 == 8) == (shift_in and sN > 0)
    _lib.check(L.admfem_impose_dirichlet(ctypes.c_void_p(dI.data_ptr()), _lib.ptr(dV), sN,
                                         _lib.ptr(dbd), _lib.ptr(dbv), bd.size, _lib.ptr(dr), N,
                                         ctypes.c_void_p(goi.data_ptr()), _lib.ptr(gov),
                                         _lib.ptr(gorhs), _lib.ptr(cnt), _lib.stream_ptr()))
    n = int(cnt[0])
    out = (host(goi.reshape(-1))[:2 * n].reshape(n, 2), host(gov)[:n], host(gorhs))
    del in_keep, out_keep
    return out, (dI, dV, dbd, dbv)


@pytest.mark.parametrize("shift_in,shift_out", [(False, False), (True, False), (False, True)])
def test_impose_dirichlet_many_tiles_shuffled(shift_in, shift_out):
    """Triplets in random order over many 2,048-triplet tiles (kept, lifted
    and dropped ones interleaved, the last tile partial), extra interior
    boundary nodes with repeats, and index arrays off 16-byte alignment:
    output, rhs and the gradient op bitwise against the oracle."""
    gm, om = pair(40, seed=11)
    rng = np.random.default_rng(12)
    I, V = om.laplace_coo(rng.random(om.ngauss) + 0.2)
    perm = rng.permutation(V.size)
    I, V = np.ascontiguousarray(I[perm]), np.ascontiguousarray(V[perm])
    extra = rng.choice(om.nnode, size=60, replace=False) + 1
    bd = np.concatenate([om.bcnode(), extra, extra[:7]])
    bdval = rng.standard_normal(bd.size)
    rhs = rng.standard_normal(om.nnode)
    N, sN = om.nnode, V.size
    assert sN > 20 * 2048 and sN % 2048
    (goi, gov, gorhs), (dI, dV, dbd, dbv) = _dirichlet_dev(I, V, bd, bdval, rhs, N,
                                                           shift_in, shift_out)
    oi, ov, orhs = O.dirichlet(I, V, bd, rhs, bdval)
    np.testing.assert_array_equal(goi, oi)
    np.testing.assert_array_equal(gov, ov)
    np.testing.assert_array_equal(gorhs, orhs)
    n = oi.shape[0]
    L = _lib.lib()
    g_ov, g_or = rng.standard_normal(n), rng.standard_normal(N)
    gv = torch.zeros(sN, dtype=torch.float64, device="cuda")
    gr = torch.zeros(N, dtype=torch.float64, device="cuda")
    gb = torch.zeros(bd.size, dtype=torch.float64, device="cuda")
    dgov, dgor = dev(g_ov), dev(g_or)
    _lib.check(L.admfem_impose_dirichlet_grad(_lib.ptr(dgov), _lib.ptr(dgor),
                                              ctypes.c_void_p(dI.data_ptr()), _lib.ptr(dV), sN,
                                              _lib.ptr(dbd), _lib.ptr(dbv), bd.size, N,
                                              _lib.ptr(gv), _lib.ptr(gr), _lib.ptr(gb),
                                              _lib.stream_ptr()))
    ogv, ogr, ogb = O.dirichlet_backward(g_ov, g_or, I, V, bd, bdval, N)
    np.testing.assert_array_equal(host(gv), ogv)
    np.testing.assert_array_equal(host(gr), ogr)
    np.testing.assert_array_equal(host(gb), ogb)


def test_impose_dirichlet_no_triplets():
    """An empty triplet list: only the boundary diagonals, rhs = bdval on
    the boundary and the input elsewhere."""
    rng = np.random.default_rng(3)
    N = 9
    bd = np.array([2, 5, 5, 9])
    bdval = rng.standard_normal(bd.size)
    rhs = rng.standard_normal(N)
    I = np.zeros((0, 2), dtype=np.int64)
    V = np.zeros(0)
    (goi, gov, gorhs), _ = _dirichlet_dev(I, V, bd, bdval, rhs, N)
    oi, ov, orhs = O.dirichlet(I, V, bd, rhs, bdval)
    np.testing.assert_array_equal(goi, oi)
    np.testing.assert_array_equal(gov, ov)
    np.testing.assert_array_equal(gorhs, orhs)
    assert goi.shape[0] == 3


def test_impose_dirichlet_rejects_out_of_range_index():
    L = _lib.lib()
    I = dev(np.array([[0, 0], [0, 5]]), torch.int64)
    V = dev(np.array([1.0, 2.0]))
    bd = dev(np.array([1]), torch.int64)
    bv = dev(np.array([0.0]))
    r = dev(np.zeros(2))
    out_i = torch.zeros((3, 2), dtype=torch.int64, device="cuda")
    out_v = torch.zeros(3, dtype=torch.float64, device="cuda")
    out_r = torch.zeros(2, dtype=torch.float64, device="cuda")
    cnt = np.zeros(1, dtype=np.int64)
    rc = L.admfem_impose_dirichlet(_lib.ptr(I), _lib.ptr(V), 2, _lib.ptr(bd), _lib.ptr(bv), 1,
                                   _lib.ptr(r), 2, _lib.ptr(out_i), _lib.ptr(out_v),
                                   _lib.ptr(out_r), _lib.ptr(cnt), _lib.stream_ptr())
    assert rc == _lib.ERR_ARG


# ---------------------------------------------------------------- fused system
def _oracle_system(om, kap, f, bd, bdval, dt=None, u_prev=None, rho=None):
    I, V = om.laplace_coo(kap)
    F = om.source(f)
    if dt is None:
        AI, AV, rhs = I, V, F
    else:
        MI, MV = om.mass_coo(rho if rho is not None else np.ones(om.ngauss))
        AI = np.concatenate([MI, I])
        AV = np.concatenate([MV, dt * V])
        rhs = O.csr_matrix(MI, MV, om.nnode) @ u_prev + dt * F
    oi, ov, orhs = O.dirichlet(AI, AV, bd, rhs, bdval)
    return O.csr_matrix(oi, ov, om.nnode), orhs


# heat: False (Poisson), "rho" (a mass density field: the per-element matrix
# layout), "unit" (rho = 1: the matrix rows rebuilt from the coordinates)
@pytest.mark.parametrize("m,heat", [(12, False), (12, "rho"), (260, "rho"), (12, "unit"),
                                    (260, "unit"), (260, False)])
def test_system_assembly_matches_oracle(m, heat):
    print(f"system assembly m={m} heat={heat}", flush=True)
    gm, om = pair(m, seed=7)
    rng = np.random.default_rng(5)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 1))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    bdval = rng.standard_normal(bd.size)
    S = admfem.DirichletSystem(gm, bd)
    if heat:
        dt, up = 0.05, rng.standard_normal(om.nnode)
        rho = rng.random(om.ngauss) + 0.5 if heat == "rho" else None
        S.assemble(kappa=kap, f=f, bdval=bdval, rho=rho, u_prev=up, lap_coef=dt, mass_coef=1.0,
                   src_coef=dt)
        A0, r0 = _oracle_system(om, kap, f, bd, bdval, dt=dt, u_prev=up, rho=rho)
    else:
        S.assemble(kappa=kap, f=f, bdval=bdval)
        A0, r0 = _oracle_system(om, kap, f, bd, bdval)
    A, r = S.csr()
    A0 = A0.tocsr()
    A0.sort_indices()
    np.testing.assert_array_equal(A.indptr, A0.indptr)    # pattern exact
    np.testing.assert_array_equal(A.indices, A0.indices)
    assert relerr(A.data, A0.data) < 1e-13
    assert relerr(r, r0) < 1e-13


@pytest.mark.parametrize("heat", [False, True])
def test_shuffled_elements_system_matches_oracle(heat):
    """Elements listed in random order: the element order is not banded, so
    the nodes are numbered by reverse Cuthill-McKee instead."""
    print(f"shuffled elements heat={heat}", flush=True)
    c, e = meshgen.jittered_square(40, seed=11)
    e = e[np.random.default_rng(4).permutation(e.shape[0])]
    gm, om = admfem.Mesh(c, e, zero_based=True), O.OracleMesh(c, e)
    rng = np.random.default_rng(8)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 4))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    bdval = rng.standard_normal(bd.size)
    S = admfem.DirichletSystem(gm, bd)
    if heat:
        up = rng.standard_normal(om.nnode)
        S.assemble(kappa=kap, f=f, bdval=bdval, u_prev=up, lap_coef=0.05, mass_coef=1.0,
                   src_coef=0.05)
        A0, r0 = _oracle_system(om, kap, f, bd, bdval, dt=0.05, u_prev=up)
    else:
        S.assemble(kappa=kap, f=f, bdval=bdval)
        A0, r0 = _oracle_system(om, kap, f, bd, bdval)
    A, r = S.csr()
    A0 = A0.tocsr()
    A0.sort_indices()
    np.testing.assert_array_equal(A.indptr, A0.indptr)
    np.testing.assert_array_equal(A.indices, A0.indices)
    assert relerr(A.data, A0.data) < 1e-13
    assert relerr(r, r0) < 1e-13
    u, st = S.solve(precision="fp64", rtol=1e-13)
    assert st["status"] == 0
    assert relerr(host(u), O.DirectSolver(A0).solve(r0)) < 1e-10


@pytest.mark.parametrize("m,prec,tol", [(16, "fp64", 1e-10), (40, "fp64", 1e-10),
                                        (40, "mixed", 1e-8)])
def test_solve_matches_direct(m, prec, tol):
    print(f"solve m={m} {prec}", flush=True)
    gm, om = pair(m, seed=8)
    rng = np.random.default_rng(6)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 2))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    S = admfem.DirichletSystem(gm, bd).assemble(kappa=kap, f=f)
    u, st = S.solve(precision=prec, rtol=1e-13 if prec == "fp64" else 1e-11)
    A0, r0 = _oracle_system(om, kap, f, bd, np.zeros(bd.size))
    u0 = O.DirectSolver(A0).solve(r0)
    print(st, flush=True)
    assert st["status"] == 0
    assert relerr(host(u), u0) < tol


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replayed_iterations_match_one_by_one(prec):
    """The solver replays groups of 8 iterations from a captured HIP graph;
    with check_every=1 every iteration is launched on its own.  Same kernels
    in the same order, so the solutions are bitwise equal: converged, again
    (the cached graph replayed), with another right-hand side (mixed: a new
    graph, it holds b), and stopped at a max_iter that is no multiple of 8."""
    print(f"graph vs one-by-one {prec}", flush=True)
    gm, om = pair(40, seed=8)
    rng = np.random.default_rng(7)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 2))
    S = admfem.DirichletSystem(gm, om.bcnode()).assemble(kappa=kap, f=rng.uniform(-1, 1,
                                                                                  om.ngauss))
    b2 = dev(rng.uniform(-1, 1, gm.nnode))

    def run(b, max_iter, every):  # through the C-ABI: a stopped solve returns its code
        x = torch.empty(gm.nnode, dtype=torch.float64, device="cuda")
        o = admfem.solve_opts(prec, 1e-10, max_iter, check_every=every)
        st = _lib.SolveStats()
        rc = _lib.lib().admfem_system_solve(S._h, _lib.ptr(b), _lib.ptr(x), ctypes.byref(o),
                                            ctypes.byref(st), _lib.stream_ptr())
        return rc, host(x), st.as_dict()

    for b, max_iter in ((None, 100000), (None, 100000), (b2, 100000), (b2, 21)):
        rg, ug, sg = run(b, max_iter, 0)
        re_, ue, se = run(b, max_iter, 1)
        print(max_iter, rg, sg, se, flush=True)
        assert rg == re_ == (_lib.OK if max_iter > 21 else _lib.ERR_NOT_CONVERGED)
        assert sg["iterations"] == se["iterations"]
        assert np.array_equal(ug, ue)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-8)])
def test_forward_adjoint_gradient_matches_oracle_chain(prec, tol):
    gm, om = pair(24, seed=9)
    rng = np.random.default_rng(7)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 3))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    interior = np.setdiff1d(np.arange(om.nnode), bd - 1)
    obs = rng.choice(interior, size=interior.size // 20, replace=False)
    uobs = rng.standard_normal(obs.size) * 1e-2
    S = admfem.DirichletSystem(gm, bd)
    out = S.forward_adjoint(obs, uobs, precision=prec,
                            rtol=1e-13 if prec == "fp64" else 1e-11, kappa=kap, f=f)

    def dl(u):
        g = np.zeros_like(u)
        g[obs] = 2 * (u[obs] - uobs)
        return g

    ref = O.forward_adjoint(om, kap, f, bd, np.zeros(bd.size), dl)
    assert relerr(host(out["u"]), ref["u"]) < tol
    assert relerr(host(out["grad_kappa"]), ref["grad_kappa"]) < tol
    assert out["loss"] == pytest.approx(((ref["u"][obs] - uobs) ** 2).sum(), rel=tol)


@pytest.mark.parametrize("shuffle", [False, True])
def test_kappa_gradient_entry_point_matches_oracle(shuffle):
    """admfem_system_kappa_gradient on the oracle's own u and lambda (user
    order) gives the oracle's gradient chain SparseSolverGrad ->
    ImposeDirichletGrad -> FemLaplaceScalarGrad; shuffled elements take the
    RCM node order."""
    c, e = meshgen.jittered_square(30, seed=13)
    if shuffle:
        e = e[np.random.default_rng(6).permutation(e.shape[0])]
    gm, om = admfem.Mesh(c, e, zero_based=True), O.OracleMesh(c, e)
    rng = np.random.default_rng(14)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), 5))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    interior = np.setdiff1d(np.arange(om.nnode), bd - 1)
    obs = rng.choice(interior, size=40, replace=False)
    uobs = rng.standard_normal(obs.size) * 1e-2

    def dl(u):
        g = np.zeros_like(u)
        g[obs] = 2 * (u[obs] - uobs)
        return g

    ref = O.forward_adjoint(om, kap, f, bd, np.zeros(bd.size), dl)
    S = admfem.DirichletSystem(gm, bd).assemble(kappa=kap, f=f)
    g = S.kappa_gradient(ref["u"], ref["lam"])
    assert relerr(host(g), ref["grad_kappa"]) < 1e-12


def fan_mesh(k=12):
    """Centre node, a ring of k nodes and an outer ring of 2k boundary nodes:
    the centre row has k interior neighbours (wider than the 8-entry fast
    path of the SpMV)."""
    ang1 = 2 * np.pi * np.arange(k) / k
    ang2 = 2 * np.pi * np.arange(2 * k) / (2 * k)
    coords = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(ang1), np.sin(ang1)], 1),
                             2.0 * np.stack([np.cos(ang2), np.sin(ang2)], 1)])
    r1 = lambda i: 1 + (i % k)
    r2 = lambda j: 1 + k + (j % (2 * k))
    el = [[0, r1(i), r1(i + 1)] for i in range(k)]
    for i in range(k):
        el += [[r1(i), r2(2 * i), r2(2 * i + 1)], [r1(i), r2(2 * i + 1), r1(i + 1)],
               [r1(i + 1), r2(2 * i + 1), r2(2 * i + 2)]]
    return coords, np.array(el)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-8)])
def test_wide_rows_fan_mesh(prec, tol):
    c, e = fan_mesh(12)
    gm = admfem.Mesh(c, e, zero_based=True)
    om = O.OracleMesh(c, e)
    rng = np.random.default_rng(12)
    kap = rng.random(om.ngauss) + 0.5
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    assert bd.size == 24
    S = admfem.DirichletSystem(gm, bd).assemble(kappa=kap, f=f)
    A0, r0 = _oracle_system(om, kap, f, bd, np.zeros(bd.size))
    A, r = S.csr()
    assert np.diff(A.indptr).max() == 13  # centre: itself + 12 ring nodes
    assert relerr(A.toarray(), A0.toarray()) < 1e-13
    u, st = S.solve(precision=prec, rtol=1e-13 if prec == "fp64" else 1e-11)
    u0 = O.DirectSolver(A0).solve(r0)
    assert st["status"] == 0
    assert relerr(host(u), u0) < tol
