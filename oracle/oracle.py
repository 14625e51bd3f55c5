"""CPU oracle for the AdFem P1/P2 forward/adjoint path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, as the checker or the timed CPU baseline.  The product
library (``adfem.jl_amd``) never imports it.

It wraps ``liboracle.so`` (``oracle.c``: a serial C restatement of the
reference's element loops, each function citing its reference file:line) and
adds what the reference delegates to third-party code:

* the sparse direct solve ``K \\ F`` and its adjoint -- ADCME's ``SparseSolver``
  (Eigen ``SparseLU``; call sites README.md:54,75) restated with SciPy's
  SuperLU (``scipy.sparse.linalg.splu``), the same LU family;
* the gradient chain that TF's symbolic differentiation composes from the
  reference's ``...Grad`` ops (SparseSolverGrad -> ImposeDirichletGrad ->
  FemLaplaceScalarGrad), see ``forward_adjoint``.
"""
import ctypes
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_I, _L = ctypes.c_int, ctypes.c_int64


def lib():
    """Load ``liboracle.so`` (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
    L = ctypes.CDLL(path)
    sig = {
        "orc_tri_rule": (_I, [_I, _dp, _dp, _dp]),
        "orc_mesh_init": (_I, [_dp, _I, _ip, _I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_mesh_init_p2": (_I, [_dp, _I, _ip, _I, _I, _lp, _ip, _dp, _dp, _dp]),
        "orc_mesh_edges": (_L, [_ip, _I, _I, ctypes.c_void_p, ctypes.c_void_p]),
        "orc_laplace_forward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp, _lp, _dp]),
        "orc_laplace_backward": (None, [_I, _I, _I, _dp, _dp, _dp, _dp, _dp]),
        "orc_mass_forward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _lp, _dp]),
        "orc_mass_backward": (None, [_I, _I, _I, _dp, _dp, _dp, _dp]),
        "orc_source_forward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp]),
        "orc_source_backward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp]),
        "orc_dirichlet_forward": (_L, [_lp, _dp, _L, _lp, _dp, _I, _dp, _I, _lp, _dp, _dp]),
        "orc_dirichlet_backward": (None, [_dp, _dp, _dp, _dp, _dp, _lp, _dp, _L, _lp, _dp, _I, _I]),
        "orc_coo_to_csr": (_L, [_lp, _dp, _L, _I, _lp, _ip, _dp]),
        "orc_bcnode": (_I, [_ip, _I, _I, _lp]),
        "orc_fem_to_gauss_forward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp]),
        "orc_fem_to_gauss_backward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp]),
        "orc_fem_grad_forward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp]),
        "orc_fem_grad_backward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp]),
        "orc_laplace_term_forward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_laplace_term_backward": (None, [_ip, _I, _I, _I, _dp, _dp, _dp, _dp, _dp, _dp,
                                             _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class OracleMesh:
    """P1/P2 mesh precompute (deps/MFEM/Common.cpp:20-142) on the host.

    ``coords`` is nnode x 2, ``elems`` nelem x 3 with 0-based vertex ids.
    After construction ``elems`` holds the orientation-fixed connectivity
    (deps/MFEM/API.cpp:58-65 hands that back to Julia).  ``degree`` 1 or 2;
    ``order`` -1 takes the reference's default (2 for P1, 4 for P2,
    src/MFEM/MFEM.jl:71-77).  ``dofs`` (nelem x nd) are the element dofs:
    the vertices, then for P2 nnode + the MFEM id of edges (0,1),(1,2),(2,0).
    ``hs`` holds the linear vertex functions (FemToGaussPoints), ``h``,
    ``hx``, ``hy`` the degree's basis (stride nd).
    """

    def __init__(self, coords, elems, order=-1, degree=1):
        L = lib()
        if degree not in (1, 2):
            raise ValueError("degree must be 1 or 2")
        if order == -1:
            order = 2 if degree == 1 else 4
        self.degree = degree
        self.coords = _f64(coords).reshape(-1, 2)
        self.elems = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 3).copy()
        self.nnode = self.coords.shape[0]
        self.nelem = self.elems.shape[0]
        q = np.zeros(8)
        nq = L.orc_tri_rule(order, q.copy(), q.copy(), q.copy())
        if nq < 0:
            raise ValueError(f"unsupported integration order {order}")
        self.nq = nq
        ng = nq * self.nelem
        self.ngauss = ng
        self.gx, self.gy, self.w = np.zeros(ng), np.zeros(ng), np.zeros(ng)
        self.area = np.zeros(self.nelem)
        self.hs, self.hx, self.hy = np.zeros(3 * ng), np.zeros(3 * ng), np.zeros(3 * ng)
        L.orc_mesh_init(self.coords, self.nnode, self.elems, self.nelem, order,
                        self.gx, self.gy, self.w, self.area, self.hs, self.hx, self.hy)
        if degree == 1:
            self.nd, self.dofs, self.ndof, self.h = 3, self.elems, self.nnode, self.hs
        else:
            ee = np.zeros(3 * self.nelem, dtype=np.int64)
            self.nedge = L.orc_mesh_edges(self.elems, self.nelem, self.nnode, None, ee.ctypes.data)
            self.elem_edge = ee.reshape(-1, 3)
            self.nd, self.ndof = 6, self.nnode + self.nedge
            self.dofs = np.zeros((self.nelem, 6), dtype=np.int32)
            self.h, self.hx, self.hy = np.zeros(6 * ng), np.zeros(6 * ng), np.zeros(6 * ng)
            L.orc_mesh_init_p2(self.coords, self.nnode, self.elems, self.nelem, order, ee,
                               self.dofs, self.h, self.hx, self.hy)

    # --- getters (deps/MFEM/API.cpp:17-65) ---
    def edges(self):
        L = lib()
        ne = L.orc_mesh_edges(self.elems, self.nelem, self.nnode, None, None)
        out = np.zeros(2 * ne, dtype=np.int64)
        L.orc_mesh_edges(self.elems, self.nelem, self.nnode, out.ctypes.data, None)
        return out.reshape(2, ne).T.copy()

    def bcnode(self):
        """Boundary dofs, 1-based: vertices ascending; for P2 then
        nnode + the ids of boundary edges, ascending (src/MFEM/MCore.jl:439-450
        with get_edge_dof, src/MFEM/MUtils.jl:149-163)."""
        out = np.zeros(self.nnode, dtype=np.int64)
        n = lib().orc_bcnode(self.elems, self.nelem, self.nnode, out)
        vert = out[:n].copy()
        if self.degree == 1:
            return vert
        # bcedge (MCore.jl:385-406): edges used by an odd number of elements
        ids, cnt = np.unique(self.elem_edge.ravel(), return_counts=True)
        return np.concatenate([vert, self.nnode + ids[cnt % 2 == 1] + 1])

    # --- forward / backward ops ---
    def laplace_coo(self, kappa):
        n = self.nd * self.nd * self.ngauss
        ind, vv = np.zeros(2 * n, dtype=np.int64), np.zeros(n)
        lib().orc_laplace_forward(self.dofs, self.nd, self.nelem, self.nq, self.w, self.hx,
                                  self.hy, _f64(kappa), ind, vv)
        return ind.reshape(n, 2), vv

    def laplace_backward(self, grad_vv):
        g = np.zeros(self.ngauss)
        lib().orc_laplace_backward(self.nd, self.nelem, self.nq, self.w, self.hx, self.hy,
                                   _f64(grad_vv), g)
        return g

    def mass_coo(self, rho):
        n = self.nd * self.nd * self.ngauss
        ind, vv = np.zeros(2 * n, dtype=np.int64), np.zeros(n)
        lib().orc_mass_forward(self.dofs, self.nd, self.nelem, self.nq, self.w, self.h,
                               _f64(rho), ind, vv)
        return ind.reshape(n, 2), vv

    def mass_backward(self, grad_vv):
        g = np.zeros(self.ngauss)
        lib().orc_mass_backward(self.nd, self.nelem, self.nq, self.w, self.h, _f64(grad_vv), g)
        return g

    def source(self, f):
        rhs = np.zeros(self.ndof)
        lib().orc_source_forward(self.dofs, self.nd, self.nelem, self.nq, self.w, self.h,
                                 _f64(f), rhs)
        return rhs

    def source_backward(self, grad_rhs):
        g = np.zeros(self.ngauss)
        lib().orc_source_backward(self.dofs, self.nd, self.nelem, self.nq, self.w, self.h,
                                  _f64(grad_rhs), g)
        return g

    # FemToGaussPoints (vertex values, linear hs), DofToGaussPoints (dofs, h),
    # FemGrad, ComputeLaplaceTerm
    def fem_to_gauss(self, u):
        out = np.zeros(self.ngauss)
        lib().orc_fem_to_gauss_forward(self.elems, 3, self.nelem, self.nq, self.hs, _f64(u), out)
        return out

    def fem_to_gauss_backward(self, grad_out):
        g = np.zeros(self.nnode)
        lib().orc_fem_to_gauss_backward(self.elems, 3, self.nelem, self.nq, self.hs,
                                        _f64(grad_out), g)
        return g

    def dof_to_gauss(self, u):
        out = np.zeros(self.ngauss)
        lib().orc_fem_to_gauss_forward(self.dofs, self.nd, self.nelem, self.nq, self.h, _f64(u),
                                       out)
        return out

    def dof_to_gauss_backward(self, grad_out):
        g = np.zeros(self.ndof)
        lib().orc_fem_to_gauss_backward(self.dofs, self.nd, self.nelem, self.nq, self.h,
                                        _f64(grad_out), g)
        return g

    def fem_grad(self, u):
        out = np.zeros(2 * self.ngauss)
        lib().orc_fem_grad_forward(self.dofs, self.nd, self.nelem, self.nq, self.hx, self.hy,
                                   _f64(u), out)
        return out.reshape(-1, 2)

    def fem_grad_backward(self, grad_out):
        g = np.zeros(self.ndof)
        lib().orc_fem_grad_backward(self.dofs, self.nd, self.nelem, self.nq, self.hx, self.hy,
                                    _f64(grad_out).reshape(-1), g)
        return g

    def laplace_term(self, u, nu):
        out = np.zeros(self.ndof)
        lib().orc_laplace_term_forward(self.dofs, self.nd, self.nelem, self.nq, self.w, self.hx,
                                       self.hy, _f64(nu), _f64(u), out)
        return out

    def laplace_term_backward(self, u, nu, grad_out):
        gn, gu = np.zeros(self.ngauss), np.zeros(self.ndof)
        lib().orc_laplace_term_backward(self.dofs, self.nd, self.nelem, self.nq, self.w,
                                        self.hx, self.hy, _f64(nu), _f64(u), _f64(grad_out),
                                        gn, gu)
        return gn, gu


def dirichlet(indices, vv, bd, rhs, bdval):
    """ImposeDirichlet forward (deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:27-60).

    ``bd`` is 1-based as in Julia.  Returns (oindices [n,2], ov [n], orhs)."""
    indices = _i64(indices).reshape(-1)
    sN = indices.size // 2
    bd, bdval, rhs = _i64(bd), _f64(bdval), _f64(rhs)
    cap = sN + bd.size
    oi, ov, orhs = np.zeros(2 * cap, dtype=np.int64), np.zeros(cap), np.zeros(rhs.size)
    n = lib().orc_dirichlet_forward(indices, _f64(vv), sN, bd, bdval, bd.size, rhs, rhs.size,
                                    oi, ov, orhs)
    return oi[:2 * n].reshape(n, 2).copy(), ov[:n].copy(), orhs


def dirichlet_backward(grad_ov, grad_orhs, indices, vv, bd, bdval, N):
    """ImposeDirichletGrad (ImposeDirichlet.h:63-93).

    Returns (grad_vv, grad_rhs, grad_bdval)."""
    indices = _i64(indices).reshape(-1)
    sN = indices.size // 2
    bd, bdval = _i64(bd), _f64(bdval)
    gv, gr, gb = np.zeros(sN), np.zeros(N), np.zeros(bd.size)
    lib().orc_dirichlet_backward(gv, gr, gb, _f64(grad_ov), _f64(grad_orhs), indices, _f64(vv),
                                 sN, bd, bdval, bd.size, N)
    return gv, gr, gb


def pcl_impose_dirichlet(indices, bd, outdof):
    """pcl_ImposeDirichlet (deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:98-111),
    called as src/pcl.jl:14-21 does: ``indices`` 1-based [sN, 2], ``bd``
    1-based.  J [sN, outdof]: J[k, s] = 1 for the s-th triplet k whose row
    and column are both off the boundary."""
    ind = _i64(indices).reshape(-1, 2)
    bds = set(int(b) - 1 for b in _i64(bd))
    J = np.zeros((ind.shape[0], outdof))
    s = 0
    for k in range(ind.shape[0]):
        if int(ind[k, 0]) - 1 not in bds and int(ind[k, 1]) - 1 not in bds:
            J[k, s] = 1.0
            s += 1
    return J


def pcl_laplace_jacobian(mesh):
    """pcl_FemLaplaceScalar_Jacobian (deps/MFEM/FemLaplace1/FemLaplaceScalar.h:65-91):
    H [ngauss, nd^2 ngauss], H[s, t] = D D^T w of triplet t's Gauss point s
    (t in the s-th run of nd^2 triplets), zero elsewhere.  The reference's
    loop is the forward's (h:7-25) without the kappa factor; here that
    forward is evaluated at kappa = 1."""
    ng = mesh.ngauss
    _, vv = mesh.laplace_coo(np.ones(ng))
    nd2 = vv.size // ng
    H = np.zeros((ng, vv.size))
    t = np.arange(vv.size)
    H[t // nd2, t] = vv
    return H


def coo_to_csr(indices, vv, N):
    """Sum duplicates in triplet order, keep explicit zeros: the matrix that
    ``sparse(I,J,V)`` (src/MFEM/MCore.jl:118-119) / Eigen ``setFromTriplets``
    build from the reference's COO.  Returns (row_ptr int64, col int32, val)."""
    indices = _i64(indices).reshape(-1)
    sN = indices.size // 2
    rp = np.zeros(N + 1, dtype=np.int64)
    col, val = np.zeros(max(sN, 1), dtype=np.int32), np.zeros(max(sN, 1))
    nnz = lib().orc_coo_to_csr(indices, _f64(vv), sN, N, rp, col, val)
    return rp, col[:nnz].copy(), val[:nnz].copy()


def csr_matrix(indices, vv, N):
    rp, col, val = coo_to_csr(indices, vv, N)
    return sp.csr_matrix((val, col, rp), shape=(N, N))


class DirectSolver:
    """``K \\ F`` restated with SuperLU; ``solve_adjoint`` is the transposed
    solve of ADCME's SparseSolverGrad (Kᵀ λ = ∂L/∂u)."""

    def __init__(self, A):
        self.A = sp.csc_matrix(A)
        self.lu = spl.splu(self.A)

    def solve(self, b):
        return self.lu.solve(_f64(b))

    def solve_adjoint(self, g):
        return self.lu.solve(_f64(g), trans="T")


def forward_adjoint(mesh, kappa, f, bd, bdval, dloss_du, *, rho=None, dt=None,
                    u_prev=None):
    """One forward + adjoint solve composed from the reference's ops.

    Poisson (``dt is None``): K = laplace(κ), F = source(f).
    Heat step: A = mass(ρ) + dt·laplace(κ) (src-level pattern of
    docs/src/heatequation.md:65,81 restated on P1), rhs = M·u_prev + dt·F.
    Then (K_bc, F_bc) = ImposeDirichlet(A, rhs, bd, bdval), u = K_bc \\ F_bc,
    g = dloss_du(u), and the gradient chain of the reference's Grad ops:
      SparseSolverGrad:   λ = K_bcᵀ \\ g;  ∂vv_k = −λ[i_k]·u[j_k];  ∂F = λ
      ImposeDirichletGrad (ImposeDirichlet.h:63-93) → ∂vv (pre-BC), ∂rhs
      FemLaplaceScalarGrad (FemLaplaceScalar.h:28-54) → ∂L/∂κ.
    Returns dict(u, lam, grad_kappa, loss_grad, A_bc (csr), rhs_bc, K_coo).
    """
    N = mesh.ndof
    Ki, Kv = mesh.laplace_coo(kappa)
    F = mesh.source(f)
    if dt is None:
        Ai, Av, rhs = Ki, Kv, F
    else:
        rho = np.ones(mesh.ngauss) if rho is None else rho
        Mi, Mv = mesh.mass_coo(rho)
        Msp = csr_matrix(Mi, Mv, N)
        Ai = np.concatenate([Mi, Ki])
        Av = np.concatenate([Mv, dt * Kv])
        rhs = Msp @ _f64(u_prev) + dt * F
    oi, ov, orhs = dirichlet(Ai, Av, bd, rhs, bdval)
    A = csr_matrix(oi, ov, N)
    solver = DirectSolver(A)
    u = solver.solve(orhs)
    g = _f64(dloss_du(u))
    lam = solver.solve_adjoint(g)
    grad_ov = -lam[oi[:, 0]] * u[oi[:, 1]]
    gvv, grhs, gbd = dirichlet_backward(grad_ov, lam, Ai, Av, bd, bdval, N)
    if dt is None:
        gk = mesh.laplace_backward(gvv)
    else:
        nM = Mi.shape[0]
        gk = mesh.laplace_backward(dt * gvv[nM:])
    return dict(u=u, lam=lam, grad_kappa=gk, grad_rhs=grhs, grad_bdval=gbd,
                A_bc=A, rhs_bc=orhs, solver=solver)
