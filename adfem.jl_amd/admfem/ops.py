"""Differentiable operators: the reference's TF custom ops and their
registered gradient ops, as torch.autograd.Functions over libadmfem.

Mirrors the PyObject methods of src/MFEM/MCore.jl / MUtils.jl:

    compute_fem_laplace_matrix1(kappa, mesh)   MCore.jl:100-107   (FemLaplaceScalar)
    compute_fem_mass_matrix1(rho, mesh)        MCore.jl:170-181   (ComputeFemMassMatrixMfem)
    compute_fem_source_term1(f, mesh)          MCore.jl:65-72     (FemSourceScalar)
    impose_Dirichlet_boundary_conditions(A, rhs, bdnode, bdval)
                                               MUtils.jl:201-225  (ImposeDirichlet)
    A \\ rhs  -> solve(A, rhs)                  ADCME SparseSolver

Sparse matrices are ``SparseTensor`` (COO triplets on the GPU, duplicates
kept, as ADCME's RawSparseTensor).  torch tensors must live on the GPU;
numpy inputs take the reference's Array paths (MCore.jl:77-82, 112-120) and
return numpy / scipy results.  ``loss.backward()`` plays the role of
ADCME's ``gradients(loss, kappa)``.
"""
import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .system import solve_opts

_DEV = "cuda"


class SparseTensor:
    """n x n matrix as COO triplets: ``indices`` [nnz, 2] int64 (0-based) and
    differentiable ``values`` [nnz] float64, both on the GPU."""

    def __init__(self, indices, values, n):
        # the reference converts indices to Int64 and values to Float64
        # (convert_to_tensor in MUtils.jl:210); .to() keeps the autograd graph
        self.indices = torch.as_tensor(indices).to(device=_DEV, dtype=torch.int64).contiguous()
        self.values = torch.as_tensor(values).to(device=_DEV, dtype=torch.float64)
        self.n = int(n)

    @property
    def shape(self):
        return (self.n, self.n)

    def __add__(self, other):
        if not isinstance(other, SparseTensor) or other.n != self.n:
            raise TypeError("SparseTensor + SparseTensor of the same size")
        return SparseTensor(torch.cat([self.indices, other.indices]),
                            torch.cat([self.values, other.values]), self.n)

    def __mul__(self, a):
        return SparseTensor(self.indices, self.values * a, self.n)

    __rmul__ = __mul__

    def to_scipy(self):
        """Host CSR with duplicates summed (explicit zeros kept)."""
        i = self.indices.detach().cpu().numpy()
        v = self.values.detach().cpu().numpy()
        A = sp.coo_matrix((v, (i[:, 0], i[:, 1])), shape=self.shape).tocsr()
        A.sum_duplicates()
        return A

    def __matmul__(self, x):
        i, j = self.indices[:, 0], self.indices[:, 1]
        y = torch.zeros(self.n, dtype=x.dtype, device=x.device)
        return y.index_add_(0, i, self.values * x[j])

    def __truediv__(self, rhs):  # not Julia's \, kept explicit: use solve()
        raise TypeError("use admfem.ops.solve(A, rhs) for A \\ rhs")


def _dev(t, dtype=torch.float64):
    return t.to(device=_DEV, dtype=dtype).contiguous()


class _LaplaceCOO(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, mesh, fwd, bwd):
        nt = mesh.elem_ndof ** 2 * mesh.ngauss
        ind = torch.empty((nt, 2), dtype=torch.int64, device=_DEV)
        vv = torch.empty(nt, dtype=torch.float64, device=_DEV)
        k = _dev(kappa)
        _lib.check(fwd(mesh.handle, _lib.ptr(k), _lib.ptr(ind), _lib.ptr(vv), _lib.stream_ptr()))
        ctx.mesh, ctx.bwd = mesh, bwd
        ctx.mark_non_differentiable(ind)
        return ind, vv

    @staticmethod
    def backward(ctx, g_ind, g_vv):
        g = _dev(g_vv)
        out = torch.empty(ctx.mesh.ngauss, dtype=torch.float64, device=_DEV)
        _lib.check(ctx.bwd(ctx.mesh.handle, _lib.ptr(g), _lib.ptr(out), _lib.stream_ptr()))
        return out, None, None, None


def compute_fem_laplace_matrix1(kappa, mesh=None):
    """FemLaplaceScalar (deps/MFEM/FemLaplace1/FemLaplaceScalar.h:3-54).
    ``compute_fem_laplace_matrix1(mesh)``: kappa = 1 (MCore.jl:137)."""
    if mesh is None:
        kappa, mesh = np.ones(kappa.ngauss), kappa
    if len(kappa) != mesh.ngauss:
        raise ValueError("length(kappa) must be get_ngauss(mesh)")
    L = _lib.lib()
    if isinstance(kappa, np.ndarray):  # Array path: sparse(I+1, J+1, V) (MCore.jl:112-120)
        ind, vv = _LaplaceCOO.apply(torch.as_tensor(kappa), mesh, L.admfem_fem_laplace_scalar,
                                    L.admfem_fem_laplace_scalar_grad)
        return SparseTensor(ind, vv, mesh.ndof).to_scipy().tocsc()
    ind, vv = _LaplaceCOO.apply(kappa, mesh, L.admfem_fem_laplace_scalar,
                                L.admfem_fem_laplace_scalar_grad)
    return SparseTensor(ind, vv, mesh.ndof)


def compute_fem_mass_matrix1(rho, mesh=None):
    """ComputeFemMassMatrixMfem (ComputeFemMassMatrixMfem.h:4-56);
    ``compute_fem_mass_matrix1(mesh)``: rho = 1 (MCore.jl:179-181)."""
    if mesh is None:
        rho, mesh = np.ones(rho.ngauss), rho
    if len(rho) != mesh.ngauss:
        raise ValueError("length(rho) must be get_ngauss(mesh)")
    L = _lib.lib()
    r = torch.as_tensor(rho) if isinstance(rho, np.ndarray) else rho
    ind, vv = _LaplaceCOO.apply(r, mesh, L.admfem_fem_mass_matrix, L.admfem_fem_mass_matrix_grad)
    return SparseTensor(ind, vv, mesh.ndof)


class _Source(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, mesh):
        out = torch.empty(mesh.ndof, dtype=torch.float64, device=_DEV)
        ff = _dev(f)
        _lib.check(_lib.lib().admfem_fem_source_scalar(mesh.handle, _lib.ptr(ff), _lib.ptr(out),
                                                       _lib.stream_ptr()))
        ctx.mesh = mesh
        return out

    @staticmethod
    def backward(ctx, g):
        g = _dev(g)
        out = torch.empty(ctx.mesh.ngauss, dtype=torch.float64, device=_DEV)
        _lib.check(_lib.lib().admfem_fem_source_scalar_grad(ctx.mesh.handle, _lib.ptr(g),
                                                            _lib.ptr(out), _lib.stream_ptr()))
        return out, None


def compute_fem_source_term1(f, mesh):
    """FemSourceScalar (deps/MFEM/FemSource1/FemSourceScalar.h:4-32)."""
    if len(f) != mesh.ngauss:
        raise ValueError("length(f) must be get_ngauss(mesh)")
    if isinstance(f, np.ndarray):
        return _Source.apply(torch.as_tensor(f), mesh).cpu().numpy()
    return _Source.apply(f, mesh)


class _ToGauss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, mesh, fwd, bwd):
        uu = _dev(u)
        out = torch.empty(mesh.ngauss, dtype=torch.float64, device=_DEV)
        _lib.check(fwd(mesh.handle, _lib.ptr(uu), _lib.ptr(out), _lib.stream_ptr()))
        ctx.mesh, ctx.bwd, ctx.n = mesh, bwd, uu.numel()
        return out

    @staticmethod
    def backward(ctx, g):
        g = _dev(g)
        gu = torch.empty(ctx.n, dtype=torch.float64, device=_DEV)
        _lib.check(ctx.bwd(ctx.mesh.handle, _lib.ptr(g), _lib.ptr(gu), _lib.stream_ptr()))
        return gu, None, None, None


def _nodal(u, mesh, n=None):
    n = mesh.ndof if n is None else n
    if len(u) != n:
        raise ValueError(f"length(u) must be {n}")
    return isinstance(u, np.ndarray)


def fem_to_gauss_points(u, mesh):
    """FemToGaussPointsMfem (deps/MFEM/FemToGaussPoints/FemToGaussPointsMfem.h:
    6-33; src/MFEM/MUtils.jl:231-250): vertex values (nnode) -> Gauss points,
    linear interpolation for either degree."""
    L = _lib.lib()
    host = _nodal(u, mesh, mesh.nnode)
    out = _ToGauss.apply(torch.as_tensor(u) if host else u, mesh,
                         L.admfem_fem_to_gauss_points, L.admfem_fem_to_gauss_points_grad)
    return out.cpu().numpy() if host else out


def dof_to_gauss_points(u, mesh):
    """DofToGaussPointsMfem (deps/MFEM/DofToGaussPoints/DofToGaussPointsMfem.h:
    6-32; src/MFEM/MUtils.jl:253-270): dof values (ndof) -> Gauss points with
    the degree's basis; for P1 the dofs are the nodes."""
    L = _lib.lib()
    host = _nodal(u, mesh)
    out = _ToGauss.apply(torch.as_tensor(u) if host else u, mesh,
                         L.admfem_dof_to_gauss_points, L.admfem_dof_to_gauss_points_grad)
    return out.cpu().numpy() if host else out


class _FemGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, mesh):
        uu = _dev(u)
        out = torch.empty((mesh.ngauss, 2), dtype=torch.float64, device=_DEV)
        _lib.check(_lib.lib().admfem_fem_grad(mesh.handle, _lib.ptr(uu), _lib.ptr(out),
                                              _lib.stream_ptr()))
        ctx.mesh = mesh
        return out

    @staticmethod
    def backward(ctx, g):
        g = _dev(g)
        gu = torch.empty(ctx.mesh.ndof, dtype=torch.float64, device=_DEV)
        _lib.check(_lib.lib().admfem_fem_grad_grad(ctx.mesh.handle, _lib.ptr(g), _lib.ptr(gu),
                                                   _lib.stream_ptr()))
        return gu, None


def eval_grad_on_gauss_pts1(u, mesh):
    """FemGradMfem (deps/MFEM/FemGrad/FemGradMfem.h:8-41; src/MFEM/MCore.jl:
    239-246): (du/dx, du/dy) at every Gauss point, shape (ngauss, 2)."""
    host = _nodal(u, mesh)
    out = _FemGrad.apply(torch.as_tensor(u) if host else u, mesh)
    return out.cpu().numpy() if host else out


class _LaplaceTerm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, nu, mesh):
        uu, nn = _dev(u), _dev(nu)
        out = torch.empty(mesh.ndof, dtype=torch.float64, device=_DEV)
        _lib.check(_lib.lib().admfem_fem_laplace_term(mesh.handle, _lib.ptr(uu), _lib.ptr(nn),
                                                      _lib.ptr(out), _lib.stream_ptr()))
        ctx.save_for_backward(uu, nn)
        ctx.mesh = mesh
        return out

    @staticmethod
    def backward(ctx, g):
        uu, nn = ctx.saved_tensors
        g = _dev(g)
        gu = torch.empty_like(uu) if ctx.needs_input_grad[0] else None
        gn = torch.empty_like(nn) if ctx.needs_input_grad[1] else None
        _lib.check(_lib.lib().admfem_fem_laplace_term_grad(
            ctx.mesh.handle, _lib.ptr(g), _lib.ptr(uu), _lib.ptr(nn), _lib.ptr(gu),
            _lib.ptr(gn), _lib.stream_ptr()))
        return gu, gn, None


def compute_fem_laplace_term1(u, nu, mesh):
    """ComputeLaplaceTermMfem (deps/MFEM/ComputeLaplaceTermMfem/
    ComputeLaplaceTermMfem.h:4-39; src/MFEM/MCore.jl:740-763): K(nu) u,
    matrix-free."""
    if len(u) != mesh.ndof or len(nu) != mesh.ngauss:
        raise ValueError("u must have mesh.ndof entries and nu get_ngauss(mesh)")
    host = isinstance(u, np.ndarray) and isinstance(nu, np.ndarray)
    t = (lambda a: torch.as_tensor(a) if isinstance(a, np.ndarray) else a)
    out = _LaplaceTerm.apply(t(u), t(nu), mesh)
    return out.cpu().numpy() if host else out


class _Dirichlet(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vv, rhs, bdval, ind, bd, N):
        sN, nb = ind.shape[0], bd.numel()
        vv, rhs, bdval = _dev(vv), _dev(rhs), _dev(bdval)
        oi = torch.empty((sN + nb, 2), dtype=torch.int64, device=_DEV)
        ov = torch.empty(sN + nb, dtype=torch.float64, device=_DEV)
        orhs = torch.empty(N, dtype=torch.float64, device=_DEV)
        cnt = np.zeros(1, dtype=np.int64)
        _lib.check(_lib.lib().admfem_impose_dirichlet(
            _lib.ptr(ind), _lib.ptr(vv), sN, _lib.ptr(bd), _lib.ptr(bdval), nb, _lib.ptr(rhs), N,
            _lib.ptr(oi), _lib.ptr(ov), _lib.ptr(orhs), _lib.ptr(cnt), _lib.stream_ptr()))
        n = int(cnt[0])
        ctx.save_for_backward(vv, bdval, ind, bd)
        ctx.N = N
        oi, ov = oi[:n].contiguous(), ov[:n].contiguous()
        ctx.mark_non_differentiable(oi)
        return oi, ov, orhs

    @staticmethod
    def backward(ctx, g_oi, g_ov, g_orhs):
        vv, bdval, ind, bd = ctx.saved_tensors
        g_ov = _dev(g_ov) if g_ov is not None else torch.zeros(0, dtype=torch.float64,
                                                                 device=_DEV)
        g_orhs = _dev(g_orhs) if g_orhs is not None else torch.zeros(ctx.N, dtype=torch.float64,
                                                                     device=_DEV)
        gv = torch.empty_like(vv)
        gr = torch.empty(ctx.N, dtype=torch.float64, device=_DEV)
        gb = torch.empty_like(bdval)
        _lib.check(_lib.lib().admfem_impose_dirichlet_grad(
            _lib.ptr(g_ov), _lib.ptr(g_orhs), _lib.ptr(ind), _lib.ptr(vv), ind.shape[0],
            _lib.ptr(bd), _lib.ptr(bdval), bd.numel(), ctx.N, _lib.ptr(gv), _lib.ptr(gr),
            _lib.ptr(gb), _lib.stream_ptr()))
        return gv, gr, gb, None, None, None


def impose_Dirichlet_boundary_conditions(A, rhs, bdnode=None, bdval=None):
    """ImposeDirichlet (deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:27-93).
    ``bdnode`` 1-based as returned by bcnode.  ``(A, bdnode)`` is the
    homogeneous helper (MUtils.jl:220-225) returning the matrix only.
    A SciPy sparse ``A`` with host arrays takes the Array path
    (MUtils.jl:184-198): the same kernels, returning ``(B, r)`` as SciPy CSC
    and NumPy."""
    if sp.issparse(A):
        if bdnode is None or bdval is None:
            raise ValueError("the SparseMatrixCSC form takes (A, rhs, bdnode, bdval)")
        C = A.tocoo()
        ind = torch.as_tensor(np.stack([C.row, C.col], 1).astype(np.int64)).to(_DEV)
        T = SparseTensor(ind, torch.as_tensor(C.data.astype(np.float64)).to(_DEV), A.shape[0])
        B, r = impose_Dirichlet_boundary_conditions(T, np.asarray(rhs, dtype=np.float64),
                                                    bdnode, bdval)
        return B.to_scipy().tocsc(), r.detach().cpu().numpy()
    homogeneous = bdval is None and bdnode is None
    if homogeneous:  # impose_Dirichlet_boundary_conditions(A, bdnode)
        bdnode, rhs = rhs, torch.zeros(A.n, dtype=torch.float64, device=_DEV)
    bd = torch.as_tensor(np.asarray(bdnode, dtype=np.int64)).to(_DEV)
    if bdval is None:
        bdval = torch.zeros(bd.numel(), dtype=torch.float64, device=_DEV)
    if not torch.is_tensor(rhs):
        rhs = torch.as_tensor(np.asarray(rhs, dtype=np.float64))
    if not torch.is_tensor(bdval):
        bdval = torch.as_tensor(np.asarray(bdval, dtype=np.float64))
    oi, ov, orhs = _Dirichlet.apply(A.values, rhs, bdval, A.indices.contiguous(), bd, A.n)
    B = SparseTensor(oi, ov, A.n)
    return B if homogeneous else (B, orhs)


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vv, rhs, ind, n, opts):
        vv, rhs = _dev(vv), _dev(rhs)
        u = torch.empty(n, dtype=torch.float64, device=_DEV)
        st = _lib.SolveStats()
        _lib.check(_lib.lib().admfem_sparse_solve(
            _lib.ptr(ind), _lib.ptr(vv), ind.shape[0], n, _lib.ptr(rhs), _lib.ptr(u),
            ctypes.byref(opts), ctypes.byref(st), _lib.stream_ptr()))
        ctx.save_for_backward(vv, ind, u)
        ctx.n, ctx.opts = n, opts
        return u

    @staticmethod
    def backward(ctx, g_u):
        vv, ind, u = ctx.saved_tensors
        g_u = _dev(g_u)
        gv = torch.empty_like(vv)
        gr = torch.empty(ctx.n, dtype=torch.float64, device=_DEV)
        st = _lib.SolveStats()
        _lib.check(_lib.lib().admfem_sparse_solve_grad(
            _lib.ptr(ind), _lib.ptr(vv), ind.shape[0], ctx.n, _lib.ptr(u), _lib.ptr(g_u),
            _lib.ptr(gv), _lib.ptr(gr), ctypes.byref(ctx.opts), ctypes.byref(st),
            _lib.stream_ptr()))
        return gv, gr, None, None, None


def solve(A, rhs, precision="fp64", rtol=1e-12, max_iter=1000000):
    """``A \\ rhs`` for an SPD SparseTensor (ADCME SparseSolver, with its
    gradient SparseSolverGrad)."""
    opts = solve_opts(precision, rtol, max_iter)
    return _Solve.apply(A.values, rhs, A.indices.contiguous(), A.n, opts)


def solve_pattern_info():
    """The pattern the general solve keeps (the last one built):
    dict(n, triplets, slots, reordered), ``reordered`` True when the solver
    works in reverse Cuthill-McKee order of it rather than the caller's."""
    info = np.zeros(4, dtype=np.int64)
    _lib.check(_lib.lib().admfem_sparse_solve_info(_lib.ptr(info)))
    return dict(n=int(info[0]), triplets=int(info[1]), slots=int(info[2]),
                reordered=bool(info[3]))
