"""Dirichlet-eliminated systems on the GPU: the fused forward + adjoint path.

``DirichletSystem(mesh, bdnode)`` owns the post-BC CSR pattern of
``impose_Dirichlet_boundary_conditions`` (src/MFEM/MUtils.jl:201-213) for a
fixed boundary node set, assembles
    A = mass_coef*M(rho) + lap_coef*K(kappa),
    rhs = src_coef*F(f) + M(rho) u_prev   (then Dirichlet lifting)
straight into it, solves with Jacobi-PCG (FP64 or FP32 + FP64 residual
replacement) and returns dL/dkappa by the adjoint method.  All vectors are
torch tensors on the GPU (float64); node vectors in the user's node order.
"""
import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib

_DEV = "cuda"


def _f64(t, n=None, name="array"):
    if t is None:
        return None
    if not torch.is_tensor(t):
        t = torch.as_tensor(np.asarray(t, dtype=np.float64))
    t = t.to(device=_DEV, dtype=torch.float64).contiguous()
    if n is not None and t.numel() != n:
        raise ValueError(f"{name}: expected {n} values, got {t.numel()}")
    return t


def solve_opts(precision="fp64", rtol=1e-10, max_iter=100000, check_every=0,
               replace_delta=0.0):
    prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
    return _lib.SolveOpts(prec, int(max_iter), float(rtol), int(check_every),
                          float(replace_delta))


class DirichletSystem:
    def __init__(self, mesh, bdnode):
        self.mesh = mesh
        bd = np.ascontiguousarray(np.asarray(bdnode, dtype=np.int64).ravel())
        self.bdnode = bd
        h = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.admfem_system_create(mesh.handle, _lib.ptr(bd), bd.size, ctypes.byref(h)))
        self._h = h
        info = np.zeros(4, dtype=np.int64)
        _lib.check(L.admfem_system_info(h, _lib.ptr(info)))
        self.n, self.nnz, self.nbd_unique, self.device_bytes = (int(v) for v in info)
        self._keep = []  # tensors referenced by the last assembled problem
        self.lap_coef = 1.0

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().admfem_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def problem(self, kappa=None, rho=None, f=None, bdval=None, u_prev=None,
                lap_coef=1.0, mass_coef=0.0, src_coef=1.0):
        ng, n = self.mesh.ngauss, self.n
        t = dict(kappa=_f64(kappa, ng, "kappa"), rho=_f64(rho, ng, "rho"), f=_f64(f, ng, "f"),
                 bdval=_f64(bdval, self.bdnode.size, "bdval"), u_prev=_f64(u_prev, n, "u_prev"))
        self._keep = list(t.values())
        self.lap_coef = float(lap_coef)
        return _lib.Problem(*[_lib.ptr(t[k]) for k in ("kappa", "rho", "f", "bdval", "u_prev")],
                            float(lap_coef), float(mass_coef), float(src_coef))

    def assemble(self, **kw):
        P = self.problem(**kw)
        _lib.check(_lib.lib().admfem_system_assemble(self._h, ctypes.byref(P),
                                                     _lib.stream_ptr()))
        return self

    def csr(self):
        """(scipy CSR matrix, rhs) of the last assembly, user node order."""
        rp = np.zeros(self.n + 1, dtype=np.int32)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        rhs = np.zeros(self.n)
        torch.cuda.synchronize()
        _lib.check(_lib.lib().admfem_system_csr(self._h, _lib.ptr(rp), _lib.ptr(col),
                                                _lib.ptr(val), _lib.ptr(rhs)))
        return sp.csr_matrix((val, col, rp), shape=(self.n, self.n)), rhs

    def solve(self, b=None, precision="fp64", rtol=1e-10, max_iter=100000, **kw):
        """Solve A x = b (b=None: the assembled rhs).  Returns (x, stats)."""
        b = _f64(b, self.n, "b")
        x = torch.empty(self.n, dtype=torch.float64, device=_DEV)
        o = solve_opts(precision, rtol, max_iter, **kw)
        st = _lib.SolveStats()
        _lib.check(_lib.lib().admfem_system_solve(self._h, _lib.ptr(b), _lib.ptr(x),
                                                  ctypes.byref(o), ctypes.byref(st),
                                                  _lib.stream_ptr()))
        return x, st.as_dict()

    def kappa_gradient(self, u, lam, scale=None):
        scale = self.lap_coef if scale is None else scale
        g = torch.empty(self.mesh.ngauss, dtype=torch.float64, device=_DEV)
        u, lam = _f64(u, self.n), _f64(lam, self.n)  # held until the kernel is queued
        _lib.check(_lib.lib().admfem_system_kappa_gradient(
            self._h, _lib.ptr(u), _lib.ptr(lam), float(scale), _lib.ptr(g), _lib.stream_ptr()))
        return g

    def forward_adjoint(self, obs_idx, obs_val, precision="fp64", rtol=1e-10,
                        max_iter=100000, want_lam=True, replace_delta=0.0, **problem):
        """Assemble, u = A^-1 rhs, L = sum (u[obs_idx]-obs_val)^2 (obs_idx
        0-based user node ids), lam = A^-1 dL/du, dL/dkappa.  Returns dict."""
        P = self.problem(**problem)
        idx = torch.as_tensor(np.asarray(obs_idx), dtype=torch.int32).to(_DEV).contiguous() \
            if not torch.is_tensor(obs_idx) else obs_idx.to(device=_DEV, dtype=torch.int32).contiguous()
        ov = _f64(obs_val, idx.numel(), "obs_val")
        u = torch.empty(self.n, dtype=torch.float64, device=_DEV)
        lam = torch.empty(self.n, dtype=torch.float64, device=_DEV) if want_lam else None
        gk = torch.empty(self.mesh.ngauss, dtype=torch.float64, device=_DEV)
        loss = ctypes.c_double()
        st = (_lib.SolveStats * 2)()
        o = solve_opts(precision, rtol, max_iter, replace_delta=replace_delta)
        _lib.check(_lib.lib().admfem_system_forward_adjoint(
            self._h, ctypes.byref(P), _lib.ptr(idx), _lib.ptr(ov), idx.numel(), ctypes.byref(o),
            _lib.ptr(u), _lib.ptr(lam), _lib.ptr(gk), ctypes.byref(loss), st,
            _lib.stream_ptr()))
        return dict(u=u, lam=lam, grad_kappa=gk, loss=loss.value,
                    stats=(st[0].as_dict(), st[1].as_dict()))
