"""Serving engine: continuous batching of forward+adjoint requests on one GPU.

Each request is the reference's forward + gradient pipeline
(compute_fem_laplace_matrix1 / compute_fem_mass_matrix1 /
compute_fem_source_term1 -> impose_Dirichlet_boundary_conditions -> K\\F ->
gradients of L = sum (u[obs] - obs_val)^2 w.r.t. kappa) on one of the
registered meshes.  The engine (libadmfem, C++) admits requests as they
arrive and as device memory frees up under an HBM budget, advances all
admitted systems together with one batched launch per CG step, and reports
per-request latency.
"""
import ctypes

import numpy as np

from . import _lib

c_int, c_double, c_vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p

MODE_BATCHED, MODE_SEQ_FP64 = 0, 1
SRC_HOST, SRC_DEVICE, SRC_SYNTH = 0, 1, 2


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_double), ("max_active", c_int), ("mode", c_int),
                ("rtol", c_double), ("max_iter", c_int), ("replace_delta", c_double),
                ("chunk", c_int), ("use_graph", c_int)]


class Request(ctypes.Structure):
    _fields_ = [("topo", c_int), ("kind", c_int), ("dt", c_double), ("arrival", c_double),
                ("source", c_int), ("coords", c_vp), ("kappa", c_vp), ("f", c_vp),
                ("u_prev", c_vp), ("obs_idx", c_vp), ("obs_val", c_vp), ("nobs", c_int),
                ("seed", ctypes.c_ulonglong), ("u_out", c_vp), ("grad_out", c_vp)]


class Result(ctypes.Structure):
    _fields_ = [("status", c_int), ("done", c_int), ("t_arrival", c_double),
                ("t_admit", c_double), ("t_done", c_double), ("loss", c_double),
                ("iters_fwd", c_int), ("iters_adj", c_int), ("relres_fwd", c_double),
                ("relres_adj", c_double), ("device_bytes", c_double)]


_SIGS = {
    "admfem_server_create": (c_int, [ctypes.POINTER(ServerOpts), c_vp]),
    "admfem_server_destroy": (c_int, [c_vp]),
    "admfem_server_add_topology": (c_int, [c_vp, c_vp, c_double, c_vp]),
    "admfem_server_submit": (c_int, [c_vp, ctypes.POINTER(Request), c_vp]),
    "admfem_server_run": (c_int, [c_vp, c_double]),
    "admfem_server_result": (c_int, [c_vp, ctypes.c_int64, ctypes.POINTER(Result)]),
    "admfem_server_stats": (c_int, [c_vp, c_vp]),
    "admfem_server_spmv_repeat": (c_int, [c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp]),
}
_bound = False


def _L():
    global _bound
    L = _lib.lib()
    if not _bound:
        for n, (res, args) in _SIGS.items():
            fn = getattr(L, n)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def exported_symbols():
    return list(_SIGS)


def _on_device(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Server:
    """``Server(budget_bytes, mode="batched"|"seq_fp64", rtol=1e-8, ...)``.

    Topologies are DirichletSystem objects (mesh + boundary nodes); requests
    use them by id.  Host input arrays are kept alive here until the request
    is admitted; outputs (u in user node order, dL/dkappa) are filled in when
    the request completes.
    """

    def __init__(self, budget_bytes, max_active=1024, mode="batched", rtol=1e-8,
                 max_iter=200000, replace_delta=0.0, chunk=0, use_graph=True):
        m = {"batched": MODE_BATCHED, "seq_fp64": MODE_SEQ_FP64}[mode]
        o = ServerOpts(float(budget_bytes), int(max_active), m, float(rtol), int(max_iter),
                       float(replace_delta), int(chunk), int(bool(use_graph)))
        h = c_vp()
        _lib.check(_L().admfem_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._topos = []
        self._keep = {}
        self._out = {}

    def close(self):
        if getattr(self, "_h", None):
            _L().admfem_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_topology(self, system, jitter=0.0):
        tid = c_int()
        _lib.check(_L().admfem_server_add_topology(self._h, system._h, float(jitter),
                                                  ctypes.byref(tid)))
        self._topos.append(system)
        return tid.value

    def submit(self, topo, kind="poisson", dt=0.0, arrival=0.0, seed=None, coords=None,
               kappa=None, f=None, u_prev=None, obs_idx=None, obs_val=None, want_u=True,
               want_grad=True):
        """Synthetic inputs when ``seed`` is given (generated on the device),
        torch device tensors (kept alive here until the request is done) or
        host arrays otherwise.  Returns the request id."""
        S = self._topos[topo]
        k = {"poisson": 0, "heat": 1}[kind]
        keep = []
        if seed is not None:
            src, nobs = SRC_SYNTH, 0
            ci = ki = fi = ui = oi = ov = None
        elif _on_device(kappa):
            import torch
            src = SRC_DEVICE

            def dev(t, dtype):
                return None if t is None else t.to(device="cuda", dtype=dtype).contiguous()
            ci, ki, fi, ui = (dev(t, torch.float64) for t in (coords, kappa, f, u_prev))
            oi, ov = dev(obs_idx, torch.int32), dev(obs_val, torch.float64)
            nobs = 0 if oi is None else oi.numel()
            keep = [ci, ki, fi, ui, oi, ov]
            torch.cuda.current_stream().synchronize()  # the engine reads them on its stream
        else:
            src = SRC_HOST
            ci, ki, fi, ui = _f64(coords), _f64(kappa), _f64(f), _f64(u_prev)
            oi = None if obs_idx is None else np.ascontiguousarray(obs_idx, dtype=np.int32)
            ov = _f64(obs_val)
            nobs = 0 if oi is None else oi.size
            keep = [ci, ki, fi, ui, oi, ov]
        u = np.zeros(S.n) if want_u else None
        g = np.zeros(S.mesh.ngauss) if want_grad else None
        r = Request(topo, k, float(dt), float(arrival), src, _lib.ptr(ci), _lib.ptr(ki),
                    _lib.ptr(fi), _lib.ptr(ui), _lib.ptr(oi), _lib.ptr(ov), int(nobs),
                    int(seed or 0), _lib.ptr(u), _lib.ptr(g))
        rid = ctypes.c_int64()
        _lib.check(_L().admfem_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self._keep[rid.value] = keep
        self._out[rid.value] = (u, g)
        return rid.value

    def run(self, timeout=1e9):
        _lib.check(_L().admfem_server_run(self._h, float(timeout)))

    def result(self, rid):
        r = Result()
        _lib.check(_L().admfem_server_result(self._h, int(rid), ctypes.byref(r)))
        d = {k: getattr(r, k) for k, _ in Result._fields_}
        d["u"], d["grad"] = self._out[rid]
        d["latency"] = d["t_done"] - d["t_arrival"]
        return d

    def spmv_repeat(self, topos, repeat):
        """Instrumentation: the batched SpMV kernel over one synthetic Poisson
        system per entry of ``topos``, ``repeat`` launches on torch's current
        stream.  Returns (algorithmic bytes per launch, average seconds per
        launch from HIP events on that stream)."""
        t = np.ascontiguousarray(topos, dtype=np.int32)
        b, sec = ctypes.c_double(), ctypes.c_double()
        _lib.check(_L().admfem_server_spmv_repeat(self._h, _lib.ptr(t), int(t.size), int(repeat),
                                                  ctypes.byref(b), ctypes.byref(sec),
                                                  _lib.stream_ptr()))
        return b.value, sec.value

    def stats(self):
        s = np.zeros(8)
        _lib.check(_L().admfem_server_stats(self._h, _lib.ptr(s)))
        keys = ["done", "wall_s", "peak_bytes", "peak_active", "steps", "iterations", "failed",
                "mean_batch"]
        return dict(zip(keys, s.tolist()))
