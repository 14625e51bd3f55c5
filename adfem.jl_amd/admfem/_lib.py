"""ctypes binding of libadmfem.so (the C-ABI declared in include/admfem.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the SONAME libamdhip64.so.7): device pointers and
streams are then shared between torch tensors and the library.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadmfem.so")

c_int, c_i64, c_double, c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
c_ll_p = ctypes.POINTER(ctypes.c_longlong)

OK, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_NOT_CONVERGED, ERR_NAN, ERR_BUDGET, ERR_BREAKDOWN = range(8)
PREC_FP64, PREC_MIXED = 0, 1


class Problem(ctypes.Structure):
    _fields_ = [("kappa", c_vp), ("rho", c_vp), ("f", c_vp), ("bdval", c_vp),
                ("u_prev", c_vp), ("lap_coef", c_double), ("mass_coef", c_double),
                ("src_coef", c_double)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("precision", c_int), ("max_iter", c_int), ("rtol", c_double),
                ("check_every", c_int), ("replace_delta", c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("replacements", c_int), ("status", c_int),
                ("rel_residual", c_double), ("seconds", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AdmfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"admfem error {code}: {msg}")
        self.code = code


# name -> (restype, argtypes)
_SIGS = {
    "admfem_last_error": (ctypes.c_char_p, []),
    "admfem_version": (c_int, []),
    "admfem_system_spmv_repeat": (c_int, [c_vp, c_int, c_int, c_vp, c_vp]),
    "admfem_sparse_solve": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_vp,
                                    ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats), c_vp]),
    "admfem_sparse_solve_grad": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_vp,
                                         ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats),
                                         c_vp]),
    "admfem_sparse_solve_info": (c_int, [c_vp]),
    # A. reference-named
    "init_nnfem_mesh": (c_ll_p, [c_vp, c_int, c_vp, c_int, c_int, c_int, c_int, c_vp]),
    "mfem_get_ngauss": (c_int, []),
    "mfem_get_gauss": (None, [c_vp, c_vp]),
    "mfem_get_gauss_weights": (None, [c_vp]),
    "mfem_get_area": (None, [c_vp]),
    "mfem_get_elem_ndof": (c_int, []),
    "mfem_get_ndof": (c_int, []),
    "mfem_get_connectivity": (None, [c_vp]),
    "mfem_get_element_to_vertices": (None, [c_vp]),
    "FemLaplaceScalar_forward_Julia": (None, [c_vp, c_vp, c_vp]),
    "FemSourceScalar_forward_Julia": (None, [c_vp, c_vp]),
    "FemToGaussPointsMfem_Julia": (None, [c_vp, c_vp]),
    "DofToGaussPointsMfem_forward_Julia": (None, [c_vp, c_vp]),
    "ComputeLaplaceTermMfem_forward_Julia": (None, [c_vp, c_vp, c_vp]),
    "pcl_FemLaplaceScalar_Jacobian": (None, [c_vp]),
    "pcl_ImposeDirichlet": (None, [c_vp, c_vp, c_vp, c_int, c_int]),
    "admfem_default_mesh": (c_vp, []),
    # B. meshes
    "admfem_mesh_create": (c_int, [c_vp, c_int, c_vp, c_int, c_int, c_int, c_vp]),
    "admfem_mesh_destroy": (c_int, [c_vp]),
    "admfem_mesh_info": (c_int, [c_vp, c_vp]),
    "admfem_mesh_gauss": (c_int, [c_vp, c_vp, c_vp]),
    "admfem_mesh_weights": (c_int, [c_vp, c_vp]),
    "admfem_mesh_area": (c_int, [c_vp, c_vp]),
    "admfem_mesh_elems": (c_int, [c_vp, c_vp]),
    "admfem_mesh_dofs": (c_int, [c_vp, c_vp]),
    "admfem_mesh_edges": (c_int, [c_vp, c_vp]),
    "admfem_mesh_bcnode": (c_int, [c_vp, c_vp, c_vp]),
    "admfem_mesh_pattern": (c_int, [c_vp, c_vp, c_vp]),
    # C. operators
    "admfem_fem_laplace_scalar": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_laplace_scalar_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_mass_matrix": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_mass_matrix_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_source_scalar": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_source_scalar_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_to_gauss_points": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_to_gauss_points_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_dof_to_gauss_points": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_dof_to_gauss_points_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_grad_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_laplace_term": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "admfem_fem_laplace_term_grad": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "admfem_impose_dirichlet": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_int, c_vp, c_int,
                                        c_vp, c_vp, c_vp, c_vp, c_vp]),
    "admfem_impose_dirichlet_grad": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_int,
                                             c_int, c_vp, c_vp, c_vp, c_vp]),
    # D. systems
    "admfem_system_create": (c_int, [c_vp, c_vp, c_int, c_vp]),
    "admfem_system_destroy": (c_int, [c_vp]),
    "admfem_system_info": (c_int, [c_vp, c_vp]),
    "admfem_system_pattern": (c_int, [c_vp, c_vp, c_vp]),
    "admfem_system_csr": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "admfem_system_assemble": (c_int, [c_vp, ctypes.POINTER(Problem), c_vp]),
    "admfem_system_solve": (c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(SolveOpts),
                                    ctypes.POINTER(SolveStats), c_vp]),
    "admfem_system_kappa_gradient": (c_int, [c_vp, c_vp, c_vp, c_double, c_vp, c_vp]),
    "admfem_system_kappa_gradient_noalloc": (c_int, [c_vp, c_vp, c_vp, c_double, c_vp, c_vp]),
    "admfem_system_forward_adjoint": (c_int, [c_vp, ctypes.POINTER(Problem), c_vp, c_vp, c_int,
                                              ctypes.POINTER(SolveOpts), c_vp, c_vp, c_vp,
                                              c_vp, ctypes.POINTER(SolveStats), c_vp]),
}

_LIB = None


def lib():
    """Load the library (raises if it is not built: there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python adfem.jl_amd/build.py` "
                "(the HIP library is the only implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGS)


def check(rc):
    if rc != OK:
        msg = lib().admfem_last_error().decode(errors="replace")
        raise AdmfemError(rc, msg)
    return rc


def stream_ptr(stream=None):
    """HIP stream handle of a torch stream (current stream by default)."""
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    """Device (or host) address of a tensor / numpy array, None for None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)
