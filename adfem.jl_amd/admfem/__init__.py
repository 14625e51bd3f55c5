"""admfem -- MI355X-native drop-in for AdFem.jl's P1/P2 forward/adjoint path.

The Python mirror of the reference's Julia operator surface
(src/MFEM/MFEM.jl, MCore.jl, MUtils.jl) over libadmfem.so's C-ABI.
Every compute call runs the library's HIP kernels; there is no CPU
fallback -- a missing library raises.
"""
from ._lib import AdmfemError, lib  # noqa: F401
from .mesh import (Mesh, bcedge, bcnode, fem_nodes, gauss_nodes, gauss_weights,  # noqa: F401
                   get_area, get_ngauss)
from .system import DirichletSystem, solve_opts  # noqa: F401
from . import meshgen  # noqa: F401
from .ops import (SparseTensor, compute_fem_laplace_matrix1, compute_fem_mass_matrix1,  # noqa: F401,E402
                  compute_fem_source_term1, impose_Dirichlet_boundary_conditions, solve,
                  fem_to_gauss_points, dof_to_gauss_points, eval_grad_on_gauss_pts1,
                  compute_fem_laplace_term1)
