/*
 * libadmfem -- MI355X (gfx950) drop-in for AdFem's P1-triangle forward and
 * adjoint path.
 *
 * Two layers share one implementation:
 *
 *  1. Reference-named symbols (section A).  Same names, argument order and
 *     meaning as the reference library's C-ABI (deps/MFEM/API.cpp and the
 *     *_forward_Julia functions), on a process-wide default mesh, with HOST
 *     pointers, so that the reference's Julia `ccall`s bind unchanged.
 *
 *  2. Handle-based entry points (sections B-E), DEVICE pointers and a HIP
 *     stream argument (hipStream_t passed as void*; NULL = default stream).
 *     They return an admfem status code; admfem_last_error() describes the
 *     most recent failure on the calling thread.  Section C mirrors the
 *     reference's TensorFlow custom ops X / XGrad (forward op and its
 *     registered gradient op) with the same inputs and outputs, for
 *     ADCME's load_op_and_grad or a ChainRules rrule to bind.
 *
 * Indices: node/element ids are 0-based in arrays handed to kernels
 * (COO indices, elems), and 1-based where the reference's Julia API is
 * 1-based (bdnode, connectivity getters, the edge table).
 * Per-Gauss-point arrays are element-major: g = nq*e + k.
 */
#ifndef ADMFEM_H
#define ADMFEM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define ADMFEM_OK 0
#define ADMFEM_ERR_ARG 1          /* invalid argument / size */
#define ADMFEM_ERR_HIP 2          /* a HIP runtime call failed */
#define ADMFEM_ERR_UNSUPPORTED 3  /* element degree / quadrature not built */
#define ADMFEM_ERR_NOT_CONVERGED 4
#define ADMFEM_ERR_NAN 5          /* NaN/Inf met in the solve */
#define ADMFEM_ERR_BUDGET 6       /* request larger than the HBM budget */
#define ADMFEM_ERR_BREAKDOWN 7    /* CG breakdown (p.Ap <= 0: not SPD) */

const char *admfem_last_error(void);
int admfem_version(void);

typedef struct admfem_mesh_s *admfem_mesh;
typedef struct admfem_system_s *admfem_system;

/* ======================================================================
 * A. Reference-named C-ABI (default mesh, host pointers)
 * ==================================================================== */

/* deps/MFEM/API.cpp:4-14 (init_nnfem_mesh -> NNFEM_Mesh::init,
 * deps/MFEM/Common.cpp:20-142).  vertices: 3*num_vertices (x,y,z; z unused),
 * element_indices: 3*num_elements, 0-based.  Builds the default mesh
 * (replacing a previous one, with the reference's warning).  Returns a
 * malloc()ed array of 2*nedges 1-based vertex ids [min...; max...] that the
 * caller frees (Julia: unsafe_wrap(own=true)), and writes nedges.  Degrees
 * 1 (P1) and 2 (P2) and orders 0-4 (1-, 3- and 6-point rules; -1: 2 for P1,
 * 4 for P2) are built; other values return NULL with admfem_last_error()
 * set. */
long long *init_nnfem_mesh(double *vertices, int num_vertices,
                           int *element_indices, int num_elements, int order,
                           int lorder, int degree, long long *nedges);
/* deps/MFEM/API.cpp:17-19 */
int mfem_get_ngauss(void);
/* deps/MFEM/API.cpp:21-24: Gauss point coordinates, ngauss each. */
void mfem_get_gauss(double *x, double *y);
/* deps/MFEM/API.cpp:26-34 */
void mfem_get_gauss_weights(double *w);
/* deps/MFEM/API.cpp:36-39 */
void mfem_get_area(double *a);
/* deps/MFEM/API.cpp:41-43 */
int mfem_get_elem_ndof(void);
/* deps/MFEM/API.cpp:45-47 */
int mfem_get_ndof(void);
/* deps/MFEM/API.cpp:49-56: 1-based dofs, element-major. */
void mfem_get_connectivity(long long *conn);
/* deps/MFEM/API.cpp:58-65: column-major nelem x 3, 1-based, oriented. */
void mfem_get_element_to_vertices(long long *elems);
/* deps/MFEM/FemLaplace1/FemLaplaceScalar.h:59-61 (forward of h:3-26):
 * indices [2*nd*nd*ngauss] (i,j pairs, 0-based), vv [nd*nd*ngauss],
 * kappa [ngauss]; nd = elem_ndof (3 for P1, 6 for P2). */
void FemLaplaceScalar_forward_Julia(long long *indices, double *vv,
                                    const double *kappa);
/* deps/MFEM/FemSource1/FemSourceScalar.h:35-37 (forward of h:4-15):
 * rhs [ndof] is ACCUMULATED into (the caller zeroes it). */
void FemSourceScalar_forward_Julia(double *rhs, const double *f);
/* deps/MFEM/FemToGaussPoints/FemToGaussPointsMfem.h:36-48 (called from
 * src/MFEM/MUtils.jl:248): out [ngauss] = u [nnode] interpolated to the
 * Gauss points (overwritten). */
void FemToGaussPointsMfem_Julia(double *out, const double *u);
/* deps/MFEM/DofToGaussPoints/DofToGaussPointsMfem.h:35-37 (MUtils.jl:268):
 * out [ngauss] from u [ndof] with the degree's basis (P1: the nodes). */
void DofToGaussPointsMfem_forward_Julia(double *out, const double *u);
/* deps/MFEM/ComputeLaplaceTermMfem/ComputeLaplaceTermMfem.h:42-44
 * (src/MFEM/MCore.jl:744-745): out [ndof] += K(nu) u, ACCUMULATED (the
 * caller zeroes it). */
void ComputeLaplaceTermMfem_forward_Julia(double *out, const double *nu, const double *u);
/* deps/MFEM/FemLaplace1/FemLaplaceScalar.h:65-91 (src/pcl.jl:35-39): the
 * Jacobian of FemLaplaceScalar's values with respect to kappa, H [ngauss x
 * nd*nd*ngauss] column-major, H[s + t*ngauss] = d vv[t] / d kappa[s].  Only
 * the nonzeros (t in the s-th run of nd*nd triplets) are written; the caller
 * zeroes H, as the reference's does. */
void pcl_FemLaplaceScalar_Jacobian(double *H);
/* deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:98-111 (src/pcl.jl:14-21):
 * the Jacobian of ImposeDirichlet's output values with respect to its input
 * values, J [sN x outdof] column-major: J[k + s*sN] = 1 for the s-th kept
 * triplet k (neither index a boundary node).  indices: 1-based [sN x 2]
 * column-major; bd: 1-based [bdN].  Only the ones are written (the caller
 * zeroes J).  Host-side index bookkeeping: no device work. */
void pcl_ImposeDirichlet(double *J, const long long *indices, const long long *bd, int bdN,
                         int sN);
/* The default mesh as a handle, for sections B-E (NULL if none). */
admfem_mesh admfem_default_mesh(void);

/* ======================================================================
 * B. Mesh handles (replace the process-global `mmesh`,
 *    deps/MFEM/Common.cpp:7, so several meshes can live at once)
 * ==================================================================== */

/* coords: host nnode x 2 row-major; elems: host nelem x 3 row-major 0-based
 * (orientation is fixed as NNFEM_Mesh::init does, Common.cpp:40).
 * order: integration order 0-4 (-1 -> the reference default: 2 for P1, 4
 * for P2, src/MFEM/MFEM.jl:71-77); degree 1 (P1) or 2 (P2: dofs are the
 * vertices, then nnode + the MFEM id of each element's edges (0,1),(1,2),
 * (2,0), Common.cpp:63-72). */
int admfem_mesh_create(const double *coords, int nnode, const int *elems,
                       int nelem, int order, int degree, admfem_mesh *out);
int admfem_mesh_destroy(admfem_mesh m);
/* info[0..7] = nnode, nelem, ngauss, ndof, elem_ndof, nedges,
 *              nnz of the assembled (pre-BC) matrix, quadrature points/elem */
int admfem_mesh_info(admfem_mesh m, int64_t info[8]);
/* Host-pointer getters, same layouts as section A. */
int admfem_mesh_gauss(admfem_mesh m, double *x, double *y);
int admfem_mesh_weights(admfem_mesh m, double *w);
int admfem_mesh_area(admfem_mesh m, double *a);
int admfem_mesh_elems(admfem_mesh m, int *elems); /* nelem x 3 row-major 0-based, oriented */
/* Element dofs, nelem x elem_ndof row-major, 0-based (what
 * mfem_get_connectivity returns 1-based, deps/MFEM/API.cpp:49-56). */
int admfem_mesh_dofs(admfem_mesh m, int *dofs);
int admfem_mesh_edges(admfem_mesh m, long long *edges); /* 2*nedges, as init_nnfem_mesh */
/* Boundary nodes (edges used by one element), ascending, 1-based
 * (src/MFEM/MCore.jl:385-406, 439-450 bcedge/bcnode; the reference returns a
 * Set's order, which nothing downstream depends on); for P2 then the
 * boundary edges' dofs nnode + edge id + 1, ascending (MCore.jl:441-443).
 * out capacity ndof. */
int admfem_mesh_bcnode(admfem_mesh m, long long *out, int *count);
/* Pre-BC CSR pattern (sorted columns, explicit zeros kept), host pointers. */
int admfem_mesh_pattern(admfem_mesh m, int *row_ptr, int *col);

/* ======================================================================
 * C. Operators with their gradients (device pointers).  Each mirrors one
 *    reference TF op "X" and its gradient op "XGrad".
 * ==================================================================== */

/* FemLaplaceScalar (deps/MFEM/FemLaplace1/FemLaplaceScalar.cpp:14-33,
 * forward FemLaplaceScalar.h:3-26): kappa [ngauss] -> indices
 * [nd*nd*ngauss x 2] int64, vv [nd*nd*ngauss] (nd = elem_ndof: 3 for P1, 6
 * for P2).  Same triplet order as the reference. */
int admfem_fem_laplace_scalar(admfem_mesh m, const double *kappa,
                              long long *indices, double *vv, void *stream);
/* FemLaplaceScalarGrad (FemLaplaceScalar.h:28-54): grad_vv -> grad_kappa. */
int admfem_fem_laplace_scalar_grad(admfem_mesh m, const double *grad_vv,
                                   double *grad_kappa, void *stream);
/* ComputeFemMassMatrixMfem (deps/MFEM/ComputeFemMassMatrix1/
 * ComputeFemMassMatrixMfem.h:4-29): rho [ngauss] -> indices, vv. */
int admfem_fem_mass_matrix(admfem_mesh m, const double *rho,
                           long long *indices, double *vv, void *stream);
/* ComputeFemMassMatrixMfemGrad (h:31-56): grad_vv -> grad_rho
 * (overwritten). */
int admfem_fem_mass_matrix_grad(admfem_mesh m, const double *grad_vv,
                                double *grad_rho, void *stream);
/* FemSourceScalar (deps/MFEM/FemSource1/FemSourceScalar.h:4-15):
 * f [ngauss] -> rhs [ndof] (overwritten, as the op zeroes it first,
 * FemSourceScalar.cpp:73).  Atomic-free node gather. */
int admfem_fem_source_scalar(admfem_mesh m, const double *f, double *rhs,
                             void *stream);
/* FemSourceScalarGrad (FemSourceScalar.h:17-32): grad_rhs -> grad_f. */
int admfem_fem_source_scalar_grad(admfem_mesh m, const double *grad_rhs,
                                  double *grad_f, void *stream);

/* FemToGaussPointsMfem (deps/MFEM/FemToGaussPoints/FemToGaussPointsMfem.h:
 * 6-18): u [nnode] (user node order) -> out [ngauss] (overwritten), with the
 * linear vertex functions for either degree. */
int admfem_fem_to_gauss_points(admfem_mesh m, const double *u, double *out, void *stream);
/* FemToGaussPointsMfemGrad (h:20-33): grad_out [ngauss] -> grad_u [nnode]
 * (overwritten).  Atomic-free node gather. */
int admfem_fem_to_gauss_points_grad(admfem_mesh m, const double *grad_out, double *grad_u,
                                    void *stream);
/* DofToGaussPointsMfem / ...Grad (deps/MFEM/DofToGaussPoints/
 * DofToGaussPointsMfem.h:6-32): u [ndof] -> out [ngauss] with the degree's
 * basis, and grad_out [ngauss] -> grad_u [ndof] (overwritten); for P1 the
 * dofs are the nodes and these equal the two functions above. */
int admfem_dof_to_gauss_points(admfem_mesh m, const double *u, double *out, void *stream);
int admfem_dof_to_gauss_points_grad(admfem_mesh m, const double *grad_out, double *grad_u,
                                    void *stream);
/* FemGradMfem (deps/MFEM/FemGrad/FemGradMfem.h:8-22), the op behind
 * eval_grad_on_gauss_pts1 (src/MFEM/MCore.jl:239-246): u [ndof] ->
 * out [ngauss x 2] row-major (du/dx, du/dy) per Gauss point. */
int admfem_fem_grad(admfem_mesh m, const double *u, double *out, void *stream);
/* FemGradMfemGrad (h:24-41): grad_out [ngauss x 2] -> grad_u [ndof]. */
int admfem_fem_grad_grad(admfem_mesh m, const double *grad_out, double *grad_u, void *stream);
/* ComputeLaplaceTermMfem (deps/MFEM/ComputeLaplaceTermMfem/
 * ComputeLaplaceTermMfem.h:4-17), the op behind compute_fem_laplace_term1
 * (src/MFEM/MCore.jl:740-763): out [ndof] = K(nu) u without forming K
 * (overwritten); u [ndof], nu [ngauss].  Takes a stream-ordered temporary
 * of ndof doubles (u in the internal order). */
int admfem_fem_laplace_term(admfem_mesh m, const double *u, const double *nu, double *out,
                            void *stream);
/* ComputeLaplaceTermMfemGrad (h:19-39): grad_u = K(nu) grad_out (K is
 * symmetric), grad_nu[g] = w_g grad(grad_out) . grad(u) on g's element.
 * Either output may be NULL. */
int admfem_fem_laplace_term_grad(admfem_mesh m, const double *grad_out, const double *u,
                                 const double *nu, double *grad_u, double *grad_nu,
                                 void *stream);
/* ImposeDirichlet (deps/MFEM/ImposeDirichlet/ImposeDirichlet.cpp:14-40,
 * forward ImposeDirichlet.h:27-60).  indices [sN x 2] int64 0-based, vv [sN],
 * bd [bdN] int64 1-based, rhs [N], bdval [bdN].  Outputs: oindices
 * [cap x 2], ov [cap], orhs [N], with cap >= sN + bdN; *out_count (host)
 * receives the output triplet count.  Kept triplets keep their input order
 * (explicit zeros and duplicates too), then one (b,b,1.0) per distinct
 * boundary node in ascending order; duplicated bd entries: the last one
 * wins, as with the reference's std::map assignment.  Synchronises the
 * stream (the count is data dependent).  Works in a per-device workspace
 * grown to the largest call so far and kept until the process ends; calls
 * on one device (this op and its gradient) are serialised by a lock. */
int admfem_impose_dirichlet(const long long *indices, const double *vv,
                            int64_t sN, const long long *bd,
                            const double *bdval, int bdN, const double *rhs,
                            int N, long long *oindices, double *ov,
                            double *orhs, int64_t *out_count, void *stream);
/* ImposeDirichletGrad (ImposeDirichlet.h:63-93): grad_ov [count],
 * grad_orhs [N] -> grad_vv [sN], grad_rhs [N], grad_bdval [bdN]. */
int admfem_impose_dirichlet_grad(const double *grad_ov, const double *grad_orhs,
                                 const long long *indices, const double *vv,
                                 int64_t sN, const long long *bd,
                                 const double *bdval, int bdN, int N,
                                 double *grad_vv, double *grad_rhs,
                                 double *grad_bdval, void *stream);

/* ======================================================================
 * D. Systems: mesh + Dirichlet node set, assembled straight into CSR
 *    (the fused forward path: compute_fem_laplace_matrix1 / mass / source
 *    -> impose_Dirichlet_boundary_conditions -> K\F -> gradients).
 * ==================================================================== */

/* bdnode: host, 1-based (as bcnode returns), duplicates allowed (the last
 * bdval entry of a node wins).  The post-BC CSR pattern is the reference's:
 * interior-interior element pairs (explicit zeros kept) plus the boundary
 * diagonals.  P1 and P2 meshes (P2: rows are the ndof dofs, bdnode holds
 * 1-based dof ids as bcnode returns them, every node vector below is ndof
 * long). */
int admfem_system_create(admfem_mesh m, const long long *bdnode, int nbd,
                         admfem_system *out);
int admfem_system_destroy(admfem_system s);
/* info[0..3] = n, nnz (post-BC), number of distinct boundary nodes,
 *              bytes of device memory held */
int admfem_system_info(admfem_system s, int64_t info[4]);
/* Post-BC CSR pattern, host pointers (row_ptr [n+1], col [nnz]). */
int admfem_system_pattern(admfem_system s, int *row_ptr, int *col);

/* What to assemble.  A = mass_coef*M(rho) + lap_coef*K(kappa);
 * rhs = src_coef*F(f) + M(rho)*u_prev (if u_prev) ; Dirichlet values bdval
 * (per bdnode entry as given to system_create; NULL = homogeneous).
 * NULL kappa / rho mean all-ones.  Device pointers. */
typedef struct {
  const double *kappa;  /* [ngauss] or NULL */
  const double *rho;    /* [ngauss] or NULL */
  const double *f;      /* [ngauss] or NULL (no source) */
  const double *bdval;  /* [nbd] or NULL */
  const double *u_prev; /* [n] or NULL */
  double lap_coef;      /* multiplier of K (1 for Poisson, dt for heat) */
  double mass_coef;     /* multiplier of M (0 for Poisson, 1 for heat) */
  double src_coef;      /* multiplier of F (1 for Poisson, dt for heat) */
} admfem_problem;

/* Assembles A (post-BC CSR values, FP64) and rhs into the system's own
 * device buffers.  No atomics: each row gathers from its incident elements. */
int admfem_system_assemble(admfem_system s, const admfem_problem *p,
                           void *stream);
/* The last assembled system as host CSR in user node order (row_ptr [n+1],
 * col [nnz] ascending per row, val [nnz]) and rhs [n] (may be NULL):
 * the matrix and vector impose_Dirichlet_boundary_conditions returns
 * (src/MFEM/MUtils.jl:201-213), duplicates summed.  Synchronises. */
int admfem_system_csr(admfem_system s, int *row_ptr, int *col, double *val,
                      double *rhs);

#define ADMFEM_PREC_FP64 0   /* FP64 Jacobi-PCG */
#define ADMFEM_PREC_MIXED 1  /* FP32 Jacobi-PCG, FP64 residual replacement */
typedef struct {
  int precision;      /* ADMFEM_PREC_* */
  int max_iter;       /* total inner iterations */
  double rtol;        /* ||b - A x||_2 <= rtol * ||b||_2 (true residual) */
  int check_every;    /* iterations between host convergence checks (0: auto) */
  double replace_delta; /* mixed: FP64 residual replacement when the inner
                           residual fell by this factor since the last one
                           (0: 0.3); also forced after 64 iterations without
                           a 1% drop */
} admfem_solve_opts;

typedef struct {
  int iterations;
  int replacements;   /* FP64 residual replacements done (mixed: at the
                         replacement points; FP64: when the recursive
                         residual met rtol but the true one did not) */
  int status;         /* ADMFEM_OK / NOT_CONVERGED / NAN / BREAKDOWN */
  double rel_residual; /* final true residual ||b-Ax||/||b|| in FP64 */
  double seconds;
} admfem_solve_stats;

/* Solves A x = b with the system's assembled A.  b, x: device [n]; b NULL =
 * the assembled rhs.  x is the output (its input value is ignored).  The
 * solver replays its iterations from a HIP graph (8 iterations, kernel nodes
 * only), captured on the first solve of a setting on a stream of its own
 * and replayed on the caller's; the system holds both until it is
 * destroyed.  check_every = 1 launches every iteration by itself (same
 * kernels and order, the same result bit for bit). */
int admfem_system_solve(admfem_system s, const double *b, double *x,
                        const admfem_solve_opts *o, admfem_solve_stats *st,
                        void *stream);

/* Adjoint gradient, without forming dK: for every element e and Gauss
 * point k, grad_kappa[g] = -scale * w_g * (grad lam~ . grad u) on e, where
 * lam~ is lam with the system's boundary entries set to zero.  This equals
 * the reference chain SparseSolverGrad -> ImposeDirichletGrad
 * (ImposeDirichlet.h:63-93, which drops the boundary rows' gradient) ->
 * FemLaplaceScalarGrad (FemLaplaceScalar.h:28-54).  scale = lap_coef. */
int admfem_system_kappa_gradient(admfem_system s, const double *u,
                                 const double *lam, double scale,
                                 double *grad_kappa, void *stream);

/* No counterpart in the reference: admfem_system_kappa_gradient for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Same arguments, kernels and result (bit
 * for bit); it stages (u, lam) in a fixed 32 MiB device array instead of the
 * system's buffer and calls no allocating routine.  Limit: systems of at
 * most 2,097,152 nodes (larger ones return ADMFEM_ERR_UNSUPPORTED before any
 * device work).  The array is one per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU (other
 * streams or threads) race. */
int admfem_system_kappa_gradient_noalloc(admfem_system s, const double *u,
                                         const double *lam, double scale,
                                         double *grad_kappa, void *stream);

/* One forward + adjoint solve for L = sum_k (u[obs_idx[k]] - obs_val[k])^2:
 * assemble, solve A u = rhs, g = dL/du, solve A lam = g, grad_kappa.
 * obs_idx: device int32 [nobs], 0-based, distinct (an index outside [0, n)
 * or repeated fails the call with ADMFEM_ERR_ARG).  Outputs u [n], lam [n]
 * (may be NULL), grad_kappa [ngauss]; *loss (host) receives L.  Stats for
 * the forward and adjoint solves in st[0], st[1]. */
int admfem_system_forward_adjoint(admfem_system s, const admfem_problem *p,
                                  const int *obs_idx, const double *obs_val,
                                  int nobs, const admfem_solve_opts *o,
                                  double *u, double *lam, double *grad_kappa,
                                  double *loss, admfem_solve_stats st[2],
                                  void *stream);

/* ADCME's SparseSolver (the `\` of a SparseTensor, call sites README.md:54,75;
 * ADCME is not vendored in the reference): u = A \ rhs for A given as COO
 * triplets indices [nnz x 2] (0-based int64) / vv [nnz], duplicates summed
 * in triplet order; rhs, u [n].  A must be symmetric positive definite
 * (Jacobi-PCG; ADMFEM_ERR_BREAKDOWN otherwise).  Device pointers;
 * synchronises the stream.  The last pattern built (per device) is kept: a
 * call whose indices equal it (compared on the device) only sums the new
 * values into it, so the gradient call after a forward solve, and an
 * optimisation loop on a fixed mesh, skip the pattern build.  The kept
 * pattern holds about 24 B per triplet plus the solver's arrays until the
 * next pattern or the process end; calls of this op and its gradient are
 * serialised by a lock.  The solve runs in reverse Cuthill-McKee order of
 * the pattern when that narrows its bandwidth (rhs and u stay in the
 * caller's numbering). */
int admfem_sparse_solve(const long long *indices, const double *vv, int64_t nnz,
                        int n, const double *rhs, double *u,
                        const admfem_solve_opts *o, admfem_solve_stats *st,
                        void *stream);
/* SparseSolverGrad: lambda = A^-T grad_u (= A^-1 grad_u, A symmetric);
 * grad_vv[k] = -lambda[i_k] * u[j_k]; grad_rhs = lambda. */
int admfem_sparse_solve_grad(const long long *indices, const double *vv,
                             int64_t nnz, int n, const double *u,
                             const double *grad_u, double *grad_vv,
                             double *grad_rhs, const admfem_solve_opts *o,
                             admfem_solve_stats *st, void *stream);

/* Instrumentation (no reference counterpart): the pattern the general
 * sparse solve keeps (the last one built on any device): info[0..3] = n,
 * triplets it was built from, distinct (row, col) slots, and 1 when the
 * solver works in reverse Cuthill-McKee order of it (0: the caller's order,
 * kept when it is already banded).  Zeros before the first solve. */
int admfem_sparse_solve_info(int64_t info[4]);

/* Instrumentation (no reference counterpart): launches the solver's step
 * kernel (the SpMV with the iteration's vector updates fused in; FP64 or
 * FP32 matrix per `precision`) `repeat` times back to back on `stream`, on
 * the system's work vectors, for timing the dominant kernel.  *bytes (host,
 * may be NULL) receives its algorithmic bytes per launch, one canonical CSR
 * CG iteration: nnz*(s+4) + 4*(n+1) + 2*n*s for the SpMV plus 6*n*s for the
 * vector updates, with s the value size. */
int admfem_system_spmv_repeat(admfem_system s, int precision, int repeat,
                              double *bytes, void *stream);

/* ======================================================================
 * E. Serving: continuous batching of forward+adjoint requests under an HBM
 *    budget.  No reference counterpart (the reference runs one solve at a
 *    time on the host); each request is the reference's forward + gradient
 *    pipeline (compute_fem_* -> impose_Dirichlet -> K\F -> gradients) for
 *    L = sum_k (u[obs_k] - obs_val_k)^2.
 *
 *    Requests are admitted when they have arrived and their device memory
 *    fits the budget; admitted systems advance together, one batched launch
 *    per CG step for all of them (mixed precision: FP32 inner iterations,
 *    FP64 residual replacement); a system that converges moves on (forward
 *    -> adjoint -> gradient -> done) and frees its memory for the next
 *    arrivals.  Results do not depend on what else is in the batch.
 * ==================================================================== */

typedef struct admfem_server_s *admfem_server;

#define ADMFEM_MODE_BATCHED 0     /* continuous batching, mixed precision */
#define ADMFEM_MODE_SEQ_FP64 1    /* one request at a time, FP64 solver */

typedef struct {
  double hbm_budget_bytes; /* device memory for admitted requests */
  int max_active;          /* cap on concurrently admitted requests */
  int mode;                /* ADMFEM_MODE_* */
  double rtol;             /* true relative residual of every solve */
  int max_iter;
  double replace_delta;    /* mixed: residual replacement factor (0: 0.3) */
  int chunk;               /* CG iterations between host checks (0: 16;
                              rounded up to a multiple of 8) */
  int use_graph;           /* replay each chunk as a captured HIP graph */
} admfem_server_opts;

#define ADMFEM_SRC_HOST 0    /* inputs are host pointers */
#define ADMFEM_SRC_DEVICE 1  /* inputs are device pointers */
#define ADMFEM_SRC_SYNTH 2   /* inputs generated on the device from `seed` */

typedef struct {
  int topo;                /* id from admfem_server_add_topology */
  int kind;                /* 0: Poisson (K u = F); 1: heat step
                              ((M + dt K) u = M u_prev + dt F) */
  double dt;               /* heat step */
  double arrival;          /* seconds after admfem_server_run starts */
  int source;              /* ADMFEM_SRC_* */
  /* HOST / DEVICE inputs (user node / dof order; must stay valid until the
   * request is admitted): coords nnode x 2 (NULL: the topology's own),
   * kappa/f [ngauss], u_prev [n] (heat; n = the system's rows: nnode for
   * P1, ndof for P2), obs_idx (0-based dof ids, distinct) / obs_val.
   * Missing inputs and, for HOST inputs, indices outside [0, n) or
   * repeated fail admfem_server_submit with ADMFEM_ERR_ARG; for DEVICE
   * inputs the request finishes with status ADMFEM_ERR_ARG. */
  const double *coords, *kappa, *f, *u_prev;
  const int *obs_idx;
  const double *obs_val;
  int nobs;
  /* SYNTH: interior vertices jittered by U[-a, a]^2 (a from the topology),
   * kappa = exp(GRF) (32 random Fourier modes, sigma 1, length 0.1) and
   * f ~ U[-1, 1] at the Gauss points, u_prev a GRF at the dofs' nodes
   * (heat; a P2 edge dof: its edge's midpoint), 1% of the interior dofs
   * observed with obs_val = 0; all from `seed`. */
  unsigned long long seed;
  /* host outputs (may be NULL), valid once the request is done */
  double *u_out;           /* [n] */
  double *grad_out;        /* [ngauss] dL/dkappa */
} admfem_request;

typedef struct {
  int status;              /* ADMFEM_OK or the failing solve's code;
                              ADMFEM_ERR_BUDGET if it can never fit */
  int done;
  double t_arrival, t_admit, t_done; /* seconds after run start */
  double loss;
  int iters_fwd, iters_adj;
  double relres_fwd, relres_adj;     /* true FP64 relative residuals */
  double device_bytes;
} admfem_result;

int admfem_server_create(const admfem_server_opts *o, admfem_server *out);
int admfem_server_destroy(admfem_server sv);
/* Registers a mesh + boundary set (a system made by admfem_system_create; it
 * must outlive the server).  jitter: absolute amplitude of the SYNTH node
 * jitter.  Returns the topology id. */
int admfem_server_add_topology(admfem_server sv, admfem_system s, double jitter,
                               int *topo_id);
/* Queues a request; *req_id receives its id (0, 1, ...). */
int admfem_server_submit(admfem_server sv, const admfem_request *r, int64_t *req_id);
/* Serves every queued request (arrival times are measured from the first
 * call), returns when all are done or after timeout_s seconds. */
int admfem_server_run(admfem_server sv, double timeout_s);
int admfem_server_result(admfem_server sv, int64_t req_id, admfem_result *out);
/* stats[0..7] = requests done, wall seconds, peak device bytes, peak
 * concurrent requests, batched CG steps, total CG iterations, failed,
 * mean batch size */
int admfem_server_stats(admfem_server sv, double stats[8]);

/* Instrumentation (no reference counterpart): admits one synthetic Poisson
 * request (seed i) on topology topos[i] for i < nsys, as a batch, launches
 * the batched CG step kernel (the SpMV with the iteration's vector updates
 * fused in) over all of them `repeat` times back to back on `stream`, then
 * releases them.  *bytes (may be NULL) receives the algorithmic bytes per
 * launch: the sum over the systems of one canonical CSR CG iteration,
 * nnz*(4+4) + 4*(n+1) + 2*n*4 for the SpMV plus 6*n*4 for the vector
 * updates (FP32 values).  *seconds (may be
 * NULL) receives the average time per launch, from HIP events recorded on
 * `stream` right before the first and after the last launch.  Not to be
 * called while admfem_server_run is running on the same server. */
int admfem_server_spmv_repeat(admfem_server sv, const int *topos, int nsys, int repeat,
                              double *bytes, double *seconds, void *stream);

#ifdef __cplusplus
}
#endif
#endif
