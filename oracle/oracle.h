/*
 * CPU oracle for the AdFem P1/P2 forward/adjoint path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU loops, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the product library links, loads or calls this code.
 *
 * Each function names the reference file:line it restates.  All arrays are
 * host memory, indices 0-based unless stated, doubles are FP64.
 */
#ifndef ADMFEM_ORACLE_H
#define ADMFEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quadrature rule of the reference triangle (area 1/2) for an integration
 * order (0-2: centroid / 3-point; 3-4: 6-point): fills xi/eta/weight
 * (capacity 8) and returns the point count, or -1 for an order this
 * restatement does not carry. */
int orc_tri_rule(int order, double *xi, double *eta, double *weight);

/*
 * Mesh precompute, restating deps/MFEM/Common.cpp:20-142 (NNFEM_Mesh::init)
 * for P1 elements.
 *   coords  nnode x 2 row-major
 *   elems   nelem x 3 row-major, 0-based; REWRITTEN in place with the
 *           orientation fix of FinalizeTriMesh(1,0,true) (Common.cpp:40):
 *           a triangle with negative Jacobian determinant has its first two
 *           vertices swapped.
 *   outputs (nq = points of the rule, ngauss = nq*nelem, g = nq*e + k):
 *     gx, gy [ngauss]      Gauss points (Common.cpp:110-111)
 *     w      [ngauss]      ip.weight*area/0.5 (Common.cpp:116)
 *     area   [nelem]       Heron's formula (Common.cpp:9-15, 83)
 *     h, hx, hy [ngauss*3] shape values / physical gradients, [g][r]
 *                          (Common.cpp:119-126, CalcPhysShape/DShape)
 * Returns nq, or -1 on an unsupported order.
 */
int orc_mesh_init(const double *coords, int nnode, int *elems, int nelem,
                  int order, double *gx, double *gy, double *w, double *area,
                  double *h, double *hx, double *hy);

/* P2 precompute (Common.cpp:20-142 with degree 2): elems orientation-fixed
 * (orc_mesh_init first), elem_edge[3e+j] the MFEM id of element edge j
 * ((0,1),(1,2),(2,0), from orc_mesh_edges).  Outputs dofs [nelem x 6] (the
 * vertices, then nnode + edge id; Common.cpp:63-72) and h/hx/hy
 * [ngauss*6] (CalcPhysShape / CalcPhysDShape of MFEM's order-2 H1 triangle:
 * vertex functions l(2l-1), then 4 l0 l1, 4 l1 l2, 4 l2 l0).  Gauss points,
 * weights and areas are those of orc_mesh_init.  Returns nq or -1. */
int orc_mesh_init_p2(const double *coords, int nnode, const int *elems, int nelem,
                     int order, const int64_t *elem_edge, int *dofs, double *h,
                     double *hx, double *hy);

/* Edge table of the finalized mesh (MFEM numbering: first appearance over the
 * oriented elements' edges (0,1),(1,2),(2,0)); returns nedges and, when out is
 * non-NULL, writes [min+1 ...; max+1 ...] (2*nedges, 1-based) as
 * deps/MFEM/Common.cpp:133-140 does; elem_edge (optional, 3*nelem) receives
 * each element edge's 0-based id.  nedges upper bound is 3*nelem. */
int64_t orc_mesh_edges(const int *elems, int nelem, int nnode, int64_t *out,
                       int64_t *elem_edge);

/* In the element ops below, dofs is nelem x nd (the element's global dof ids,
 * 0-based; for P1 nd = 3 and dofs = elems) and h/hx/hy are [ngauss*nd]. */

/* FemLaplaceScalar_forward (deps/MFEM/FemLaplace1/FemLaplaceScalar.h:3-26):
 * nd*nd triplets per Gauss point, p-major then q. indices [2*nd*nd*ngauss]. */
void orc_laplace_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                         const double *hx, const double *hy,
                         const double *kappa, int64_t *indices, double *vv);
/* FemLaplaceScalar_backward (FemLaplaceScalar.h:28-54). */
void orc_laplace_backward(int nd, int nelem, int nq, const double *w, const double *hx,
                          const double *hy, const double *grad_vv,
                          double *grad_kappa);

/* ComputeFemMassMatrix1_forward/_backward
 * (deps/MFEM/ComputeFemMassMatrix1/ComputeFemMassMatrixMfem.h:4-56).
 * The backward accumulates into grad_rho (caller zeroes, as the op does). */
void orc_mass_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                      const double *h, const double *rho, int64_t *indices,
                      double *vv);
void orc_mass_backward(int nd, int nelem, int nq, const double *w, const double *h,
                       const double *grad_vv, double *grad_rho);

/* FemSourceScalar_forward/_backward (deps/MFEM/FemSource1/FemSourceScalar.h:4-32).
 * Forward accumulates into rhs (caller zeroes). */
void orc_source_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                        const double *h, const double *f, double *rhs);
void orc_source_backward(const int *dofs, int nd, int nelem, int nq, const double *w,
                         const double *h, const double *grad_rhs,
                         double *grad_f);

/* ImposeDirichlet_forward + Copy (deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:27-60).
 * bd is 1-based; later duplicates of a node overwrite earlier values (std::map
 * assignment).  Output capacity must be sN + bdN.  Returns the output count. */
int64_t orc_dirichlet_forward(const int64_t *indices, const double *vv,
                              int64_t sN, const int64_t *bd, const double *bdval,
                              int bdN, const double *rhs, int N,
                              int64_t *oindices, double *ov, double *orhs);
/* ImposeDirichlet_backward (ImposeDirichlet.h:63-93).  grad_vv_ipt, grad_rhs_ipt
 * and grad_bdval are zeroed here first (the TF op hands zero-initialised
 * outputs to the reference loop). */
void orc_dirichlet_backward(double *grad_vv_ipt, double *grad_rhs_ipt,
                            double *grad_bdval, const double *grad_ov,
                            const double *grad_orhs, const int64_t *indices,
                            const double *vv, int64_t sN, const int64_t *bd,
                            const double *bdval, int bdN, int N);

/* COO (triplets in order) -> CSR with duplicates summed in triplet order and
 * explicit zeros kept; column indices ascending.  row_ptr [N+1]; col/val
 * capacity sN.  Returns nnz. */
int64_t orc_coo_to_csr(const int64_t *indices, const double *vv, int64_t sN,
                       int N, int64_t *row_ptr, int32_t *col, double *val);

/* FemToGaussPointsMfem_forward/_backward
 * (deps/MFEM/FemToGaussPoints/FemToGaussPointsMfem.h:6-33): out[g] =
 * sum_r u[node_r] hs(r, j).  DofToGaussPointsMfem (DofToGaussPointsMfem.h:
 * 6-32) is the same loop over the element's dofs with h, identical for P1.
 * The backward accumulates into grad_u (caller zeroes). */
void orc_fem_to_gauss_forward(const int *dofs, int nd, int nelem, int nq, const double *h,
                              const double *u, double *out);
void orc_fem_to_gauss_backward(const int *dofs, int nd, int nelem, int nq, const double *h,
                               const double *grad_out, double *grad_u);

/* FemGradMfem_forward/_backward (deps/MFEM/FemGrad/FemGradMfem.h:8-41):
 * out[2g] = sum_r u[dof_r] hx(r, j), out[2g+1] = sum_r u[dof_r] hy(r, j).
 * The backward accumulates into grad_u (caller zeroes). */
void orc_fem_grad_forward(const int *dofs, int nd, int nelem, int nq, const double *hx,
                          const double *hy, const double *u, double *out);
void orc_fem_grad_backward(const int *dofs, int nd, int nelem, int nq, const double *hx,
                           const double *hy, const double *grad_out, double *grad_u);

/* ComputeLaplaceTermMfem_forward/_backward
 * (deps/MFEM/ComputeLaplaceTermMfem/ComputeLaplaceTermMfem.h:4-39):
 * out[dof_p] += nu_j sum_q (hx_p hx_q + hy_p hy_q) u[dof_q] w_j.  Both
 * accumulate (caller zeroes out, grad_nu and grad_u). */
void orc_laplace_term_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                              const double *hx, const double *hy, const double *nu,
                              const double *u, double *out);
void orc_laplace_term_backward(const int *dofs, int nd, int nelem, int nq, const double *w,
                               const double *hx, const double *hy, const double *nu,
                               const double *u, const double *grad_out, double *grad_nu,
                               double *grad_u);

/* Boundary nodes of a triangle mesh: vertices of edges used by exactly one
 * element (src/MFEM/MCore.jl:385-406, 439-450), returned ascending, 1-based.
 * out capacity nnode.  Returns the count. */
int orc_bcnode(const int *elems, int nelem, int nnode, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
