// P2 (degree 2) element kernels: the COO operators of FemLaplaceScalar and
// ComputeFemMassMatrixMfem, the load vector of FemSourceScalar, the
// DofToGaussPoints / FemGrad maps, the matrix-free Laplace term, and their
// gradient ops, for the 6-dof quadratic triangle of NNFEM_Mesh::init with
// degree 2 (deps/MFEM/Common.cpp:46-72): dofs are the three vertices, then
// nnode + the MFEM id of edges (0,1), (1,2), (2,0).
//
// The basis is MFEM's order-2 H1 triangle, nodal at the vertices and edge
// midpoints: with barycentric l0 = 1 - xi - eta, l1 = xi, l2 = eta,
//   phi_k = l_k (2 l_k - 1) (k < 3), phi_3 = 4 l0 l1, phi_4 = 4 l1 l2,
//   phi_5 = 4 l2 l0,
// and its physical gradients combine the constant grad l_k of the affine map
// (CalcPhysDShape, Common.cpp:119-126):
//   grad phi_k = (4 l_k - 1) grad l_k, grad phi_3 = 4 (l1 grad l0 + l0 grad l1), ...
//
// Layouts as the P1 kernels: Gauss arrays in user element order
// (g = nq e + k), dof arrays in user dof order, COO triplet
// t = 36 g + 6 p + q.  Maps to Gauss points run one thread per Gauss point;
// the COO ops stage each block's basis gradients in LDS and write its run of
// triplets; maps to dofs gather per dof over its incident (element, local
// dof) pairs in ascending element order -- the reference's addition order --
// with one writer per output and no atomics.  Gauss weights come from the
// mesh's array (the same bits as recomputing them).
#include "p2basis.h"

namespace admfem {

namespace {

// FemLaplaceScalar_forward / ComputeFemMassMatrix1_forward, P2
// (FemLaplaceScalar.h:3-26, ComputeFemMassMatrixMfem.h:4-29): triplet
// t = 36 g + 6 p + q, value D_pq * coef[g] * w[g].  A block owns P2_E
// consecutive elements: one thread per (element, Gauss point) puts the six
// basis gradients in LDS (Laplace), the block's coef and w go to LDS too,
// then all threads write the block's contiguous run of triplets, coalesced,
// with streaming stores.
constexpr int P2_NT = 256, P2_E = 16;

template <bool MASS>
__global__ void __launch_bounds__(P2_NT)
    p2_coo_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
                  const int *__restrict__ conn_dof, const double *__restrict__ x,
                  const double *__restrict__ y, const double *__restrict__ gw,
                  const double *__restrict__ coef, long long *__restrict__ ind,
                  double *__restrict__ vv, bool ind16) {
#pragma clang fp contract(off)
  __shared__ double lgx[P2_E][MAXQ][6], lgy[P2_E][MAXQ][6];
  __shared__ int cd[P2_E][6];
  __shared__ double cf[P2_E * MAXQ], cw[P2_E * MAXQ];  // coef[g], w[g] of the block
  const int nq = rule.nq;
  const int e0 = blockIdx.x * P2_E;
  const int ne = min(P2_E, nelem - e0);
  for (int l = threadIdx.x; l < ne * nq; l += P2_NT) {
    const int64_t g = (int64_t)nq * e0 + l;
    cf[l] = coef[g];
    cw[l] = gw[g];
  }
  for (int i = threadIdx.x; i < ne * 6; i += P2_NT) {
    const int el = i / 6, r = i - 6 * el;
    cd[el][r] = conn_dof[(size_t)r * nelem + e0 + el];
  }
  if (!MASS && (int)threadIdx.x < ne * nq) {
    const int el = threadIdx.x / nq, k = threadIdx.x - nq * el;
    const ElemGeom G = p2_geom(conn, nelem, e0 + el, x, y);
    double gx[6], gy[6];
    p2_grad(rule.xi[k], rule.eta[k], G, gx, gy);
#pragma unroll
    for (int r = 0; r < 6; r++) {
      lgx[el][k][r] = gx[r];
      lgy[el][k][r] = gy[r];
    }
  }
  __syncthreads();
  const int64_t t0 = (int64_t)36 * nq * e0;
  const int nt = 36 * nq * ne;
  for (int l = threadIdx.x; l < nt; l += P2_NT) {
    const int gl = l / 36, pq = l - 36 * gl;
    const int p = pq / 6, q = pq - 6 * p;
    const int el = gl / nq, k = gl - nq * el;
    const int64_t t = t0 + l;
    double v;
    if (MASS) {
      double h[6];
      p2_shape(rule.xi[k], rule.eta[k], h);
      v = h[p] * h[q] * cf[gl] * cw[gl];
    } else {
      v = (lgx[el][k][p] * lgx[el][k][q] + lgy[el][k][p] * lgy[el][k][q]) * cf[gl] * cw[gl];
    }
    store_triplet(ind, vv, t, cd[el][p], cd[el][q], v, ind16);
  }
}

// The two backward ops (FemLaplaceScalar.h:28-54, ComputeFemMassMatrixMfem.h:
// 31-56): one thread per Gauss point, sum over its 36 triplets in order.
template <bool MASS>
__global__ void __launch_bounds__(256) p2_coo_grad_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
                                   const double *__restrict__ x, const double *__restrict__ y,
                                   const double *__restrict__ gw,
                                   const double *__restrict__ gvv, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  const int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  // the point's 36 incoming gradients, all requested before the sum (as
  // 16-byte loads: 36 g doubles is 16-byte aligned)
  double t[36];
  const double2 *src = reinterpret_cast<const double2 *>(gvv + 36 * g);
#pragma unroll
  for (int i = 0; i < 18; i++) {
    const double2 v2 = src[i];
    t[2 * i] = v2.x;
    t[2 * i + 1] = v2.y;
  }
  // keep the loads above: left to itself the scheduler moves them down to
  // their uses, a few at a time, each group waiting for the last
  __builtin_amdgcn_sched_barrier(0);
  const ElemGeom G = p2_geom(conn, nelem, e, x, y);
  const double w = gw[g];
  double a[6], b[6];
  if (MASS) {
    p2_shape(rule.xi[k], rule.eta[k], a);
  } else {
    p2_grad(rule.xi[k], rule.eta[k], G, a, b);
  }
  double v = 0.0;
#pragma unroll
  for (int p = 0; p < 6; p++)
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const double n = MASS ? a[p] * a[q] * w : (a[p] * a[q] + b[p] * b[q]) * w;
      v += t[6 * p + q] * n;
    }
  out[g] = v;
}

enum { G_SOURCE_T = 0, G_DOF_TO_GAUSS = 1, G_FEM_GRAD = 2, G_LAP_NU = 3 };

// Per Gauss point:
//   G_SOURCE_T      FemSourceScalar_backward: out[g] = sum_r h_r w grad_rhs[dof_r]
//   G_DOF_TO_GAUSS  DofToGaussPointsMfem_forward: out[g] = sum_r u[dof_r] h_r
//   G_FEM_GRAD      FemGradMfem_forward: out[2g], out[2g+1] = sum_r u[dof_r] grad h_r
//   G_LAP_NU        ComputeLaplaceTermMfem_backward, nu part:
//                   out[g] = w grad(a) . grad(b) (the term is linear in nu)
template <int OP>
__global__ void p2_gauss_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
                                const int *__restrict__ conn_dof, const double *__restrict__ x,
                                const double *__restrict__ y, const double *__restrict__ gw,
                                const double *__restrict__ a,
                                const double *__restrict__ b, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  const int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  int d[6];
#pragma unroll
  for (int r = 0; r < 6; r++) d[r] = conn_dof[(size_t)r * nelem + e];
  if (OP == G_DOF_TO_GAUSS) {
    double h[6];
    p2_shape(rule.xi[k], rule.eta[k], h);
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < 6; r++) v += a[d[r]] * h[r];
    out[g] = v;
    return;
  }
  const double w = gw[g];
  if (OP == G_SOURCE_T) {
    double h[6];
    p2_shape(rule.xi[k], rule.eta[k], h);
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < 6; r++) v += h[r] * w * a[d[r]];
    out[g] = v;
    return;
  }
  const ElemGeom G = p2_geom(conn, nelem, e, x, y);
  double gx[6], gy[6];
  p2_grad(rule.xi[k], rule.eta[k], G, gx, gy);
  if (OP == G_FEM_GRAD) {
    double vx = 0.0, vy = 0.0;
#pragma unroll
    for (int r = 0; r < 6; r++) vx += a[d[r]] * gx[r];
#pragma unroll
    for (int r = 0; r < 6; r++) vy += a[d[r]] * gy[r];
    out[2 * g] = vx;
    out[2 * g + 1] = vy;
  } else {  // G_LAP_NU
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
    for (int r = 0; r < 6; r++) {
      ax += gx[r] * a[d[r]];
      ay += gy[r] * a[d[r]];
      bx += gx[r] * b[d[r]];
      by += gy[r] * b[d[r]];
    }
    out[g] = w * (ax * bx + ay * by);
  }
}

enum { D_SOURCE = 0, D_DOF_TO_GAUSS_T = 1, D_FEM_GRAD_T = 2, D_LAPLACE = 3 };

// Per dof, over its incident (element, local dof p) pairs, ascending element:
//   D_SOURCE          FemSourceScalar_forward: sum_k f[g] h_p(k) w_k
//   D_DOF_TO_GAUSS_T  DofToGaussPointsMfem_backward: sum_k in[g] h_p(k)
//   D_FEM_GRAD_T      FemGradMfem_backward: sum_k in[2g] dh_p/dx + in[2g+1] dh_p/dy
//   D_LAPLACE         ComputeLaplaceTermMfem_forward (= K(nu) u), with u and
//                     the thread order in the internal dof order (vertex
//                     dofs permuted like the nodes, edge dofs as they are),
//                     so an incidence's six values are local gathers:
//                     sum_k nu_k w_k grad h_p . grad u_e(k)
template <int OP, int NQ>
__global__ void __launch_bounds__(256) p2_dof_kernel(int ndof, int nnode, const int *__restrict__ iperm, int nelem,
                              QuadRule rule, const int *__restrict__ conn,
                              const int *__restrict__ conn_dof, const double *__restrict__ x,
                              const double *__restrict__ y, const double *__restrict__ gw,
                              const int *__restrict__ d2e_ptr, const int *__restrict__ d2e,
                              const double *__restrict__ in, const double *__restrict__ nu,
                              double *__restrict__ out, int accumulate) {
#pragma clang fp contract(off)
  constexpr bool CHEAP = OP == D_DOF_TO_GAUSS_T || OP == D_SOURCE;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ndof) return;
  const int dof = (OP == D_LAPLACE && r < nnode) ? iperm[r] : r;  // user dof id
  double acc = accumulate ? out[dof] : 0.0;
  const int t1 = d2e_ptr[dof + 1];
  // the cheap ops are unrolled over incidences so that their loads overlap;
  // the gradient ops keep one incidence in flight (registers, no spills)
#pragma unroll(CHEAP ? 2 : 1)
  for (int t = d2e_ptr[dof]; t < t1; t++) {
    const int e = d2e[t] >> 3, p = d2e[t] & 7;
    if (OP == D_DOF_TO_GAUSS_T) {
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        double h[6];
        p2_shape(rule.xi[k], rule.eta[k], h);
        acc += in[(size_t)NQ * e + k] * h[p];
      }
    } else if (OP == D_SOURCE) {
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        double h[6];
        p2_shape(rule.xi[k], rule.eta[k], h);
        const size_t g = (size_t)NQ * e + k;
        acc += in[g] * h[p] * gw[g];
      }
    } else if (OP == D_FEM_GRAD_T) {
      const ElemGeom G = p2_geom(conn, nelem, e, x, y);
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        double gx[6], gy[6];
        p2_grad(rule.xi[k], rule.eta[k], G, gx, gy);
        const size_t g = (size_t)NQ * e + k;
        acc += in[2 * g] * gx[p] + in[2 * g + 1] * gy[p];
      }
    } else {  // D_LAPLACE
      const ElemGeom G = p2_geom(conn, nelem, e, x, y);
      double ue[6];
#pragma unroll
      for (int q = 0; q < 6; q++)
        ue[q] = in[q < 3 ? conn[(size_t)q * nelem + e] : conn_dof[(size_t)q * nelem + e]];
#pragma unroll
      for (int k = 0; k < NQ; k++) {
        double gx[6], gy[6];
        p2_grad(rule.xi[k], rule.eta[k], G, gx, gy);
        double ux = 0.0, uy = 0.0;
#pragma unroll
        for (int q = 0; q < 6; q++) {
          ux += gx[q] * ue[q];
          uy += gy[q] * ue[q];
        }
        const size_t g = (size_t)NQ * e + k;
        acc += nu[g] * gw[g] * (gx[p] * ux + gy[p] * uy);
      }
    }
  }
  out[dof] = acc;
}

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// dst = src in the internal dof order: vertex dof r <- src[iperm[r]], edge
// dofs unchanged
__global__ void p2_permute_dofs(int ndof, int nnode, const int *__restrict__ iperm,
                                const double *__restrict__ src, double *__restrict__ dst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < ndof) dst[r] = src[r < nnode ? iperm[r] : r];
}

template <int OP>
void gauss_op(const Mesh &m, const double *a, const double *b, double *out, hipStream_t st) {
  p2_gauss_kernel<OP><<<nb(m.ngauss, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_dof.p, m.x.p, m.y.p, m.gw.p, a, b, out);
  HIPCHECK(hipGetLastError());
}

template <int OP>
void dof_op(const Mesh &m, const double *in, const double *nu, double *out, bool accumulate,
            hipStream_t st) {
  auto go = [&](auto kern) {
    kern<<<nb(m.ndof, 256), 256, 0, st>>>((int)m.ndof, m.nnode, m.d_iperm.p, m.nelem, m.rule,
                                          m.conn.p, m.conn_dof.p,
                                          m.x.p, m.y.p, m.gw.p, m.d_d2e_ptr.p, m.d_d2e.p, in, nu,
                                          out, accumulate ? 1 : 0);
  };
  if (m.rule.nq == 6) go(p2_dof_kernel<OP, 6>);
  else if (m.rule.nq == 3) go(p2_dof_kernel<OP, 3>);
  else go(p2_dof_kernel<OP, 1>);
  HIPCHECK(hipGetLastError());
}

}  // namespace

void launch_p2_coo(const Mesh &m, bool mass, const double *coef, long long *ind, double *vv,
                   hipStream_t st) {
  const unsigned grid = nb(m.nelem, P2_E);
  const bool ind16 = ((uintptr_t)ind & 15) == 0;
  if (mass)
    p2_coo_kernel<true><<<grid, P2_NT, 0, st>>>(m.nelem, m.rule, m.conn.p, m.conn_dof.p, m.x.p,
                                                m.y.p, m.gw.p, coef, ind, vv, ind16);
  else
    p2_coo_kernel<false><<<grid, P2_NT, 0, st>>>(m.nelem, m.rule, m.conn.p, m.conn_dof.p, m.x.p,
                                                 m.y.p, m.gw.p, coef, ind, vv, ind16);
  HIPCHECK(hipGetLastError());
}

void launch_p2_coo_grad(const Mesh &m, bool mass, const double *gvv, double *out,
                        hipStream_t st) {
  if (mass)
    p2_coo_grad_kernel<true><<<nb(m.ngauss, 256), 256, 0, st>>>(m.nelem, m.rule, m.conn.p,
                                                                 m.x.p, m.y.p, m.gw.p, gvv, out);
  else
    p2_coo_grad_kernel<false><<<nb(m.ngauss, 256), 256, 0, st>>>(m.nelem, m.rule, m.conn.p,
                                                                  m.x.p, m.y.p, m.gw.p, gvv, out);
  HIPCHECK(hipGetLastError());
}

void launch_p2_source(const Mesh &m, const double *f, double *rhs, bool accumulate,
                      hipStream_t st) {
  dof_op<D_SOURCE>(m, f, nullptr, rhs, accumulate, st);
}

void launch_p2_source_grad(const Mesh &m, const double *grhs, double *gf, hipStream_t st) {
  gauss_op<G_SOURCE_T>(m, grhs, nullptr, gf, st);
}

void launch_p2_dof_to_gauss(const Mesh &m, const double *u, double *out, hipStream_t st) {
  gauss_op<G_DOF_TO_GAUSS>(m, u, nullptr, out, st);
}

void launch_p2_dof_to_gauss_grad(const Mesh &m, const double *g, double *gu, hipStream_t st) {
  dof_op<D_DOF_TO_GAUSS_T>(m, g, nullptr, gu, false, st);
}

void launch_p2_fem_grad(const Mesh &m, const double *u, double *out, hipStream_t st) {
  gauss_op<G_FEM_GRAD>(m, u, nullptr, out, st);
}

void launch_p2_fem_grad_grad(const Mesh &m, const double *g, double *gu, hipStream_t st) {
  dof_op<D_FEM_GRAD_T>(m, g, nullptr, gu, false, st);
}

void launch_p2_laplace_term(const Mesh &m, const double *u, const double *nu, double *out,
                            hipStream_t st) {
  double *u_int = nullptr;  // u in the internal dof order, a stream-ordered temporary
  HIPCHECK(hipMallocAsync((void **)&u_int, sizeof(double) * m.ndof, st));
  p2_permute_dofs<<<nb(m.ndof, 256), 256, 0, st>>>((int)m.ndof, m.nnode, m.d_iperm.p, u, u_int);
  dof_op<D_LAPLACE>(m, u_int, nu, out, false, st);
  HIPCHECK(hipFreeAsync(u_int, st));
}

void launch_p2_laplace_term_nu_grad(const Mesh &m, const double *go, const double *u,
                                    double *gnu, hipStream_t st) {
  gauss_op<G_LAP_NU>(m, go, u, gnu, st);
}

}  // namespace admfem
