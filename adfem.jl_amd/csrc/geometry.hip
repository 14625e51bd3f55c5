// Element-level kernels: mesh geometry (Gauss points, weights, areas), the
// reference-layout COO operators (FemLaplaceScalar, ComputeFemMassMatrixMfem)
// with their gradient ops, and the load vector (FemSourceScalar) with its
// gradient.  All atomic-free: a COO op's block stages its elements' factors
// in LDS and writes its contiguous run of triplets; the load vector gathers
// per node over its incident elements.  Gauss weights come from the mesh's
// array (geometry_kernel), the same bits as recomputing them per use.
#include "geom.h"

namespace admfem {

namespace {

struct ElemCoords {
  double x[3], y[3];
};

__device__ __forceinline__ ElemCoords load_elem(const int *__restrict__ conn,
                                                int nelem, int e,
                                                const double *__restrict__ x,
                                                const double *__restrict__ y) {
  ElemCoords c;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int v = conn[(size_t)k * nelem + e];
    c.x[k] = x[v];
    c.y[k] = y[v];
  }
  return c;
}

// NNFEM_Mesh::init per element (deps/MFEM/Common.cpp:83-116): Heron area,
// w = ip.weight*area/0.5, Gauss points sum_k hs_k * vertex_k.
__global__ void geometry_kernel(int nelem, QuadRule rule,
                                const int *__restrict__ conn,
                                const double *__restrict__ x,
                                const double *__restrict__ y,
                                double *__restrict__ gx,
                                double *__restrict__ gy,
                                double *__restrict__ gw,
                                double *__restrict__ area) {
#pragma clang fp contract(off)
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  ElemCoords c = load_elem(conn, nelem, e, x, y);
  double A = heron_area(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
  area[e] = A;
  for (int k = 0; k < rule.nq; k++) {
    size_t g = (size_t)rule.nq * e + k;
    double h0 = 1 - rule.xi[k] - rule.eta[k], h1 = rule.xi[k], h2 = rule.eta[k];
    gx[g] = c.x[0] * h0 + c.x[1] * h1 + c.x[2] * h2;
    gy[g] = c.y[0] * h0 + c.y[1] * h1 + c.y[2] * h2;
    gw[g] = gauss_weight(rule.w[k], A);
  }
}

// FemLaplaceScalar_forward (deps/MFEM/FemLaplace1/FemLaplaceScalar.h:3-26)
// and ComputeFemMassMatrix1_forward (ComputeFemMassMatrixMfem.h:4-29):
// triplet t = 9*g + 3*p + q of Gauss point g = nq*e + k, value
// D_pq * coef[g] * w[g] with D_pq = grad phi_p . grad phi_q (Laplace) or
// h_p h_q (mass), w the mesh's Gauss weight (ip.weight * area / 0.5, the
// same bits as the per-element computation).  A block owns CO_E consecutive
// elements: one thread per element puts its 6 distinct D_pq and its user
// vertex ids in LDS, the block's coef and w go to LDS too (each is used by
// 9 triplets), then all threads write the block's contiguous run of
// 9 * nq * CO_E triplets, coalesced, with streaming (non-temporal) stores.
constexpr int CO_NT = 256, CO_E = 64;

template <bool MASS>
__global__ void __launch_bounds__(CO_NT)
    coo_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
               const int *__restrict__ conn_user, const double *__restrict__ x,
               const double *__restrict__ y, const double *__restrict__ gw,
               const double *__restrict__ coef, long long *__restrict__ ind,
               double *__restrict__ vv, bool ind16) {
#pragma clang fp contract(off)
  __shared__ double dd[CO_E][6];        // D_pq, packed upper triangle (Laplace)
  __shared__ int cu[CO_E][3];
  __shared__ double cf[CO_E * MAXQ], cw[CO_E * MAXQ];  // coef[g], w[g] of the block
  const int e0 = blockIdx.x * CO_E;
  const int ne = min(CO_E, nelem - e0);
  const int nq = rule.nq;
  for (int l = threadIdx.x; l < ne * nq; l += CO_NT) {
    const int64_t g = (int64_t)nq * e0 + l;
    cf[l] = coef[g];
    cw[l] = gw[g];
  }
  if ((int)threadIdx.x < ne) {
    const int e = e0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 3; k++) cu[threadIdx.x][k] = conn_user[(size_t)k * nelem + e];
    if (!MASS) {
      const ElemCoords c = load_elem(conn, nelem, e, x, y);
      const ElemGeom G = elem_geom(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
      int i = 0;
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = p; q < 3; q++) dd[threadIdx.x][i++] = G.gx[p] * G.gx[q] + G.gy[p] * G.gy[q];
    }
  }
  __syncthreads();
  const int64_t t0 = (int64_t)9 * nq * e0;
  const int nt = 9 * nq * ne;
  for (int l = threadIdx.x; l < nt; l += CO_NT) {
    const int gl = l / 9, pq = l - 9 * gl;
    const int p = pq / 3, q = pq - 3 * p;
    const int el = gl / nq, k = gl - nq * el;
    const int64_t t = t0 + l;
    double d;
    if (MASS) {
      d = shape(p, rule.xi[k], rule.eta[k]) * shape(q, rule.xi[k], rule.eta[k]);
    } else {
      const int lo = min(p, q), hi = max(p, q);
      d = dd[el][lo == 0 ? hi : (lo == 1 ? 2 + hi : 5)];
    }
    store_triplet(ind, vv, t, cu[el][p], cu[el][q], d * cf[gl] * cw[gl], ind16);
  }
}

// FemLaplaceScalar_backward (FemLaplaceScalar.h:28-54): one thread per Gauss
// point, grad_kappa[g] = sum_pq grad_vv[9g+3p+q] * (D D^T w)_pq.
__global__ void laplace_coo_grad_kernel(int nelem, QuadRule rule,
                                        const int *__restrict__ conn,
                                        const double *__restrict__ x,
                                        const double *__restrict__ y,
                                        const double *__restrict__ gvv,
                                        double *__restrict__ gk) {
#pragma clang fp contract(off)
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  ElemCoords c = load_elem(conn, nelem, e, x, y);
  ElemGeom G = elem_geom(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
  double w = gauss_weight(rule.w[k], G.area);
  double v = 0.0;
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      double d = (G.gx[p] * G.gx[q] + G.gy[p] * G.gy[q]) * w;
      v += gvv[9 * g + 3 * p + q] * d;
    }
  gk[g] = v;
}

// ComputeFemMassMatrix1_backward (ComputeFemMassMatrixMfem.h:31-56), output
// overwritten (the op zeroes it first, ComputeFemMassMatrixMfem.cpp:139).
__global__ void mass_coo_grad_kernel(int nelem, QuadRule rule,
                                     const int *__restrict__ conn,
                                     const double *__restrict__ x,
                                     const double *__restrict__ y,
                                     const double *__restrict__ gvv,
                                     double *__restrict__ gr) {
#pragma clang fp contract(off)
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  ElemCoords c = load_elem(conn, nelem, e, x, y);
  double A = heron_area(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
  double w = gauss_weight(rule.w[k], A);
  double v = 0.0;
  for (int l = 0; l < 3; l++)
    for (int s = 0; s < 3; s++) {
      double nn = shape(l, rule.xi[k], rule.eta[k]) * shape(s, rule.xi[k], rule.eta[k]) * w;
      v += nn * gvv[9 * g + 3 * l + s];
    }
  gr[g] = v;
}

// FemSourceScalar_forward (FemSourceScalar.h:4-15) as a per-node gather:
// node r sums f[g]*h_p(g)*w[g] over its incident elements in ascending
// element order and their Gauss points in order -- the same sequence of
// additions the reference's element loop applies to rhs[dof].
template <int NQ>
__global__ void source_kernel(int n, QuadRule rule, const int *__restrict__ n2e_ptr,
                              const int *__restrict__ n2e_elem,
                              const uint32_t *__restrict__ slots,
                              const int *__restrict__ iperm, const double *__restrict__ gw,
                              const double *__restrict__ f, double *__restrict__ rhs_user,
                              int accumulate) {
#pragma clang fp contract(off)
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int u = iperm[r];
  double acc = accumulate ? rhs_user[u] : 0.0;
  const int t1 = n2e_ptr[r + 1];
#pragma unroll 4
  for (int t = n2e_ptr[r]; t < t1; t++) {  // unrolled: the incidences' loads overlap
    const int e = n2e_elem[t];
    const int p = slots[t] & 3u;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const size_t g = (size_t)NQ * e + k;
      acc += f[g] * shape(p, rule.xi[k], rule.eta[k]) * gw[g];
    }
  }
  rhs_user[u] = acc;
}

// FemSourceScalar_backward (FemSourceScalar.h:17-32).
__global__ void source_grad_kernel(int nelem, QuadRule rule,
                                   const int *__restrict__ conn_user,
                                   const double *__restrict__ gw,
                                   const double *__restrict__ grhs,
                                   double *__restrict__ gf) {
#pragma clang fp contract(off)
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  const double w = gw[g];
  double v = 0.0;
  for (int r = 0; r < 3; r++)
    v += shape(r, rule.xi[k], rule.eta[k]) * w * grhs[conn_user[(size_t)r * nelem + e]];
  gf[g] = v;
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_geometry(const Mesh &m, hipStream_t st) {
  geometry_kernel<<<nblk(m.nelem, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.x.p, m.y.p, m.gx.p, m.gy.p, m.gw.p, m.area.p);
  HIPCHECK(hipGetLastError());
}

void launch_laplace_coo(const Mesh &m, const double *kappa, long long *ind,
                        double *vv, hipStream_t st) {
  if (m.degree == 2) return launch_p2_coo(m, false, kappa, ind, vv, st);
  coo_kernel<false><<<nblk(m.nelem, CO_E), CO_NT, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_user.p, m.x.p, m.y.p, m.gw.p, kappa, ind, vv,
      ((uintptr_t)ind & 15) == 0);
  HIPCHECK(hipGetLastError());
}

void launch_laplace_coo_grad(const Mesh &m, const double *gvv, double *gk,
                             hipStream_t st) {
  if (m.degree == 2) return launch_p2_coo_grad(m, false, gvv, gk, st);
  laplace_coo_grad_kernel<<<nblk(m.ngauss, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.x.p, m.y.p, gvv, gk);
  HIPCHECK(hipGetLastError());
}

void launch_mass_coo(const Mesh &m, const double *rho, long long *ind,
                     double *vv, hipStream_t st) {
  if (m.degree == 2) return launch_p2_coo(m, true, rho, ind, vv, st);
  coo_kernel<true><<<nblk(m.nelem, CO_E), CO_NT, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_user.p, m.x.p, m.y.p, m.gw.p, rho, ind, vv,
      ((uintptr_t)ind & 15) == 0);
  HIPCHECK(hipGetLastError());
}

void launch_mass_coo_grad(const Mesh &m, const double *gvv, double *gr,
                          hipStream_t st) {
  if (m.degree == 2) return launch_p2_coo_grad(m, true, gvv, gr, st);
  mass_coo_grad_kernel<<<nblk(m.ngauss, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.x.p, m.y.p, gvv, gr);
  HIPCHECK(hipGetLastError());
}

void launch_source(const Mesh &m, const double *f, double *rhs_user,
                   bool accumulate, hipStream_t st) {
  if (m.degree == 2) return launch_p2_source(m, f, rhs_user, accumulate, st);
  auto go = [&](auto kern) {
    kern<<<nblk(m.nnode, 256), 256, 0, st>>>(m.nnode, m.rule, m.d_n2e_ptr.p, m.d_n2e_elem.p,
                                             m.d_slots_full.p, m.d_iperm.p, m.gw.p, f, rhs_user,
                                             accumulate ? 1 : 0);
  };
  if (m.rule.nq == 3) go(source_kernel<3>);
  else if (m.rule.nq == 6) go(source_kernel<6>);
  else go(source_kernel<1>);
  HIPCHECK(hipGetLastError());
}

void launch_source_grad(const Mesh &m, const double *grhs_user, double *gf,
                        hipStream_t st) {
  if (m.degree == 2) return launch_p2_source_grad(m, grhs_user, gf, st);
  source_grad_kernel<<<nblk(m.ngauss, 256), 256, 0, st>>>(m.nelem, m.rule, m.conn_user.p,
                                                         m.gw.p, grhs_user, gf);
  HIPCHECK(hipGetLastError());
}

}  // namespace admfem
