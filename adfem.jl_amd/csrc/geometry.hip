// Element-level kernels: mesh geometry (Gauss points, weights, areas), the
// reference-layout COO operators (FemLaplaceScalar, ComputeFemMassMatrixMfem)
// with their gradient ops, and the load vector (FemSourceScalar) with its
// gradient.  All atomic-free: COO ops write one triplet per thread; the
// load vector gathers per node over its incident elements.
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

// FemLaplaceScalar_forward (deps/MFEM/FemLaplace1/FemLaplaceScalar.h:3-26):
// triplet t = 9*g + 3*p + q of Gauss point g = nq*e + k.  One thread per
// triplet so the 24-byte outputs are written coalesced.
__global__ void laplace_coo_kernel(int nelem, QuadRule rule,
                                   const int *__restrict__ conn,
                                   const int *__restrict__ conn_user,
                                   const double *__restrict__ x,
                                   const double *__restrict__ y,
                                   const double *__restrict__ kappa,
                                   long long *__restrict__ ind,
                                   double *__restrict__ vv) {
#pragma clang fp contract(off)
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)9 * rule.nq * nelem;
  if (t >= total) return;
  int64_t g = t / 9;
  int pq = (int)(t - 9 * g);
  int p = pq / 3, q = pq - 3 * p;
  int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  ElemCoords c = load_elem(conn, nelem, e, x, y);
  ElemGeom G = elem_geom(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
  double d = G.gx[p] * G.gx[q] + G.gy[p] * G.gy[q];
  double w = gauss_weight(rule.w[k], G.area);
  ind[2 * t] = conn_user[(size_t)p * nelem + e];
  ind[2 * t + 1] = conn_user[(size_t)q * nelem + e];
  vv[t] = d * kappa[g] * w;
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

// ComputeFemMassMatrix1_forward (ComputeFemMassMatrixMfem.h:4-29).
__global__ void mass_coo_kernel(int nelem, QuadRule rule,
                                const int *__restrict__ conn,
                                const int *__restrict__ conn_user,
                                const double *__restrict__ x,
                                const double *__restrict__ y,
                                const double *__restrict__ rho,
                                long long *__restrict__ ind,
                                double *__restrict__ vv) {
#pragma clang fp contract(off)
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)9 * rule.nq * nelem;
  if (t >= total) return;
  int64_t g = t / 9;
  int pq = (int)(t - 9 * g);
  int p = pq / 3, q = pq - 3 * p;
  int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  ElemCoords c = load_elem(conn, nelem, e, x, y);
  double A = heron_area(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
  double w = gauss_weight(rule.w[k], A);
  double hp = shape(p, rule.xi[k], rule.eta[k]), hq = shape(q, rule.xi[k], rule.eta[k]);
  ind[2 * t] = conn_user[(size_t)p * nelem + e];
  ind[2 * t + 1] = conn_user[(size_t)q * nelem + e];
  vv[t] = hp * hq * rho[g] * w;
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
__global__ void source_kernel(int n, QuadRule rule, int nelem,
                              const int *__restrict__ conn,
                              const double *__restrict__ x,
                              const double *__restrict__ y,
                              const int *__restrict__ n2e_ptr,
                              const int *__restrict__ n2e_elem,
                              const uint32_t *__restrict__ slots,
                              const int *__restrict__ iperm,
                              const double *__restrict__ f,
                              double *__restrict__ rhs_user, int accumulate) {
#pragma clang fp contract(off)
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int u = iperm[r];
  double acc = accumulate ? rhs_user[u] : 0.0;
  for (int t = n2e_ptr[r]; t < n2e_ptr[r + 1]; t++) {
    int e = n2e_elem[t];
    int p = slots[t] & 3u;
    ElemCoords c = load_elem(conn, nelem, e, x, y);
    double A = heron_area(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
    for (int k = 0; k < rule.nq; k++) {
      double w = gauss_weight(rule.w[k], A);
      acc += f[(size_t)rule.nq * e + k] * shape(p, rule.xi[k], rule.eta[k]) * w;
    }
  }
  rhs_user[u] = acc;
}

// FemSourceScalar_backward (FemSourceScalar.h:17-32).
__global__ void source_grad_kernel(int nelem, QuadRule rule,
                                   const int *__restrict__ conn,
                                   const int *__restrict__ conn_user,
                                   const double *__restrict__ x,
                                   const double *__restrict__ y,
                                   const double *__restrict__ grhs,
                                   double *__restrict__ gf) {
#pragma clang fp contract(off)
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  ElemCoords c = load_elem(conn, nelem, e, x, y);
  double A = heron_area(c.x[0], c.y[0], c.x[1], c.y[1], c.x[2], c.y[2]);
  double w = gauss_weight(rule.w[k], A);
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
  int64_t total = 9 * m.ngauss;
  laplace_coo_kernel<<<nblk(total, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_user.p, m.x.p, m.y.p, kappa, ind, vv);
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
  int64_t total = 9 * m.ngauss;
  mass_coo_kernel<<<nblk(total, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_user.p, m.x.p, m.y.p, rho, ind, vv);
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
  source_kernel<<<nblk(m.nnode, 256), 256, 0, st>>>(
      m.nnode, m.rule, m.nelem, m.conn.p, m.x.p, m.y.p, m.d_n2e_ptr.p,
      m.d_n2e_elem.p, m.d_slots_full.p, m.d_iperm.p, f, rhs_user,
      accumulate ? 1 : 0);
  HIPCHECK(hipGetLastError());
}

void launch_source_grad(const Mesh &m, const double *grhs_user, double *gf,
                        hipStream_t st) {
  if (m.degree == 2) return launch_p2_source_grad(m, grhs_user, gf, st);
  source_grad_kernel<<<nblk(m.ngauss, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_user.p, m.x.p, m.y.p, grhs_user, gf);
  HIPCHECK(hipGetLastError());
}

}  // namespace admfem
