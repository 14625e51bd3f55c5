// Nodal fields <-> Gauss points, and the matrix-free Laplace term, for P1
// (P2 meshes dispatch to p2.hip, except FemToGaussPoints, which maps vertex
// values with the linear functions for either degree):
//   FemToGaussPointsMfem / DofToGaussPointsMfem (deps/MFEM/FemToGaussPoints/
//     FemToGaussPointsMfem.h:6-33, DofToGaussPoints/DofToGaussPointsMfem.h:
//     6-32): out[g] = sum_r u[node_r] h_r(g)
//   FemGradMfem (deps/MFEM/FemGrad/FemGradMfem.h:8-41):
//     out[g] = (sum_r u[node_r] dh_r/dx, sum_r u[node_r] dh_r/dy)
//   ComputeLaplaceTermMfem (deps/MFEM/ComputeLaplaceTermMfem/
//     ComputeLaplaceTermMfem.h:4-39): out = K(nu) u without forming K.
// Fields are in the user's node order, Gauss arrays in the user's element
// order (g = nq e + k).  Forward maps to Gauss points run one thread per
// Gauss point; every map to the nodes (the backward ops and the Laplace
// term) is a per-node gather over the node's incident elements: one writer
// per output, no atomics.
#include "geom.h"

namespace admfem {

namespace {

struct Tri {
  double x[3], y[3];
};

__device__ __forceinline__ Tri load_tri(const int *__restrict__ conn, int nelem, int e,
                                        const double *__restrict__ x,
                                        const double *__restrict__ y) {
  Tri t;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int v = conn[(size_t)k * nelem + e];
    t.x[k] = x[v];
    t.y[k] = y[v];
  }
  return t;
}

__global__ void to_gauss_kernel(int nelem, QuadRule rule, const int *__restrict__ conn_user,
                                const double *__restrict__ u, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  const int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  const double xi = rule.xi[k], eta = rule.eta[k];
  out[g] = u[conn_user[e]] * shape(0, xi, eta) +
           u[conn_user[(size_t)nelem + e]] * shape(1, xi, eta) +
           u[conn_user[2 * (size_t)nelem + e]] * shape(2, xi, eta);
}

__global__ void fem_grad_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
                                const int *__restrict__ conn_user, const double *__restrict__ x,
                                const double *__restrict__ y, const double *__restrict__ u,
                                double *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  const int e = (int)(g / rule.nq);
  const Tri t = load_tri(conn, nelem, e, x, y);
  const ElemGeom G = elem_geom(t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2]);
  double gx = 0.0, gy = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) gx += u[conn_user[(size_t)r * nelem + e]] * G.gx[r];
#pragma unroll
  for (int r = 0; r < 3; r++) gy += u[conn_user[(size_t)r * nelem + e]] * G.gy[r];
  out[2 * g] = gx;
  out[2 * g + 1] = gy;
}

// Per node: the transposes of the two maps above, and K(nu) u.
// NODE_LAPLACE reads its field in the internal node order (permuted once by
// the launcher), so each incidence's three values are local gathers through
// the element's internal vertex ids instead of random ones in user order.
enum { NODE_TO_GAUSS_T = 0, NODE_GRAD_T = 1, NODE_LAPLACE = 2 };

template <int OP, int NQ>
__global__ void __launch_bounds__(256) node_gather_kernel(int n, int nelem, QuadRule rule, const int *__restrict__ conn,
                                   const int *__restrict__ conn_user,
                                   const double *__restrict__ x, const double *__restrict__ y,
                                   const int *__restrict__ n2e_ptr,
                                   const int *__restrict__ n2e_elem,
                                   const uint32_t *__restrict__ slots,
                                   const int *__restrict__ iperm,
                                   const double *__restrict__ gw,   // mesh Gauss weights
                                   const double *__restrict__ in,   // Gauss array or field
                                   const double *__restrict__ nu,   // NODE_LAPLACE
                                   double *__restrict__ out_user) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double acc = 0.0;
  const int t1 = n2e_ptr[r + 1];
  // unrolled over incidences so that their loads overlap
#pragma unroll 4
  for (int t = n2e_ptr[r]; t < t1; t++) {
    const int e = n2e_elem[t];
    const int p = (int)(slots[t] & 3u);
    if (OP == NODE_TO_GAUSS_T) {
#pragma unroll
      for (int k = 0; k < NQ; k++)
        acc += in[(size_t)NQ * e + k] * shape(p, rule.xi[k], rule.eta[k]);
    } else {
      const Tri tr = load_tri(conn, nelem, e, x, y);
      double gx[3], gy[3];
      elem_grads(tr.x[0], tr.y[0], tr.x[1], tr.y[1], tr.x[2], tr.y[2], gx, gy);
      if (OP == NODE_GRAD_T) {
#pragma unroll
        for (int k = 0; k < NQ; k++) {
          const size_t g = (size_t)NQ * e + k;
          acc += in[2 * g] * gx[p] + in[2 * g + 1] * gy[p];
        }
      } else {  // K(nu) u: sum_k nu_k w_k grad phi_p . grad u_e
        double ux = 0.0, uy = 0.0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const double uq = in[conn[(size_t)q * nelem + e]];
          ux += gx[q] * uq;
          uy += gy[q] * uq;
        }
        double cw = 0.0;
#pragma unroll
        for (int k = 0; k < NQ; k++) cw += nu[(size_t)NQ * e + k] * gw[(size_t)NQ * e + k];
        acc += cw * (gx[p] * ux + gy[p] * uy);
      }
    }
  }
  out_user[iperm[r]] = acc;
}

// ComputeLaplaceTermMfemGrad, nu part: w_g grad(grad_out) . grad(u) per Gauss
// point (the term is linear in nu).
__global__ void laplace_term_nu_grad_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
                                            const int *__restrict__ conn_user,
                                            const double *__restrict__ x,
                                            const double *__restrict__ y,
                                            const double *__restrict__ go,
                                            const double *__restrict__ u,
                                            double *__restrict__ gnu) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rule.nq * nelem) return;
  const int e = (int)(g / rule.nq), k = (int)(g - (int64_t)rule.nq * e);
  const Tri t = load_tri(conn, nelem, e, x, y);
  const ElemGeom G = elem_geom(t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2]);
  double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int v = conn_user[(size_t)q * nelem + e];
    ax += G.gx[q] * go[v];
    ay += G.gy[q] * go[v];
    bx += G.gx[q] * u[v];
    by += G.gy[q] * u[v];
  }
  gnu[g] = gauss_weight(rule.w[k], G.area) * (ax * bx + ay * by);
}

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// dst[r] = src[iperm[r]]: a user-order field in the internal order
__global__ void permute_kernel(int n, const int *__restrict__ iperm,
                               const double *__restrict__ src, double *__restrict__ dst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[iperm[r]];
}

template <int OP>
void node_gather(const Mesh &m, const double *in, const double *nu, double *out_user,
                 hipStream_t st) {
  auto go = [&](auto kern) {
    kern<<<nb(m.nnode, 256), 256, 0, st>>>(m.nnode, m.nelem, m.rule, m.conn.p, m.conn_user.p,
                                           m.x.p, m.y.p, m.d_n2e_ptr.p, m.d_n2e_elem.p,
                                           m.d_slots_full.p, m.d_iperm.p, m.gw.p, in, nu,
                                           out_user);
  };
  if (m.rule.nq == 3) go(node_gather_kernel<OP, 3>);
  else if (m.rule.nq == 6) go(node_gather_kernel<OP, 6>);
  else go(node_gather_kernel<OP, 1>);
  HIPCHECK(hipGetLastError());
}

}  // namespace

void launch_to_gauss(const Mesh &m, const double *u, double *out, hipStream_t st) {
  to_gauss_kernel<<<nb(m.ngauss, 256), 256, 0, st>>>(m.nelem, m.rule, m.conn_user.p, u, out);
  HIPCHECK(hipGetLastError());
}

void launch_to_gauss_grad(const Mesh &m, const double *grad_out, double *grad_u,
                          hipStream_t st) {
  node_gather<NODE_TO_GAUSS_T>(m, grad_out, nullptr, grad_u, st);
}

void launch_dof_to_gauss(const Mesh &m, const double *u, double *out, hipStream_t st) {
  if (m.degree == 2) return launch_p2_dof_to_gauss(m, u, out, st);
  launch_to_gauss(m, u, out, st);
}

void launch_dof_to_gauss_grad(const Mesh &m, const double *grad_out, double *grad_u,
                              hipStream_t st) {
  if (m.degree == 2) return launch_p2_dof_to_gauss_grad(m, grad_out, grad_u, st);
  launch_to_gauss_grad(m, grad_out, grad_u, st);
}

void launch_fem_grad(const Mesh &m, const double *u, double *out, hipStream_t st) {
  if (m.degree == 2) return launch_p2_fem_grad(m, u, out, st);
  fem_grad_kernel<<<nb(m.ngauss, 256), 256, 0, st>>>(m.nelem, m.rule, m.conn.p, m.conn_user.p,
                                                     m.x.p, m.y.p, u, out);
  HIPCHECK(hipGetLastError());
}

void launch_fem_grad_grad(const Mesh &m, const double *grad_out, double *grad_u,
                          hipStream_t st) {
  if (m.degree == 2) return launch_p2_fem_grad_grad(m, grad_out, grad_u, st);
  node_gather<NODE_GRAD_T>(m, grad_out, nullptr, grad_u, st);
}

// u (user order) -> internal order in a stream-ordered temporary, then the
// per-node gather
void launch_laplace_term(const Mesh &m, const double *u, const double *nu, double *out,
                         hipStream_t st) {
  if (m.degree == 2) return launch_p2_laplace_term(m, u, nu, out, st);
  double *u_int = nullptr;
  HIPCHECK(hipMallocAsync((void **)&u_int, sizeof(double) * std::max(m.nnode, 1), st));
  permute_kernel<<<nb(m.nnode, 256), 256, 0, st>>>(m.nnode, m.d_iperm.p, u, u_int);
  node_gather<NODE_LAPLACE>(m, u_int, nu, out, st);
  HIPCHECK(hipFreeAsync(u_int, st));
}

void launch_laplace_term_nu_grad(const Mesh &m, const double *grad_out, const double *u,
                                 double *grad_nu, hipStream_t st) {
  if (m.degree == 2) return launch_p2_laplace_term_nu_grad(m, grad_out, u, grad_nu, st);
  laplace_term_nu_grad_kernel<<<nb(m.ngauss, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.conn_user.p, m.x.p, m.y.p, grad_out, u, grad_nu);
  HIPCHECK(hipGetLastError());
}

}  // namespace admfem
