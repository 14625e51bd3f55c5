// Fused, atomic-free CSR assembly of the Dirichlet-eliminated system
//   A = mass_coef * M(rho) + lap_coef * K(kappa),
//   rhs = src_coef * F(f) + M(rho) u_prev - A[:, bd] bdval  (interior rows),
//   A[b, :] = e_b, rhs[b] = bdval  (boundary rows),
// which is what compute_fem_{laplace,mass}_matrix1 + compute_fem_source_term1
// + impose_Dirichlet_boundary_conditions produce in the reference
// (FemLaplaceScalar.h:3-26, ComputeFemMassMatrixMfem.h:4-29,
// FemSourceScalar.h:4-15, ImposeDirichlet.h:27-60), summed into CSR.
//
// A block owns AS_ROWS consecutive rows and works in two passes over LDS:
//  1. one thread per (row, incident element) pair computes that element's
//     matrix row for the row node (stiffness and mass parts), its load-vector
//     term and its Dirichlet lifting -- independent work items, so the
//     element gathers of many pairs are in flight at once;
//  2. one thread per row adds its pairs, in ascending element order, into the
//     row's CSR slots and its right-hand side.
// The block's contiguous CSR range is then written out coalesced.  Every slot
// has one writer and a fixed summation order: no atomics, deterministic.
#include "geom.h"

namespace admfem {

namespace {

constexpr int AS_NT = 256;

__global__ void __launch_bounds__(AS_NT)
    assemble_kernel(int n, int nelem, QuadRule rule,
                    const int *__restrict__ conn, const double *__restrict__ x,
                    const double *__restrict__ y,
                    const int *__restrict__ n2e_ptr,
                    const int *__restrict__ n2e_elem,
                    const uint32_t *__restrict__ slots,
                    const int *__restrict__ rp, const int *__restrict__ row_bd,
                    admfem_problem P, double *__restrict__ val,
                    double *__restrict__ rhs, int max_inc, int max_nnz) {
  extern __shared__ double lds[];
  double *ca = lds;                        // [max_inc][3] matrix-row terms
  double *rc = ca + 3 * (size_t)max_inc;   // [max_inc] right-hand-side terms
  double *vl = rc + max_inc;               // [max_nnz] the block's CSR values
  uint32_t *ws = (uint32_t *)(vl + max_nnz);  // [max_inc] slot words
  int *ip = (int *)(ws + max_inc);         // [AS_ROWS + 1] incidence offsets
  const int r0 = blockIdx.x * AS_ROWS;
  const int rend = min(r0 + AS_ROWS, n);
  const int nr = rend - r0;
  const int i0 = n2e_ptr[r0];
  const int ninc = n2e_ptr[rend] - i0;
  const int base = rp[r0];
  const int bnnz = rp[rend] - base;
  for (int k = threadIdx.x; k <= nr; k += AS_NT) ip[k] = n2e_ptr[r0 + k] - i0;
  for (int k = threadIdx.x; k < bnnz; k += AS_NT) vl[k] = 0.0;
  __syncthreads();
  // pass 1: (row, element) pairs
  for (int t = threadIdx.x; t < ninc; t += AS_NT) {
    int lo = 0, hi = nr - 1;  // row of pair t: last k with ip[k] <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ip[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int r = r0 + lo;
    const uint32_t wd = slots[i0 + t];
    ws[t] = wd;
    if (row_bd[r] >= 0) continue;  // boundary row: only its unit diagonal
    const int e = n2e_elem[i0 + t];
    const int p = wd & 3u;
    int v[3];
    double ex[3], ey[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = conn[(size_t)k * nelem + e];
      ex[k] = x[v[k]];
      ey[k] = y[v[k]];
    }
    const ElemGeom G = elem_geom(ex[0], ey[0], ex[1], ey[1], ex[2], ey[2]);
    // K_e = (sum_k w_k kappa_k) grad phi_p . grad phi_q (P1 gradients are
    // constant, so the Gauss sum factors out)
    double ck = 0.0, src = 0.0;
    double mrow[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < rule.nq; k++) {
      const size_t g = (size_t)rule.nq * e + k;
      const double w = gauss_weight(rule.w[k], G.area);
      ck += w * (P.kappa ? P.kappa[g] : 1.0);
      const double hp = shape(p, rule.xi[k], rule.eta[k]);
      if (P.mass_coef != 0.0 || P.u_prev) {
        const double wr = w * (P.rho ? P.rho[g] : 1.0) * hp;
#pragma unroll
        for (int q = 0; q < 3; q++) mrow[q] += wr * shape(q, rule.xi[k], rule.eta[k]);
      }
      if (P.f) src += P.f[g] * hp * w;
    }
    double rt = P.src_coef * src;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      double a = P.lap_coef * ck * (G.gx[p] * G.gx[q] + G.gy[p] * G.gy[q]);
      if (P.mass_coef != 0.0) a += P.mass_coef * mrow[q];
      if (P.u_prev) rt += mrow[q] * P.u_prev[v[q]];
      if (slot_of(wd, q) == SLOT_LIFT && P.bdval) rt -= a * P.bdval[row_bd[v[q]]];
      ca[3 * t + q] = a;
    }
    rc[t] = rt;
  }
  __syncthreads();
  // pass 2: rows sum their pairs in ascending element order
  for (int k = threadIdx.x; k < nr; k += AS_NT) {
    const int r = r0 + k;
    const int off = rp[r] - base;
    const int bi = row_bd[r];
    if (bi >= 0) {
      vl[off] = 1.0;  // the appended (b,b,1.0) of ImposeDirichlet.h:46-50
      rhs[r] = P.bdval ? P.bdval[bi] : 0.0;
      continue;
    }
    double acc = 0.0;
    for (int t = ip[k]; t < ip[k + 1]; t++) {
      const uint32_t wd = ws[t];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const uint32_t sl = slot_of(wd, q);
        if (sl != SLOT_LIFT) vl[off + sl] += ca[3 * t + q];
      }
      acc += rc[t];
    }
    rhs[r] = acc;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < bnnz; k += AS_NT) val[base + k] = vl[k];
}

// CSR (post-BC) -> SELL-64 off-diagonals + diagonal, FP64 and FP32 copies.
// Every entry of the slice is written, the padding ones with 0, so the
// destination needs no clearing beforehand (serving reuses device memory).
__global__ void to_sell_kernel(int n, const int *__restrict__ rp,
                               const int *__restrict__ col,
                               const double *__restrict__ val,
                               const int *__restrict__ slice_ptr,
                               double *__restrict__ sv,
                               float *__restrict__ sv32,
                               double *__restrict__ diag,
                               double *__restrict__ dinv,
                               float *__restrict__ diag32,
                               float *__restrict__ dinv32, int npad) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npad) return;
  const int s0 = slice_ptr[r >> 6], width = (slice_ptr[(r >> 6) + 1] - s0) >> 6;
  const int off = s0 + (r & 63);
  int kk = 0;
  double d = 1.0;  // padding rows: identity
  if (r < n) {
    d = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; k++) {
      double v = val[k];
      if (col[k] == r) {
        d = v;
      } else {
        sv[off + 64 * kk] = v;
        sv32[off + 64 * kk] = (float)v;
        kk++;
      }
    }
  }
  for (; kk < width; kk++) {
    sv[off + 64 * kk] = 0.0;
    sv32[off + 64 * kk] = 0.0f;
  }
  diag[r] = d;
  dinv[r] = d != 0.0 ? 1.0 / d : 0.0;
  diag32[r] = (float)d;
  dinv32[r] = d != 0.0 ? (float)(1.0 / d) : 0.0f;
}

// dL/dkappa_{e,k} = -scale * w_k * grad(lam~) . grad(u) on element e, with
// lam~ = lam zeroed on boundary rows (the gradient ImposeDirichletGrad drops,
// ImposeDirichlet.h:63-93).  One thread per element, nq outputs each.
__global__ void kappa_grad_kernel(int nelem, QuadRule rule,
                                  const int *__restrict__ conn,
                                  const double *__restrict__ x,
                                  const double *__restrict__ y,
                                  const int *__restrict__ row_bd,
                                  const double *__restrict__ u,
                                  const double *__restrict__ lam, double scale,
                                  double *__restrict__ gk) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  int v[3];
  double ex[3], ey[3], uu[3], ll[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v[k] = conn[(size_t)k * nelem + e];
    ex[k] = x[v[k]];
    ey[k] = y[v[k]];
    uu[k] = u[v[k]];
    ll[k] = row_bd[v[k]] >= 0 ? 0.0 : lam[v[k]];
  }
  ElemGeom G = elem_geom(ex[0], ey[0], ex[1], ey[1], ex[2], ey[2]);
  double dux = 0, duy = 0, dlx = 0, dly = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    dux += uu[k] * G.gx[k];
    duy += uu[k] * G.gy[k];
    dlx += ll[k] * G.gx[k];
    dly += ll[k] * G.gy[k];
  }
  const double d = dlx * dux + dly * duy;
  for (int k = 0; k < rule.nq; k++)
    gk[(size_t)rule.nq * e + k] = -scale * gauss_weight(rule.w[k], G.area) * d;
}

__global__ void gather_kernel(const int *__restrict__ idx,
                              const double *__restrict__ src,
                              double *__restrict__ dst, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// Loss L = sum_k (u[perm[obs_idx[k]]] - obs_val[k])^2 and its gradient
// g = dL/du (internal order; g zeroed beforehand).  One block, fixed order.
__global__ void __launch_bounds__(1024)
    loss_kernel(int n, const int *__restrict__ perm, const int *__restrict__ idx,
                const double *__restrict__ obs, int nobs,
                const double *__restrict__ u, double *__restrict__ g,
                double *__restrict__ loss) {
  __shared__ double sh[16];
  double acc = 0.0, bad = 0.0;
  for (int k = threadIdx.x; k < nobs; k += blockDim.x) {
    int id = idx[k];
    if (id < 0 || id >= n) {  // caller data: never index with it unchecked
      bad = 1.0;
      continue;
    }
    int r = perm[id];
    double d = u[r] - obs[k];
    g[r] = 2.0 * d;
    acc += d * d;
  }
  double t = block_sum<1024>(acc, sh);
  double b = block_sum<1024>(bad, sh);
  if (threadIdx.x == 0) {
    loss[0] = t;
    loss[1] = b;
  }
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_assemble_to(const System &s, const double *x, const double *y,
                        const admfem_problem &p, double *val, double *rhs,
                        hipStream_t st) {
  const Mesh &m = *s.mesh;
  // the block with the most (row, element) pairs / CSR entries sizes the LDS
  const int mi = std::max(s.asm_max_inc, 1), mn = std::max(s.asm_max_nnz, 1);
  const size_t lds = (size_t)mi * (4 * sizeof(double) + sizeof(uint32_t)) +
                     (size_t)mn * sizeof(double) + (AS_ROWS + 1) * sizeof(int);
  REQUIRE(lds <= 160 * 1024, ADMFEM_ERR_UNSUPPORTED, "row block too dense for LDS");
  assemble_kernel<<<nblk(s.n, AS_ROWS), AS_NT, lds, st>>>(
      s.n, m.nelem, m.rule, m.conn.p, x, y, m.d_n2e_ptr.p, m.d_n2e_elem.p,
      s.d_slots.p, s.d_rp.p, s.d_row_bd.p, p, val, rhs, mi, mn);
  HIPCHECK(hipGetLastError());
}

void launch_assemble(const System &s, const admfem_problem &p, hipStream_t st) {
  launch_assemble_to(s, s.mesh->x.p, s.mesh->y.p, p, s.val.p, s.rhs.p, st);
}

void launch_to_sell_to(const System &s, const double *val, double *sv, float *sv32,
                       double *diag, double *dinv, float *diag32, float *dinv32,
                       hipStream_t st) {
  to_sell_kernel<<<nblk(s.npad, 256), 256, 0, st>>>(s.n, s.d_rp.p, s.d_col.p, val,
                                                    s.d_slice_ptr.p, sv, sv32, diag, dinv,
                                                    diag32, dinv32, s.npad);
  HIPCHECK(hipGetLastError());
}

void launch_to_sell(System &s, hipStream_t st) {
  launch_to_sell_to(s, s.val.p, s.d_sell_val.p, s.d_sell_val32.p, s.d_diag.p, s.d_dinv.p,
                    s.d_diag32.p, s.d_dinv32.p, st);
}

void launch_kappa_gradient_to(const System &s, const double *x, const double *y,
                              const double *u_int, const double *lam_int, double scale,
                              double *gk, hipStream_t st) {
  const Mesh &m = *s.mesh;
  kappa_grad_kernel<<<nblk(m.nelem, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, x, y, s.d_row_bd.p, u_int, lam_int, scale, gk);
  HIPCHECK(hipGetLastError());
}

void launch_kappa_gradient(const System &s, const double *u_int,
                           const double *lam_int, double scale, double *gk,
                           hipStream_t st) {
  launch_kappa_gradient_to(s, s.mesh->x.p, s.mesh->y.p, u_int, lam_int, scale, gk, st);
}

// Loss and its gradient without a host round trip: g (npad, internal order)
// is cleared with a kernel, loss2 receives {L, bad-index flag}.
namespace {
__global__ void fill_zero(double *a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}
}  // namespace

void launch_loss(const System &s, const double *u_int, const int *obs_idx,
                 const double *obs_val, int nobs, double *g_int, double *loss2,
                 hipStream_t st) {
  fill_zero<<<nblk(s.npad, 256), 256, 0, st>>>(g_int, s.npad);
  loss_kernel<<<1, 1024, 0, st>>>(s.n, s.mesh->d_perm.p, obs_idx, obs_val, nobs, u_int, g_int,
                                  loss2);
  HIPCHECK(hipGetLastError());
}

void launch_gather(const int *idx, const double *src, double *dst, int n,
                   hipStream_t st) {
  if (n <= 0) return;
  gather_kernel<<<nblk(n, 256), 256, 0, st>>>(idx, src, dst, n);
  HIPCHECK(hipGetLastError());
}

double loss_and_grad(const System &s, const double *u_int, const int *obs_idx,
                     const double *obs_val, int nobs, double *g_int,
                     hipStream_t st) {
  HIPCHECK(hipMemsetAsync(g_int, 0, sizeof(double) * s.npad, st));
  DevBuf<double> d_loss(2);
  loss_kernel<<<1, 1024, 0, st>>>(s.n, s.mesh->d_perm.p, obs_idx, obs_val, nobs,
                                  u_int, g_int, d_loss.p);
  HIPCHECK(hipGetLastError());
  double h[2] = {0.0, 0.0};
  HIPCHECK(hipMemcpyAsync(h, d_loss.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  REQUIRE(h[1] == 0.0, ADMFEM_ERR_ARG, "observation index out of range [0, n)");
  return h[0];
}

}  // namespace admfem
