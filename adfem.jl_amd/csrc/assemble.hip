// Fused, atomic-free CSR assembly of the Dirichlet-eliminated system
//   A = mass_coef * M(rho) + lap_coef * K(kappa),
//   rhs = src_coef * F(f) + M(rho) u_prev - A[:, bd] bdval  (interior rows),
//   A[b, :] = e_b, rhs[b] = bdval  (boundary rows),
// which is what compute_fem_{laplace,mass}_matrix1 + compute_fem_source_term1
// + impose_Dirichlet_boundary_conditions produce in the reference
// (FemLaplaceScalar.h:3-26, ComputeFemMassMatrixMfem.h:4-29,
// FemSourceScalar.h:4-15, ImposeDirichlet.h:27-60), summed into CSR.
//
// One thread per matrix row, 256 rows per block.  Each row walks its incident
// elements (precomputed node->element list, ascending element id) and adds
// the element-matrix row into the row's slots, which live in LDS for the
// whole block; the block then writes its contiguous CSR range coalesced.
// Every slot has exactly one writer, so there are no atomics and the result
// is deterministic.
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
                    double *__restrict__ rhs) {
  extern __shared__ double lds[];
  const int r0 = blockIdx.x * AS_NT;
  const int r = r0 + threadIdx.x;
  const int rend = min(r0 + AS_NT, n);
  const int base = rp[r0];
  const int bnnz = rp[rend] - base;
  for (int k = threadIdx.x; k < bnnz; k += AS_NT) lds[k] = 0.0;
  __syncthreads();
  if (r < n) {
    const int bi = row_bd[r];
    const int off = rp[r] - base;
    if (bi >= 0) {
      lds[off] = 1.0;  // the appended (b,b,1.0) of ImposeDirichlet.h:46-50
      rhs[r] = P.bdval ? P.bdval[bi] : 0.0;
    } else {
      double src = 0.0, mu = 0.0, lift = 0.0;
      for (int t = n2e_ptr[r]; t < n2e_ptr[r + 1]; t++) {
        const int e = n2e_elem[t];
        const uint32_t wd = slots[t];
        const int p = wd & 3u;
        int v[3];
        double ex[3], ey[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          v[k] = conn[(size_t)k * nelem + e];
          ex[k] = x[v[k]];
          ey[k] = y[v[k]];
        }
        ElemGeom G = elem_geom(ex[0], ey[0], ex[1], ey[1], ex[2], ey[2]);
        // K_e = (sum_k w_k kappa_k) * grad phi_p . grad phi_q  (P1: constant
        // gradients, so the Gauss sum factors out)
        double ck = 0.0;
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
#pragma unroll
        for (int q = 0; q < 3; q++) {
          double a = P.lap_coef * ck * (G.gx[p] * G.gx[q] + G.gy[p] * G.gy[q]);
          if (P.mass_coef != 0.0) a += P.mass_coef * mrow[q];
          if (P.u_prev) mu += mrow[q] * P.u_prev[v[q]];
          const uint32_t s = slot_of(wd, q);
          if (s == SLOT_LIFT) {
            if (P.bdval) lift += a * P.bdval[row_bd[v[q]]];
          } else {
            lds[off + s] += a;
          }
        }
      }
      rhs[r] = P.src_coef * src + mu - lift;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < bnnz; k += AS_NT) val[base + k] = lds[k];
}

// CSR (post-BC) -> SELL-64 off-diagonals + diagonal, FP64 and FP32 copies.
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
  if (r >= n) {  // padding rows: identity
    diag[r] = 1.0;
    dinv[r] = 1.0;
    diag32[r] = 1.0f;
    dinv32[r] = 1.0f;
    return;
  }
  const int off = slice_ptr[r >> 6] + (r & 63);
  int kk = 0;
  double d = 0.0;
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
  // the largest per-block CSR range decides the dynamic LDS size
  size_t lds = (size_t)std::max(s.max_block_nnz, 1) * sizeof(double);
  REQUIRE(lds <= 160 * 1024, ADMFEM_ERR_UNSUPPORTED, "row block too dense for LDS");
  assemble_kernel<<<nblk(s.n, AS_NT), AS_NT, lds, st>>>(
      s.n, m.nelem, m.rule, m.conn.p, x, y, m.d_n2e_ptr.p, m.d_n2e_elem.p,
      s.d_slots.p, s.d_rp.p, s.d_row_bd.p, p, val, rhs);
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
