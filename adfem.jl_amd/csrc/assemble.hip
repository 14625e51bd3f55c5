// Fused, atomic-free assembly of the Dirichlet-eliminated system
//   A = mass_coef * M(rho) + lap_coef * K(kappa),
//   rhs = src_coef * F(f) + M(rho) u_prev - A[:, bd] bdval  (interior rows),
//   A[b, :] = e_b, rhs[b] = bdval  (boundary rows),
// which is what compute_fem_{laplace,mass}_matrix1 + compute_fem_source_term1
// + impose_Dirichlet_boundary_conditions produce in the reference
// (FemLaplaceScalar.h:3-26, ComputeFemMassMatrixMfem.h:4-29,
// FemSourceScalar.h:4-15, ImposeDirichlet.h:27-60), written straight into
// the solver's format: SELL-64 off-diagonal values (FP64), the diagonal and
// its inverse, and the right-hand side.
//
// A block owns one 64-row SELL slice (blocks mapped XCD-contiguously) and
// works in two passes over LDS:
//  1. one thread per element incident to the block's interior rows computes
//     the element's symmetric 3x3 matrix and its 3 lifted load terms -- each
//     element once per block, its gathers independent of each other;
//  2. one thread per row adds, in ascending element order, its pairs' matrix
//     row into its SELL slots (LDS) and diagonal, and their load terms.
// Each row's thread then writes its slice lanes: consecutive rows are
// consecutive addresses.  Every value has one writer and a fixed summation
// order: no atomics, deterministic.
#include "geom.h"

namespace admfem {

namespace {

constexpr int AS_NT = 128;

// index of (p, q) in the packed upper triangle [00 01 02 11 12 22]
__device__ __forceinline__ int sym3(int p, int q) {
  const int lo = min(p, q), hi = max(p, q);
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

constexpr int PF = 8;  // pair words of a row fetched before the element pass

// One element's gathered inputs: internal vertex ids, vertex coordinates,
// and its Gauss-point coefficients.
template <int NQ>
struct ElemIn {
  int v[3];
  double ex[3], ey[3], kq[NQ], fq[NQ], rq[NQ];
  template <bool MASS>
  __device__ __forceinline__ void load(const int4 rec, const double *__restrict__ x,
                                       const double *__restrict__ y, const admfem_problem &P) {
    v[0] = rec.y;
    v[1] = rec.z;
    v[2] = rec.w;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const size_t g = (size_t)NQ * rec.x + k;
      kq[k] = P.kappa ? P.kappa[g] : 1.0;
      fq[k] = P.f ? P.f[g] : 0.0;
      rq[k] = MASS && P.rho ? P.rho[g] : 1.0;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
      ex[k] = x[v[k]];
      ey[k] = y[v[k]];
    }
  }
};

// The element's symmetric matrix (packed upper triangle, 6 values) and its
// three lifted load terms, in closed form for the affine P1 map.  With
// d_p = (x_{p+2} - x_{p+1}, y_{p+1} - y_{p+2}) and det = 2 * signed area,
// grad phi_p = (d_p.y, d_p.x) / det and w_k = ipw_k * area / 0.5 =
// ipw_k * |det| (Common.cpp:116; |det| / 2 is Heron's area up to rounding),
// so the per-Gauss-point sums of FemLaplaceScalar.h:7-25,
// ComputeFemMassMatrixMfem.h:4-29 and FemSourceScalar.h:4-15 become
//   K_pq = lap * (sum_k ipw_k kappa_k) * (d_p . d_q) / |det|,
//   M_pq = |det| * sum_k ipw_k rho_k h_pk h_qk,
//   F_p  = |det| * sum_k ipw_k f_k h_pk:
// no square roots and one division per element.  The values agree with the
// per-Gauss-point form to rounding (a few ulps); the COO ops keep that form.
template <int NQ, bool MASS>
__device__ __forceinline__ void element(const ElemIn<NQ> &in, const QuadRule &rule,
                                        const admfem_problem &P, const int *__restrict__ row_bd,
                                        double *ae, double *rl) {
  const double dx[3] = {in.ex[2] - in.ex[1], in.ex[0] - in.ex[2], in.ex[1] - in.ex[0]};
  const double dy[3] = {in.ey[1] - in.ey[2], in.ey[2] - in.ey[0], in.ey[0] - in.ey[1]};
  const double ad = fabs(dx[2] * dy[1] - dx[1] * dy[2]);
  double sk = 0.0, mm[6] = {0, 0, 0, 0, 0, 0}, src[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < NQ; k++) {
    sk += rule.w[k] * in.kq[k];
    const double h[3] = {shape(0, rule.xi[k], rule.eta[k]), shape(1, rule.xi[k], rule.eta[k]),
                         shape(2, rule.xi[k], rule.eta[k])};
    if (MASS) {
#pragma unroll
      for (int p = 0; p < 3; p++) {
        const double wr = rule.w[k] * in.rq[k] * h[p];
#pragma unroll
        for (int q = p; q < 3; q++) mm[sym3(p, q)] += wr * h[q];
      }
    }
    if (P.f) {
      const double wf = rule.w[k] * in.fq[k];
#pragma unroll
      for (int p = 0; p < 3; p++) src[p] += wf * h[p];
    }
  }
  if (MASS) {
#pragma unroll
    for (int i = 0; i < 6; i++) mm[i] *= ad;
  }
  const double c = P.lap_coef * sk / ad;
  double a[6];
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = p; q < 3; q++) {
      double val = c * (dx[p] * dx[q] + dy[p] * dy[q]);
      if (MASS && P.mass_coef != 0.0) val += P.mass_coef * mm[sym3(p, q)];
      a[sym3(p, q)] = val;
    }
  double r[3];
  const double sc = P.src_coef * ad;
#pragma unroll
  for (int p = 0; p < 3; p++) r[p] = sc * src[p];
  if (MASS && P.u_prev) {
    double up[3];
#pragma unroll
    for (int q = 0; q < 3; q++) up[q] = P.u_prev[in.v[q]];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) r[p] += mm[sym3(p, q)] * up[q];
  }
  if (P.bdval) {  // Dirichlet lifting of the boundary columns
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int bq = row_bd[in.v[q]];
      if (bq >= 0) {
        const double bv = P.bdval[bq];
#pragma unroll
        for (int p = 0; p < 3; p++) r[p] -= a[sym3(p, q)] * bv;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) ae[k] = a[k];
#pragma unroll
  for (int p = 0; p < 3; p++) rl[p] = r[p];
}

template <int NQ, bool MASS>
__global__ void __launch_bounds__(AS_NT)
    assemble_kernel(int n, int npad, QuadRule rule, const double *__restrict__ x,
                    const double *__restrict__ y, const int *__restrict__ blk_eptr,
                    const int4 *__restrict__ blk_rec, const int *__restrict__ pair_ptr,
                    const uint32_t *__restrict__ pairs, const int *__restrict__ row_bd,
                    const int *__restrict__ slice_ptr, admfem_problem P, SellOut o, int max_e,
                    int wmax) {
  extern __shared__ double lds[];
  double *ae = lds;                     // [max_e][6] element matrices
  double *rl = ae + 6 * (size_t)max_e;  // [max_e][3] element load terms
  // [wmax + 2][AS_ROWS] row accumulators, slot-major: slots [0, wmax) are
  // the row's SELL positions, slot wmax its diagonal, wmax + 1 a sink for
  // lifted (boundary) columns, so pass 2 adds without branching
  double *vl = rl + 3 * (size_t)max_e;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int row = blk * AS_ROWS + (int)threadIdx.x;  // phase 2: threads < AS_ROWS
  const int e0 = blk_eptr[blk];
  const int ne = blk_eptr[blk + 1] - e0;
  // the row's pair words, fetched now so that their latency overlaps pass 1
  int bi = -1, t0 = 0, t1 = 0;
  uint32_t wb[PF];
  if (threadIdx.x < AS_ROWS && row < n) {
    bi = row_bd[row];
    if (bi < 0) {
      t0 = pair_ptr[row];
      t1 = pair_ptr[row + 1];
    }
#pragma unroll
    for (int k = 0; k < PF; k++) wb[k] = t0 + k < t1 ? pairs[t0 + k] : 0u;
  }
  for (int k = threadIdx.x; k < AS_ROWS * (wmax + 2); k += AS_NT) vl[k] = 0.0;
  // pass 1: elements (MASS: a mass term or u_prev is present), up to three
  // per thread and step with every load issued before any computes (a
  // 64-row block of a triangulation touches ~260 elements: one step)
  for (int t = threadIdx.x; t < ne; t += 3 * AS_NT) {
    ElemIn<NQ> in[3];
    const bool b1 = t + AS_NT < ne, b2 = t + 2 * AS_NT < ne;
    in[0].template load<MASS>(blk_rec[e0 + t], x, y, P);
    if (b1) in[1].template load<MASS>(blk_rec[e0 + t + AS_NT], x, y, P);
    if (b2) in[2].template load<MASS>(blk_rec[e0 + t + 2 * AS_NT], x, y, P);
    element<NQ, MASS>(in[0], rule, P, row_bd, ae + 6 * t, rl + 3 * t);
    if (b1) element<NQ, MASS>(in[1], rule, P, row_bd, ae + 6 * (t + AS_NT), rl + 3 * (t + AS_NT));
    if (b2)
      element<NQ, MASS>(in[2], rule, P, row_bd, ae + 6 * (t + 2 * AS_NT),
                        rl + 3 * (t + 2 * AS_NT));
  }
  __syncthreads();
  // pass 2: rows
  if (threadIdx.x >= AS_ROWS || row >= npad) return;
  double *my = vl + threadIdx.x;  // slot k at my[AS_ROWS * k]: no bank conflicts
  double d = 1.0, acc = 0.0;  // padding rows: identity, rhs 0
  if (row < n) {
    if (bi >= 0) {
      acc = P.bdval ? P.bdval[bi] : 0.0;  // (b,b,1.0) of ImposeDirichlet.h:46-50
    } else {
      auto add = [&](uint32_t w) {
        const int p = (int)(w & 3u), le = (int)(w >> 20);
        const double *a = ae + 6 * le;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const uint32_t code = (w >> (2 + SLOT_BITS * q)) & 63u;
          const int sl = code < PAIR_DIAG ? (int)code : wmax + (int)(code - PAIR_DIAG);
          my[AS_ROWS * sl] += a[sym3(p, q)];
        }
        acc += rl[3 * le + p];
      };
#pragma unroll
      for (int k = 0; k < PF; k++)
        if (t0 + k < t1) add(wb[k]);
      for (int t = t0 + PF; t < t1; t++) add(pairs[t]);
      d = my[AS_ROWS * wmax];
    }
  }
  const int s0 = slice_ptr[blk], width = (slice_ptr[blk + 1] - s0) >> 6;  // block = slice
  const int off = s0 + (row & 63);
  for (int k = 0; k < width; k++) o.sv[off + 64 * k] = my[AS_ROWS * k];
  o.diag[row] = d;
  if (o.dinv) o.dinv[row] = d != 0.0 ? 1.0 / d : 0.0;
  o.rhs[row] = acc;
}

// CSR (post-BC) -> SELL-64 off-diagonals + diagonal and its inverse (FP64).
// Every entry of the slice is written, the padding ones with 0, so the
// destination needs no clearing beforehand (serving reuses device memory).
__global__ void to_sell_kernel(int n, const int *__restrict__ rp,
                               const int *__restrict__ col,
                               const double *__restrict__ val,
                               const int *__restrict__ slice_ptr,
                               double *__restrict__ sv,
                               double *__restrict__ diag,
                               double *__restrict__ dinv, int npad) {
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
        kk++;
      }
    }
  }
  for (; kk < width; kk++) sv[off + 64 * kk] = 0.0;
  diag[r] = d;
  dinv[r] = d != 0.0 ? 1.0 / d : 0.0;
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
                        const admfem_problem &p, const SellOut &o, hipStream_t st) {
  const Mesh &m = *s.mesh;
  const int me = std::max(s.asm_max_e, 1), w = std::max(s.sell_wmax, 1);
  const size_t lds = (size_t)me * 9 * sizeof(double) + (size_t)AS_ROWS * (w + 2) * sizeof(double);
  REQUIRE(lds <= 160 * 1024, ADMFEM_ERR_UNSUPPORTED, "assembly block too dense for LDS");
  const unsigned grid = nblk(s.npad, AS_ROWS);
  static_assert(AS_ROWS == 64, "an assembly block is one SELL slice");
  const bool mass = p.mass_coef != 0.0 || p.u_prev;
  auto go = [&](auto kern) {
    kern<<<grid, AS_NT, lds, st>>>(s.n, s.npad, m.rule, x, y, s.d_blk_eptr.p, s.d_blk_rec.p,
                                   m.d_n2e_ptr.p, s.d_pairs.p, s.d_row_bd.p, s.d_slice_ptr.p, p,
                                   o, me, w);
  };
  if (m.rule.nq == 3) {
    if (mass) go(assemble_kernel<3, true>);
    else go(assemble_kernel<3, false>);
  } else if (m.rule.nq == 6) {
    if (mass) go(assemble_kernel<6, true>);
    else go(assemble_kernel<6, false>);
  } else {
    REQUIRE(m.rule.nq == 1, ADMFEM_ERR_UNSUPPORTED, "quadrature rule not built");
    if (mass) go(assemble_kernel<1, true>);
    else go(assemble_kernel<1, false>);
  }
  HIPCHECK(hipGetLastError());
}

void launch_assemble(System &s, const admfem_problem &p, hipStream_t st) {
  SellOut o{s.d_sell_val.p, s.d_diag.p, s.d_dinv.p, s.rhs.p};
  launch_assemble_to(s, s.mesh->x.p, s.mesh->y.p, p, o, st);
}

void launch_to_sell(System &s, hipStream_t st) {
  to_sell_kernel<<<nblk(s.npad, 256), 256, 0, st>>>(s.n, s.d_rp.p, s.d_col.p, s.val.p,
                                                    s.d_slice_ptr.p, s.d_sell_val.p,
                                                    s.d_diag.p, s.d_dinv.p, s.npad);
  HIPCHECK(hipGetLastError());
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
