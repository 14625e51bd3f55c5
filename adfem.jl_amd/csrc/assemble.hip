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
//  0. the block's distinct vertices (a few runs of consecutive rows) are
//     staged in LDS -- coordinates, and u_prev / Dirichlet values when used
//     -- while the elements' coefficients are being fetched;
//  1. one thread per element incident to the block's interior rows computes
//     the element's symmetric 3x3 matrix (or, with unit density, only its
//     Laplace factor) and its 3 lifted load terms -- each element once per
//     block, its only gathers its Gauss-point coefficients;
//  2. one thread per row adds, in ascending element order, its pairs' matrix
//     row into its SELL slots (LDS) and diagonal, and their load terms.
// Each row's thread then writes its slice lanes: consecutive rows are
// consecutive addresses.  Every value has one writer and a fixed summation
// order: no atomics, deterministic.
#include "geom.h"

namespace admfem {

namespace {

// 256 threads, one element each per pass-1 round: 78 VGPRs, 6 waves per
// SIMD (128 threads with three elements each in flight: 4 waves per SIMD);
// measured the same within 2% at 2048^2 (317.7 vs 322 us)
constexpr int AS_NT = 256;
constexpr int AS_EPT = 1;  // elements per thread in flight in pass 1

// index of (p, q) in the packed upper triangle [00 01 02 11 12 22]
__device__ __forceinline__ int sym3(int p, int q) {
  const int lo = min(p, q), hi = max(p, q);
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

constexpr int PF = 8;  // pair words of a row fetched before the element pass

// One element's inputs: its record {element id, 3 local vertex indices} and
// its Gauss-point coefficients (the only per-element gathers: the vertex
// data come from the block's staged vertices in LDS).
template <int NQ>
struct ElemIn {
  int loc;
  double kq[NQ], fq[NQ], rq[NQ];
  template <bool MASS>
  __device__ __forceinline__ void load(const int2 rec, const admfem_problem &P) {
    loc = rec.y;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const size_t g = (size_t)NQ * rec.x + k;
      kq[k] = P.kappa ? P.kappa[g] : 1.0;
      fq[k] = P.f ? P.f[g] : 0.0;
      rq[k] = MASS && P.rho ? P.rho[g] : 1.0;
    }
  }
  __device__ __forceinline__ int lv(int q) const {
    return (loc >> (BLK_VERT_BITS * q)) & (BLK_MAX_VERTS - 1);
  }
};

// The block's vertex data in LDS: coordinates, and (when present) u_prev and
// the Dirichlet value of boundary vertices (0 at interior vertices).
struct VertLds {
  double2 *xy;
  double *up, *bv;
};

// Edge vectors of the element's vertices: d_p = (x_{p+2} - x_{p+1},
// y_{p+1} - y_{p+2}), and |det| = 2 * area.
struct ElemGeo {
  double dx[3], dy[3], ad;
};
__device__ __forceinline__ ElemGeo elem_geo(const double (&ex)[3], const double (&ey)[3]) {
  ElemGeo g;
  g.dx[0] = ex[2] - ex[1];
  g.dx[1] = ex[0] - ex[2];
  g.dx[2] = ex[1] - ex[0];
  g.dy[0] = ey[1] - ey[2];
  g.dy[1] = ey[2] - ey[0];
  g.dy[2] = ey[0] - ey[1];
  g.ad = fabs(g.dx[2] * g.dy[1] - g.dx[1] * g.dy[2]);
  return g;
}

// The element's symmetric matrix (packed upper triangle, 6 values) and its
// three lifted load terms, in closed form for the affine P1 map.  With
// d_p as above and det = 2 * signed area, grad phi_p = (d_p.y, d_p.x) / det
// and w_k = ipw_k * area / 0.5 = ipw_k * |det| (Common.cpp:116; |det| / 2 is
// Heron's area up to rounding), so the per-Gauss-point sums of
// FemLaplaceScalar.h:7-25, ComputeFemMassMatrixMfem.h:4-29 and
// FemSourceScalar.h:4-15 become
//   K_pq = lap * (sum_k ipw_k kappa_k) * (d_p . d_q) / |det| = c (d_p . d_q),
//   M_pq = |det| * sum_k ipw_k rho_k h_pk h_qk,
//   F_p  = |det| * sum_k ipw_k f_k h_pk:
// no square roots and one division per element.  The values agree with the
// per-Gauss-point form to rounding (a few ulps); the COO ops keep that form.
// Outputs: a (the matrix), c (the Laplace factor), r (the load terms).
template <int NQ, bool MASS>
__device__ __forceinline__ void element(const ElemIn<NQ> &in, const VertLds &V,
                                        const QuadRule &rule, const admfem_problem &P,
                                        double (&a)[6], double &c, double (&r)[3]) {
  double ex[3], ey[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const double2 xy = V.xy[in.lv(q)];
    ex[q] = xy.x;
    ey[q] = xy.y;
  }
  const ElemGeo G = elem_geo(ex, ey);
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
    for (int i = 0; i < 6; i++) mm[i] *= G.ad;
  }
  c = P.lap_coef * sk / G.ad;
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = p; q < 3; q++) {
      double val = c * (G.dx[p] * G.dx[q] + G.dy[p] * G.dy[q]);
      if (MASS && P.mass_coef != 0.0) val += P.mass_coef * mm[sym3(p, q)];
      a[sym3(p, q)] = val;
    }
  const double sc = P.src_coef * G.ad;
#pragma unroll
  for (int p = 0; p < 3; p++) r[p] = sc * src[p];
  if (MASS && P.u_prev) {
    double up[3];
#pragma unroll
    for (int q = 0; q < 3; q++) up[q] = V.up[in.lv(q)];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) r[p] += mm[sym3(p, q)] * up[q];
  }
  if (P.bdval) {  // Dirichlet lifting of the boundary columns (value 0: nothing to lift)
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double bv = V.bv[in.lv(q)];
      if (bv != 0.0) {
#pragma unroll
        for (int p = 0; p < 3; p++) r[p] -= a[sym3(p, q)] * bv;
      }
    }
  }
}

__device__ __forceinline__ double pick3(const double (&v)[3], int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// Per-block LDS, in doubles.  Two layouts:
//  * full (a general mass density rho): [max_e][6] element matrices,
//    [max_e][3] load terms;
//  * compact (rho = 1 or no mass term: the matrix row is rebuilt in pass 2
//    from the staged coordinates, halving the per-element footprint so more
//    blocks fit a CU): [max_e][3] load terms, [max_e] Laplace factors c,
//    [max_v] staged coordinates, [max_e] packed local vertex indices (int);
// then one region used twice: first the staged vertex values read only by
// pass 1 (full: coordinates, u_prev, boundary values; compact: u_prev,
// boundary values), then, once every element is done, the row accumulators
// [wmax + 2][AS_ROWS] (slot-major: slots [0, wmax) are the row's SELL
// positions, slot wmax its diagonal, wmax + 1 a sink for lifted boundary
// columns, so pass 2 adds without branching).  max_e is even, so every
// double2 array is 16-byte aligned.
template <int NQ, bool MASS, bool COMPACT>
__global__ void __launch_bounds__(AS_NT)
    assemble_kernel(int n, int npad, QuadRule rule, const double2 *__restrict__ xy,
                    const int *__restrict__ blk_eptr, const int2 *__restrict__ blk_rec,
                    const int *__restrict__ blk_vptr, const int *__restrict__ blk_vert,
                    const int *__restrict__ pair_ptr, const uint32_t *__restrict__ pairs,
                    const int *__restrict__ row_bd, const int *__restrict__ slice_ptr,
                    admfem_problem P, SellOut o, int max_e, int max_v, int wmax) {
  extern __shared__ double lds[];
  double *ae = lds;                                      // full only
  double *rl = COMPACT ? lds : ae + 6 * (size_t)max_e;
  double *ec = rl + 3 * (size_t)max_e;                   // compact only
  VertLds V;
  double *region;
  int *eloc = nullptr;
  if (COMPACT) {
    V.xy = reinterpret_cast<double2 *>(ec + max_e);
    eloc = reinterpret_cast<int *>(V.xy + max_v);
    region = reinterpret_cast<double *>(eloc + max_e);
    V.up = region;
  } else {
    region = rl + 3 * (size_t)max_e;
    V.xy = reinterpret_cast<double2 *>(region);
    V.up = region + 2 * (size_t)max_v;
  }
  V.bv = V.up + (MASS ? (size_t)max_v : 0);
  double *vl = region;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int row = blk * AS_ROWS + (int)threadIdx.x;  // pass 2: threads < AS_ROWS
  const int e0 = blk_eptr[blk];
  const int ne = blk_eptr[blk + 1] - e0;
  const int v0 = blk_vptr[blk];
  const int nv = blk_vptr[blk + 1] - v0;
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
  // pass 1 (MASS: a mass term or u_prev is present), in chunks of AS_EPT
  // elements per thread (a 64-row block of a triangulation touches ~260
  // elements: one or two chunks).  The first chunk's records and coefficients
  // are loaded before the vertices are staged, so both latencies overlap;
  // the vertices are consecutive rows in runs, so their loads coalesce.
  auto store = [&](const ElemIn<NQ> &e, int t) {
    double a[6], c, r[3];
    element<NQ, MASS>(e, V, rule, P, a, c, r);
    if (COMPACT) {
      ec[t] = c;
      eloc[t] = e.loc;
    } else {
#pragma unroll
      for (int k = 0; k < 6; k++) ae[6 * t + k] = a[k];
    }
#pragma unroll
    for (int p = 0; p < 3; p++) rl[3 * t + p] = r[p];
  };
  for (int c0 = 0; c0 < max(ne, 1); c0 += AS_EPT * AS_NT) {  // uniform over the block
    ElemIn<NQ> in[AS_EPT];
    const int t = c0 + (int)threadIdx.x;
#pragma unroll
    for (int j = 0; j < AS_EPT; j++)
      if (t + j * AS_NT < ne) in[j].template load<MASS>(blk_rec[e0 + t + j * AS_NT], P);
    if (c0 == 0) {
      for (int k = threadIdx.x; k < nv; k += AS_NT) {
        const int v = blk_vert[v0 + k];
        V.xy[k] = xy[v];
        if (MASS && P.u_prev) V.up[k] = P.u_prev[v];
        if (P.bdval) {
          const int b = row_bd[v];
          V.bv[k] = b >= 0 ? P.bdval[b] : 0.0;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < AS_EPT; j++)
      if (t + j * AS_NT < ne) store(in[j], t + j * AS_NT);
  }
  __syncthreads();
  // pass 2: rows; each row's thread owns its slots of the accumulators and
  // clears them itself (the staged pass-1 values are no longer read)
  if (threadIdx.x >= AS_ROWS || row >= npad) return;
  double *my = vl + threadIdx.x;  // slot k at my[AS_ROWS * k]: no bank conflicts
  for (int k = 0; k < wmax + 2; k++) my[AS_ROWS * k] = 0.0;
  // compact, mass term: sum_k ipw_k h_pk h_qk of the rule (rho = 1)
  double mref[3][3];
  if (COMPACT && MASS) {
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 3; q++) mref[p][q] = 0.0;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const double h[3] = {shape(0, rule.xi[k], rule.eta[k]), shape(1, rule.xi[k], rule.eta[k]),
                           shape(2, rule.xi[k], rule.eta[k])};
#pragma unroll
      for (int p = 0; p < 3; p++) {
        const double wr = rule.w[k] * 1.0 * h[p];
#pragma unroll
        for (int q = p; q < 3; q++) mref[p][q] += wr * h[q];
      }
    }
  }
  double d = 1.0, acc = 0.0;  // padding rows: identity, rhs 0
  if (row < n) {
    if (bi >= 0) {
      acc = P.bdval ? P.bdval[bi] : 0.0;  // (b,b,1.0) of ImposeDirichlet.h:46-50
    } else {
      auto add = [&](uint32_t w) {
        const int p = (int)(w & 3u), le = (int)(w >> 20);
        double arow[3];
        if (COMPACT) {
          const int loc = eloc[le];
          double ex[3], ey[3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            const double2 c2 = V.xy[(loc >> (BLK_VERT_BITS * q)) & (BLK_MAX_VERTS - 1)];
            ex[q] = c2.x;
            ey[q] = c2.y;
          }
          const ElemGeo G = elem_geo(ex, ey);
          const double c = ec[le], dxp = pick3(G.dx, p), dyp = pick3(G.dy, p);
#pragma unroll
          for (int q = 0; q < 3; q++) {
            double val = c * (dxp * G.dx[q] + dyp * G.dy[q]);
            if (MASS && P.mass_coef != 0.0) {
              const int lo = min(p, q), hi = max(p, q);
              const double mr =
                  lo == 0 ? pick3(mref[0], hi) : (lo == 1 ? pick3(mref[1], hi) : mref[2][2]);
              val += P.mass_coef * (mr * G.ad);
            }
            arow[q] = val;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 3; q++) arow[q] = ae[6 * le + sym3(p, q)];
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const uint32_t code = (w >> (2 + SLOT_BITS * q)) & 63u;
          const int sl = code < PAIR_DIAG ? (int)code : wmax + (int)(code - PAIR_DIAG);
          my[AS_ROWS * sl] += arow[q];
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
  o.rhs[row] = acc;
}

// CSR (post-BC) -> SELL-64 off-diagonals + diagonal (FP64).
// Every entry of the slice is written, the padding ones with 0, so the
// destination needs no clearing beforehand (serving reuses device memory).
__global__ void to_sell_kernel(int n, const int *__restrict__ rp,
                               const int *__restrict__ col,
                               const double *__restrict__ val,
                               const int *__restrict__ slice_ptr,
                               double *__restrict__ sv,
                               double *__restrict__ diag,
                               int npad) {
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
}

// dL/dkappa_{e,k} = -scale * w_k * grad(lam~) . grad(u) on element e, with
// lam~ = lam zeroed on boundary rows (the gradient ImposeDirichletGrad drops,
// ImposeDirichlet.h:63-93).  One thread per element, nq outputs each.
// PAIR: u and lam come interleaved, ul[v] = (u_v, lam_v): one 16-byte
// gather per vertex instead of two 8-byte ones.
template <bool PAIR>
__global__ void kappa_grad_kernel(int nelem, QuadRule rule,
                                  const int *__restrict__ conn,
                                  const double2 *__restrict__ xy,
                                  const int *__restrict__ row_bd,
                                  const double *__restrict__ u,
                                  const double *__restrict__ lam,
                                  const double2 *__restrict__ ul, double scale,
                                  double *__restrict__ gk) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  int v[3];
  double ex[3], ey[3], uu[3], ll[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v[k] = conn[(size_t)k * nelem + e];
    const double2 c = xy[v[k]];
    ex[k] = c.x;
    ey[k] = c.y;
    if (PAIR) {
      const double2 w = ul[v[k]];
      uu[k] = w.x;
      ll[k] = row_bd[v[k]] >= 0 ? 0.0 : w.y;
    } else {
      uu[k] = u[v[k]];
      ll[k] = row_bd[v[k]] >= 0 ? 0.0 : lam[v[k]];
    }
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

// Loss L = sum_k (u[perm[obs_idx[k]]] - obs_val[k])^2 and its gradient
// g = dL/du (internal order; g zeroed beforehand).  One block, fixed order.
// loss[1] flags an index outside [0, n) or one that occurs twice: the
// observations first claim their rows in g (g[r] = k + 1), and one whose
// claim did not stick shares its row with another.  (A shared row would
// need its terms summed; the entry points require distinct indices.)
__global__ void __launch_bounds__(1024)
    loss_kernel(int n, const int *__restrict__ perm, const int *__restrict__ idx,
                const double *__restrict__ obs, int nobs,
                const double *__restrict__ u, double *__restrict__ g,
                double *__restrict__ loss) {
  __shared__ double sh[16];
  double acc = 0.0, bad = 0.0;
  for (int k = threadIdx.x; k < nobs; k += blockDim.x) {
    const int id = idx[k];
    if (id < 0 || id >= n) {  // caller data: never index with it unchecked
      bad = 1.0;
      continue;
    }
    g[perm[id]] = (double)(k + 1);
  }
  __syncthreads();  // the claims are visible to the whole block
  for (int k = threadIdx.x; k < nobs; k += blockDim.x) {
    const int id = idx[k];
    if (id >= 0 && id < n && g[perm[id]] != (double)(k + 1)) bad = 1.0;
  }
  __syncthreads();  // every check has read its claim
  for (int k = threadIdx.x; k < nobs; k += blockDim.x) {
    const int id = idx[k];
    if (id < 0 || id >= n) continue;
    const int r = perm[id];
    const double d = u[r] - obs[k];
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

void launch_assemble_to(const System &s, const double2 *xy, const admfem_problem &p,
                        const SellOut &o, hipStream_t st) {
  if (s.degree == 2) {
    launch_assemble_p2(s, xy, p, o, st);
    return;
  }
  const Mesh &m = *s.mesh;
  const bool mass = p.mass_coef != 0.0 || p.u_prev;
  const bool compact = p.rho == nullptr;
  const int me = (std::max(s.asm_max_e, 1) + 1) / 2 * 2, mv = std::max(s.asm_max_v, 1);
  const int w = std::max(s.sell_wmax, 1);
  // doubles: per-element arrays, staged vertex values read by pass 1 only,
  // row accumulators (the last two share one region)
  const size_t per_e = compact ? (size_t)me * 4 + (size_t)mv * 2 + (size_t)me / 2
                               : (size_t)me * 9;
  const size_t verts = (size_t)mv * ((compact ? 0 : 2) + (mass ? 1 : 0) + (p.bdval ? 1 : 0));
  const size_t rows = (size_t)AS_ROWS * (w + 2);
  const size_t lds = (per_e + std::max(verts, rows)) * sizeof(double);
  REQUIRE(lds <= 160 * 1024, ADMFEM_ERR_UNSUPPORTED, "assembly block too dense for LDS");
  const unsigned grid = nblk(s.npad, AS_ROWS);
  static_assert(AS_ROWS == 64, "an assembly block is one SELL slice");
  auto go = [&](auto kern) {
    kern<<<grid, AS_NT, lds, st>>>(s.n, s.npad, m.rule, xy, s.d_blk_eptr.p, s.d_blk_rec.p,
                                   s.d_blk_vptr.p, s.d_blk_vert.p, m.d_n2e_ptr.p, s.d_pairs.p,
                                   s.d_row_bd.p, s.d_slice_ptr.p, p, o, me, mv, w);
  };
  if (m.rule.nq == 3) {
    if (mass) compact ? go(assemble_kernel<3, true, true>) : go(assemble_kernel<3, true, false>);
    else compact ? go(assemble_kernel<3, false, true>) : go(assemble_kernel<3, false, false>);
  } else if (m.rule.nq == 6) {
    if (mass) compact ? go(assemble_kernel<6, true, true>) : go(assemble_kernel<6, true, false>);
    else compact ? go(assemble_kernel<6, false, true>) : go(assemble_kernel<6, false, false>);
  } else {
    REQUIRE(m.rule.nq == 1, ADMFEM_ERR_UNSUPPORTED, "quadrature rule not built");
    if (mass) compact ? go(assemble_kernel<1, true, true>) : go(assemble_kernel<1, true, false>);
    else compact ? go(assemble_kernel<1, false, true>) : go(assemble_kernel<1, false, false>);
  }
  HIPCHECK(hipGetLastError());
}

void launch_assemble(System &s, const admfem_problem &p, hipStream_t st) {
  SellOut o{s.d_sell_val.p, s.d_diag.p, s.rhs.p};
  launch_assemble_to(s, s.mesh->xy.p, p, o, st);
}

void launch_to_sell(System &s, hipStream_t st) {
  to_sell_kernel<<<nblk(s.npad, 256), 256, 0, st>>>(s.n, s.d_rp.p, s.d_col.p, s.val.p,
                                                    s.d_slice_ptr.p, s.d_sell_val.p,
                                                    s.d_diag.p, s.npad);
  HIPCHECK(hipGetLastError());
}

void launch_kappa_gradient_to(const System &s, const double2 *xy, const double *u_int,
                              const double *lam_int, double scale, double *gk, hipStream_t st) {
  if (s.degree == 2) {
    launch_kappa_gradient_p2(s, xy, u_int, lam_int, nullptr, scale, gk, st);
    return;
  }
  const Mesh &m = *s.mesh;
  kappa_grad_kernel<false><<<nblk(m.nelem, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, xy, s.d_row_bd.p, u_int, lam_int, nullptr, scale, gk);
  HIPCHECK(hipGetLastError());
}

namespace {
// dst[idx[i]] = (a[i], b[i]): one 16-byte write per entry
__global__ void scatter_pair_kernel(const int *__restrict__ idx, const double *__restrict__ a,
                                    const double *__restrict__ b, double2 *__restrict__ dst,
                                    int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = make_double2(a[i], b[i]);
}
}  // namespace

void launch_kappa_gradient_user(const System &s, const double *u, const double *lam,
                                double2 *ul_int, double scale, double *gk, hipStream_t st) {
  const Mesh &m = *s.mesh;
  if (s.n > 0)
    scatter_pair_kernel<<<nblk(s.n, 256), 256, 0, st>>>(s.d_uperm(), u, lam, ul_int, s.n);
  if (s.degree == 2) {
    launch_kappa_gradient_p2(s, m.xy.p, nullptr, nullptr, ul_int, scale, gk, st);
    return;
  }
  kappa_grad_kernel<true><<<nblk(m.nelem, 256), 256, 0, st>>>(
      m.nelem, m.rule, m.conn.p, m.xy.p, s.d_row_bd.p, nullptr, nullptr, ul_int, scale, gk);
  HIPCHECK(hipGetLastError());
}

void launch_kappa_gradient(const System &s, const double *u_int,
                           const double *lam_int, double scale, double *gk,
                           hipStream_t st) {
  launch_kappa_gradient_to(s, s.mesh->xy.p, u_int, lam_int, scale, gk, st);
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
  loss_kernel<<<1, 1024, 0, st>>>(s.n, s.d_uperm(), obs_idx, obs_val, nobs, u_int, g_int,
                                  loss2);
  HIPCHECK(hipGetLastError());
}

namespace {
__global__ void scatter_kernel(const int *__restrict__ idx, const double *__restrict__ src,
                               double *__restrict__ dst, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}
}  // namespace

void launch_scatter(const int *idx, const double *src, double *dst, int n, hipStream_t st) {
  if (n <= 0) return;
  scatter_kernel<<<nblk(n, 256), 256, 0, st>>>(idx, src, dst, n);
  HIPCHECK(hipGetLastError());
}

double loss_and_grad(const System &s, const double *u_int, const int *obs_idx,
                     const double *obs_val, int nobs, double *g_int, double *loss2,
                     hipStream_t st) {
  launch_loss(s, u_int, obs_idx, obs_val, nobs, g_int, loss2, st);
  double h[2] = {0.0, 0.0};
  HIPCHECK(hipMemcpyAsync(h, loss2, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  REQUIRE(h[1] == 0.0, ADMFEM_ERR_ARG,
          "observation index out of range [0, n) or repeated");
  return h[0];
}

}  // namespace admfem
