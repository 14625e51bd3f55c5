// The fused P2 (degree 2) system: assembly of
//   A = mass_coef * M(rho) + lap_coef * K(kappa),
//   rhs = src_coef * F(f) + M(rho) u_prev - A[:, bd] bdval  (interior rows),
//   A[b, :] = e_b, rhs[b] = bdval  (boundary rows)
// over the 6-dof quadratic triangle, written straight into the solver format
// (SELL-64 off-diagonals, diagonal, rhs) -- what the P2 branches of
// FemLaplaceScalar / ComputeFemMassMatrixMfem / FemSourceScalar followed by
// ImposeDirichlet produce in the reference (FemLaplaceScalar.h:3-26,
// ComputeFemMassMatrixMfem.h:4-29, FemSourceScalar.h:4-15,
// ImposeDirichlet.h:27-60) -- and the adjoint gradient dL/dkappa of a
// forward + adjoint pair.
//
// Assembly: a block owns one 64-row SELL slice.  Pass 1, one thread per
// element incident to the block's interior rows: the element's symmetric
// 6x6 matrix (packed upper triangle, 21 values) and its 6 lifted load terms
// into LDS.  Pass 2, one thread per row: its (row, element) pair words in
// ascending element order add the element's matrix row into the row's SELL
// slots (LDS accumulators) and its load term into the rhs.  One writer per
// value, fixed summation order, no atomics.
//
// Gradient: dL/dkappa_{e,k} = -scale * w_k * grad(lam~) . grad(u) at Gauss
// point k of element e, lam~ = lam with the boundary rows zeroed (the
// gradient ImposeDirichletGrad drops, ImposeDirichlet.h:63-93) -- the sum
// over the element's 36 triplets of FemLaplaceScalar_backward
// (FemLaplaceScalar.h:28-54) in closed form.  One thread per element.
//
// Geometry and Gauss weights (w_k = ipw_k * |det|, Common.cpp:116 with the
// area in closed form) come from the vertex coordinates given (the mesh's,
// or a serving request's own, internal node order) through the mesh's
// vertex connectivity; kappa, rho and f are in the user element order.
#include "p2basis.h"

namespace admfem {

namespace {

constexpr int P2A_NT = 256;
constexpr int P2_SYM = 21;  // packed upper triangle of a 6x6 matrix

// index of (p, q) in the packed upper triangle, rows 0..5
__device__ __forceinline__ int sym6(int p, int q) {
  const int lo = min(p, q), hi = max(p, q);
  return (lo * (11 - lo)) / 2 + hi;
}

// Vertex coordinates of element e (from xy: the mesh's own, or a serving
// request's), the gradients of its barycentric functions, and in G.area the
// factor |det| = 2 * area that turns a reference weight into the Gauss
// weight (w = ipw * area / 0.5, Common.cpp:116).
__device__ __forceinline__ ElemGeom p2_elem_geom(const int *__restrict__ conn, int nelem, int e,
                                                 const double2 *__restrict__ xy) {
  double ex[3], ey[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double2 c = xy[conn[(size_t)k * nelem + e]];
    ex[k] = c.x;
    ey[k] = c.y;
  }
  ElemGeom G;
  elem_grads(ex[0], ey[0], ex[1], ey[1], ex[2], ey[2], G.gx, G.gy);
  G.area = fabs((ex[1] - ex[0]) * (ey[2] - ey[0]) - (ex[2] - ex[0]) * (ey[1] - ey[0]));
  return G;
}

// v[0] = v[1], ..., v[N-2] = v[N-1]: the next Gauss point's value moves to
// the front (constant indices only, so the array stays in registers)
template <int N>
__device__ __forceinline__ void shift_n(double (&v)[N]) {
#pragma unroll
  for (int i = 0; i + 1 < N; i++) v[i] = v[i + 1];
}

// Element e's matrix a (packed) and lifted load terms r.  Per Gauss point
// the reference's factors: (grad h_p . grad h_q) kappa w, h_p h_q rho w,
// f h_p w; summed over the points in order.
template <int NQ, bool MASS>
__device__ __forceinline__ void p2_element(int e, int nelem, const QuadRule &rule,
                                           const int *__restrict__ conn,
                                           const double2 *__restrict__ xy,
                                           const int *__restrict__ conn2,
                                           const int *__restrict__ row_bd,
                                           const admfem_problem &P, double (&a)[P2_SYM],
                                           double (&r)[6]) {
  const ElemGeom G = p2_elem_geom(conn, nelem, e, xy);
  double mm[MASS ? P2_SYM : 1];
#pragma unroll
  for (int i = 0; i < P2_SYM; i++) a[i] = 0.0;
  if (MASS) {
#pragma unroll
    for (int i = 0; i < P2_SYM; i++) mm[i] = 0.0;
  }
#pragma unroll
  for (int p = 0; p < 6; p++) r[p] = 0.0;
  // Without a mass term every Gauss point's coefficients are requested at
  // once, before the loop (one memory latency per element instead of one
  // per point: 639 -> 556 us at 1024^2); with one, the registers go to the
  // mass matrix and each point loads its own (preloading cost occupancy:
  // 991 -> 1385 us).
  double kq[NQ], fq[NQ];
  if (!MASS) {
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const size_t g = (size_t)NQ * e + k;
      kq[k] = P.kappa ? P.kappa[g] : 1.0;
      fq[k] = P.f ? P.f[g] : 0.0;
    }
  }
#pragma unroll 1  // one Gauss point's basis in registers at a time (occupancy)
  for (int k = 0; k < NQ; k++) {
    const size_t g = (size_t)NQ * e + k;
    const double w = rule.w[k] * G.area;
    double gx[6], gy[6], h[6];
    p2_grad(rule.xi[k], rule.eta[k], G, gx, gy);
    p2_shape(rule.xi[k], rule.eta[k], h);
    const double kw = (MASS ? (P.kappa ? P.kappa[g] : 1.0) : kq[0]) * w;
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
      for (int q = p; q < 6; q++) a[sym6(p, q)] += (gx[p] * gx[q] + gy[p] * gy[q]) * kw;
    if (MASS) {
      const double rw = (P.rho ? P.rho[g] : 1.0) * w;
#pragma unroll
      for (int p = 0; p < 6; p++)
#pragma unroll
        for (int q = p; q < 6; q++) mm[sym6(p, q)] += h[p] * h[q] * rw;
    }
    if (P.f) {
      const double fw = (MASS ? P.f[g] : fq[0]) * w;
#pragma unroll
      for (int p = 0; p < 6; p++) r[p] += h[p] * fw;
    }
    if (!MASS) {
      shift_n(kq);
      shift_n(fq);
    }
  }
#pragma unroll
  for (int i = 0; i < P2_SYM; i++) {
    a[i] *= P.lap_coef;
    if (MASS) a[i] += P.mass_coef * mm[i];
  }
#pragma unroll
  for (int p = 0; p < 6; p++) r[p] *= P.src_coef;
  if (!(MASS && P.u_prev) && !P.bdval) return;
  int d[6];
#pragma unroll
  for (int q = 0; q < 6; q++) d[q] = conn2[(size_t)q * nelem + e];
  if (MASS && P.u_prev) {
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const double up = P.u_prev[d[q]];
#pragma unroll
      for (int p = 0; p < 6; p++) r[p] += mm[sym6(p, q)] * up;
    }
  }
  if (P.bdval) {  // Dirichlet lifting of the boundary columns
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int b = row_bd[d[q]];
      const double bv = b >= 0 ? P.bdval[b] : 0.0;
      if (bv != 0.0) {
#pragma unroll
        for (int p = 0; p < 6; p++) r[p] -= a[sym6(p, q)] * bv;
      }
    }
  }
}

// LDS, in doubles: [max_e][21] element matrices, [max_e][6] load terms,
// [wmax + 2][AS_ROWS] row accumulators (slot-major: slots [0, wmax) the
// row's SELL positions, wmax its diagonal, wmax + 1 a sink for lifted
// boundary columns).
template <int NQ, bool MASS>
__global__ void __launch_bounds__(P2A_NT)
    assemble_p2_kernel(int n, int nelem, QuadRule rule, const int *__restrict__ conn,
                       const double2 *__restrict__ xy, const int *__restrict__ conn2,
                       const int *__restrict__ blk_eptr, const int *__restrict__ blk_elem,
                       const int *__restrict__ inc_ptr,
                       const uint64_t *__restrict__ pairs, const int *__restrict__ row_bd,
                       const int *__restrict__ slice_ptr, admfem_problem P, SellOut o,
                       int max_e, int wmax) {
  extern __shared__ double lds[];
  double *ae = lds;
  double *rl = ae + P2_SYM * (size_t)max_e;
  double *racc = rl + 6 * (size_t)max_e;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int e0 = blk_eptr[blk], ne = blk_eptr[blk + 1] - e0;
  for (int t = threadIdx.x; t < ne; t += P2A_NT) {
    double a[P2_SYM], r[6];
    p2_element<NQ, MASS>(blk_elem[e0 + t], nelem, rule, conn, xy, conn2, row_bd, P, a, r);
#pragma unroll
    for (int i = 0; i < P2_SYM; i++) ae[P2_SYM * t + i] = a[i];
#pragma unroll
    for (int p = 0; p < 6; p++) rl[6 * t + p] = r[p];
  }
  __syncthreads();
  if (threadIdx.x >= AS_ROWS) return;
  const int row = blk * AS_ROWS + (int)threadIdx.x;
  double *my = racc + threadIdx.x;  // slot k at my[AS_ROWS * k]: no bank conflicts
  for (int k = 0; k < wmax + 2; k++) my[AS_ROWS * k] = 0.0;
  double d = 1.0, acc = 0.0;  // padding rows: identity, rhs 0
  if (row < n) {
    const int bi = row_bd[row];
    if (bi >= 0) {
      acc = P.bdval ? P.bdval[bi] : 0.0;  // (b,b,1.0) of ImposeDirichlet.h:46-50
    } else {
      const int t0 = inc_ptr[row], t1 = inc_ptr[row + 1];
      auto add = [&](uint64_t w) {
        const int p = (int)(w & 7u), le = (int)(w >> P2_PAIR_ELEM);
        const double *arow = ae + P2_SYM * le;
#pragma unroll
        for (int q = 0; q < 6; q++) {
          const uint32_t code = (uint32_t)(w >> (P2_PAIR_CODE0 + SLOT_BITS * q)) & 63u;
          const int sl = code < PAIR_DIAG ? (int)code : wmax + (int)(code - PAIR_DIAG);
          my[AS_ROWS * sl] += arow[sym6(p, q)];
        }
        acc += rl[6 * le + p];
      };
      // the row's first pair words requested together (a vertex dof has
      // about six incidences, an edge dof two), then added in order
      constexpr int PW = 8;
      uint64_t wb[PW];
#pragma unroll
      for (int k = 0; k < PW; k++) wb[k] = t0 + k < t1 ? pairs[t0 + k] : 0ull;
#pragma unroll
      for (int k = 0; k < PW; k++)
        if (t0 + k < t1) add(wb[k]);
      for (int t = t0 + PW; t < t1; t++) add(pairs[t]);
      d = my[AS_ROWS * wmax];
    }
  }
  const int s0 = slice_ptr[blk], width = (slice_ptr[blk + 1] - s0) >> 6;  // block = slice
  const int off = s0 + (row & 63);
  for (int k = 0; k < width; k++) o.sv[off + 64 * k] = my[AS_ROWS * k];
  o.diag[row] = d;
  o.rhs[row] = acc;
}

// dL/dkappa (see the file comment); u, lam in the internal dof order, as
// separate vectors or as (u, lam) pairs (PAIR).
template <int NQ, bool PAIR>
__global__ void __launch_bounds__(256)
    kappa_grad_p2_kernel(int nelem, QuadRule rule, const int *__restrict__ conn,
                         const double2 *__restrict__ xy, const int *__restrict__ conn2, const int *__restrict__ row_bd,
                         const double *__restrict__ u, const double *__restrict__ lam,
                         const double2 *__restrict__ ul, double scale,
                         double *__restrict__ gk) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  double uu[6], ll[6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int r = conn2[(size_t)q * nelem + e];
    const bool bd = row_bd[r] >= 0;
    if (PAIR) {
      const double2 v = ul[r];
      uu[q] = v.x;
      ll[q] = bd ? 0.0 : v.y;
    } else {
      uu[q] = u[r];
      ll[q] = bd ? 0.0 : lam[r];
    }
  }
  const ElemGeom G = p2_elem_geom(conn, nelem, e, xy);
#pragma unroll
  for (int k = 0; k < NQ; k++) {
    double gx[6], gy[6];
    p2_grad(rule.xi[k], rule.eta[k], G, gx, gy);
    double ux = 0.0, uy = 0.0, lx = 0.0, ly = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      ux += gx[q] * uu[q];
      uy += gy[q] * uu[q];
      lx += gx[q] * ll[q];
      ly += gy[q] * ll[q];
    }
    const size_t g = (size_t)NQ * e + k;
    gk[g] = -scale * (rule.w[k] * G.area) * (lx * ux + ly * uy);
  }
}

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_assemble_p2(const System &s, const double2 *xy, const admfem_problem &p,
                        const SellOut &o, hipStream_t st) {
  const Mesh &m = *s.mesh;
  const bool mass = p.mass_coef != 0.0 || p.u_prev;
  const int me = std::max(s.asm_max_e, 1), w = std::max(s.sell_wmax, 1);
  const size_t lds = ((size_t)(P2_SYM + 6) * me + (size_t)AS_ROWS * (w + 2)) * sizeof(double);
  REQUIRE(lds <= 160 * 1024, ADMFEM_ERR_UNSUPPORTED, "P2 assembly block too dense for LDS");
  const unsigned grid = (unsigned)(s.npad / AS_ROWS);
  auto go = [&](auto kern) {
    kern<<<grid, P2A_NT, lds, st>>>(s.n, m.nelem, m.rule, m.conn.p, xy, s.d_conn2.p,
                                    s.d_blk_eptr.p, s.d_blk_elem.p, s.d_inc_ptr.p,
                                    s.d_pairs2.p, s.d_row_bd.p, s.d_slice_ptr.p, p, o, me, w);
  };
  switch (m.rule.nq) {
    case 1: mass ? go(assemble_p2_kernel<1, true>) : go(assemble_p2_kernel<1, false>); break;
    case 3: mass ? go(assemble_p2_kernel<3, true>) : go(assemble_p2_kernel<3, false>); break;
    case 6: mass ? go(assemble_p2_kernel<6, true>) : go(assemble_p2_kernel<6, false>); break;
    default: REQUIRE(false, ADMFEM_ERR_UNSUPPORTED, "quadrature rule not built");
  }
  HIPCHECK(hipGetLastError());
}

void launch_kappa_gradient_p2(const System &s, const double2 *xy, const double *u_int,
                              const double *lam_int, const double2 *ul_int, double scale,
                              double *gk, hipStream_t st) {
  const Mesh &m = *s.mesh;
  const unsigned grid = nblk(m.nelem, 256);
  auto go = [&](auto kern) {
    kern<<<grid, 256, 0, st>>>(m.nelem, m.rule, m.conn.p, xy, s.d_conn2.p,
                               s.d_row_bd.p, u_int, lam_int, ul_int, scale, gk);
  };
  const bool pair = ul_int != nullptr;
  switch (m.rule.nq) {
    case 1: pair ? go(kappa_grad_p2_kernel<1, true>) : go(kappa_grad_p2_kernel<1, false>); break;
    case 3: pair ? go(kappa_grad_p2_kernel<3, true>) : go(kappa_grad_p2_kernel<3, false>); break;
    case 6: pair ? go(kappa_grad_p2_kernel<6, true>) : go(kappa_grad_p2_kernel<6, false>); break;
    default: REQUIRE(false, ADMFEM_ERR_UNSUPPORTED, "quadrature rule not built");
  }
  HIPCHECK(hipGetLastError());
}

}  // namespace admfem
