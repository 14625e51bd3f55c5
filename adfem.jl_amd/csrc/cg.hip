// Jacobi-preconditioned conjugate gradients on the SELL-64 form of the
// Dirichlet-eliminated system -- the GPU replacement of the reference's
// direct sparse solve K\F (ADCME SparseSolver, Eigen SparseLU [ext]).
//
// Two precisions:
//  * FP64: matrix, vectors and reductions in FP64.
//  * mixed: CG on the symmetrically scaled system S A S (S = diag^-1/2,
//    unit diagonal; the same iterates as Jacobi-PCG in exact arithmetic)
//    with the matrix and the inner vectors stored in FP32 (FP64
//    arithmetic) and the solution correction y accumulated in FP32
//    (x = x64 + S y); whenever the FP32 residual has fallen by
//    `replace_delta` (or looks converged, or has stalled) the correction is
//    folded into the FP64 solution and the residual is replaced by the true
//    FP64 residual b - A x, computed with the FP64 matrix ("reliable
//    updates", van der Vorst & Ye 2000; Clark et al. 2010).  Convergence is
//    declared only on a true FP64 residual.
//
// One iteration is one wide kernel and one single-block kernel:
//   step:    applies the pending step and forms the next direction,
//            r' = r - alpha q, x += alpha p, p' = M^-1 r' + beta p, q' = A p'
//            (FP64: a neighbour's p' formed on the fly from its own vectors;
//            mixed: q' = Â r' + beta q, a neighbour contributing its
//            r' = r - alpha q, and each replacement recomputes q = Â p), and
//            the block partials of every dot product the next scalars need
//   scalar:  alpha' = r'.M^-1 r' / p'.q' and beta' = the same quadratic form
//            of r' - alpha' q' over r'.M^-1 r', its numerator expanded as
//            r'.M^-1 r' - 2 alpha' q'.M^-1 r' + alpha'^2 q'.M^-1 q' (exact for
//            the stored vectors up to rounding, one step ahead, so no second
//            reduction is needed before p'' is formed)
// so r.M^-1 r is reduced directly every iteration (never carried by a
// recurrence) and the iterates are standard PCG's (Hestenes-Stiefel) with
// both reductions of an iteration in one pass over the vectors.  The FP64
// form keeps z = D^-1 r and w = D^-1 q instead of r and q (r = D z), so a
// neighbour's p' = z - alpha w + beta p needs no diagonal; the mixed form's
// unit diagonal needs none either.  The vectors a step gathers are
// double-buffered: it reads buffer it & 1 of the solver's own step count
// and writes the other.  The mixed form checks for a
// replacement every eighth iteration (fold x, replace r, scalar); the
// pending alpha and beta carry over it (the direction is kept).
//
// No atomics and no fences anywhere: on this part a cross-block atomic or a
// device-scope fence costs tens of ns per block, serialised, which dwarfs a
// memory-bound SpMV.  Partials are plain stores made visible by the kernel
// boundary; one single-block kernel turns them into the next scalars, so the
// wide kernels only read a handful of words and never wait on each other.
// Every kernel first reads the solver status and returns at once when the
// solve has finished, so the host enqueues iterations in chunks and polls
// the status between chunks.
#include <chrono>
#include <cmath>

#include "batched.h"
#include "geom.h"
#include "sell.h"

namespace admfem {

enum { ST_RUN = 0, ST_CONV = 1, ST_MAXIT = 2, ST_NAN = 3, ST_BREAK = 4 };

struct CGScalars {
  double rtol2, delta2;  // set by the host
  double bb, tol2;       // ||b||^2, rtol^2 ||b||^2
  double rz;             // r.D^-1 r of the current residual (mixed: scaled r.r)
  double alpha, beta;    // the pending step's
  double rr;             // FP64: ||r||^2 of the current residual (recursive);
                         // mixed: true ||b - A x||^2 at the last replacement
  double rr_ref;         // mixed: true ||b - A x||^2 at the last replacement
  double rz_ref;         // mixed: scaled r.r at the last replacement
  double rz_min;         // mixed: smallest scaled r.r since then
  int mixed;
  int status;            // ST_*
  int it;                // step kernels run (each one SpMV)
  int replace;           // mixed: frozen (converged-looking) until a replacement
  int nrepl;
  int stall;             // mixed: iterations without a new rz_min
  const double *b;       // right-hand side of the solve (read through here by
                         // the replacement kernel, so replayed graphs do not
                         // depend on which b a solve is given)
};


constexpr int NT = 256;   // threads per wide block = 4 slices of 64 rows
constexpr int RT = 1024;  // threads of the single-block scalar kernels

struct CGWork {
  int npad = 0, nblocks = 0;
  DevBuf<double> x;                          // FP64 solution
  DevBuf<double> z64[2], w64[2], p64[2];     // FP64: D^-1 r, D^-1 q, direction
  DevBuf<float> ys32, r32[2], p32[2], q32[2];  // mixed: correction, r, p, q
  DevBuf<double> dsq;    // mixed: 1 / sqrt(a_ii)
  DevBuf<float> sv32s;   // mixed: FP32 scaled off-diagonals a_ij dsq_i dsq_j
  DevBuf<double> part;  // 5 x nblocks block partials
  DevBuf<CGScalars> sc;
  CGScalars *h_sc = nullptr;  // pinned
  std::vector<double> h_part;
  // GRAPH_ITERS iterations (from an iteration index divisible by
  // GRAPH_ITERS) captured once as a graph of kernel nodes and replayed: the
  // launches of a small system cost more than its kernels.  Captured on a
  // private stream (the caller's may be the null stream, which cannot be
  // captured), replayed on the caller's; `gkey` holds every pointer and
  // setting baked into the nodes.
  hipStream_t cap = nullptr;
  hipGraphExec_t graph = nullptr;
  std::vector<const void *> gkey;
  ~CGWork() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    if (h_sc) (void)hipHostFree(h_sc);
  }
};

void cg_work_free(CGWork *w) { delete w; }

namespace {

// Blocks are dispatched round-robin over the 8 XCDs; map them so that each
// XCD works on one contiguous eighth of the rows (its L2 then holds that
// eighth's gather window).  gridDim.x % 8 == 0.
__device__ __forceinline__ int my_row() {
  const int b = blockIdx.x, nb = gridDim.x;
  return ((b & 7) * (nb >> 3) + (b >> 3)) * NT + threadIdx.x;
}

// This block's sums of NV values, stored to part[j * G + blockIdx.x].
template <int NV>
__device__ __forceinline__ void put_partials(const double (&v)[NV], double *part) {
  __shared__ double sh[NV][NT / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; j++) {
    double t = wave_sum(v[j]);
    if (lane == 0) sh[j][wv] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < NT / 64; k++) t += sh[j][k];
      part[(size_t)j * gridDim.x + blockIdx.x] = t;
    }
  }
}

// Single-block sums of NV partial arrays of length G (fixed order).
template <int NV>
__device__ __forceinline__ void sum_partials(const double *part, int G, double (&out)[NV]) {
  __shared__ double sh[NV][RT / 64];
  double a[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) a[j] = 0.0;
  // each thread adds part[k], part[k + RT], ... in that order; the loads
  // are issued 8 at a time before their adds, so they are in flight
  // together (a plain loop waits for every load before the next)
  int k = threadIdx.x;
  for (; k + 7 * RT < G; k += 8 * RT) {
    double v[NV][8];
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int j = 0; j < NV; j++) v[j][u] = part[(size_t)j * G + k + u * RT];
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int j = 0; j < NV; j++) a[j] += v[j][u];
  }
  for (; k < G; k += RT) {
#pragma unroll
    for (int j = 0; j < NV; j++) a[j] += part[(size_t)j * G + k];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; j++) {
    double t = wave_sum(a[j]);
    if (lane == 0) sh[j][wv] = t;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; j++) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < RT / 64; k++) t += sh[j][k];
    out[j] = t;
  }
}

// ---- FP64 Jacobi-PCG

// x0 = b on identity (boundary) rows, 0 elsewhere; r0 = b - A x0 (0 on those
// rows, b on the rest: interior rows of the eliminated system have no
// boundary columns) kept as z0 = D^-1 r0; w = p = 0 (with alpha = beta = 0
// the first step forms p = z0, q = A z0); partials of r.z, r.r, b.b.
__global__ void __launch_bounds__(NT)
    k_init(int n, int npad, const int *__restrict__ row_bd, const double *__restrict__ b,
           const double *__restrict__ diag, double *__restrict__ x, double *__restrict__ z,
           double *__restrict__ w, double *__restrict__ p, double *part) {
  const int i = my_row();
  double t[3] = {0.0, 0.0, 0.0};
  if (i < npad) {
    const double bi = i < n ? b[i] : 0.0;
    const bool ident = i >= n || row_bd[i] >= 0;
    const double ri = ident ? 0.0 : bi;
    const double di = diag[i];
    const double zi = di != 0.0 ? ri / di : 0.0;
    x[i] = (ident && i < n) ? bi : 0.0;
    z[i] = zi;
    w[i] = 0.0;
    p[i] = 0.0;
    t[0] = ri * zi;
    t[1] = ri * ri;
    t[2] = bi * bi;
  }
  put_partials<3>(t, part);
}

__global__ void __launch_bounds__(RT) s_init(int G, const double *part, CGScalars *S) {
  double t[3];
  sum_partials<3>(part, G, t);
  if (threadIdx.x != 0) return;
  S->rz = t[0];
  S->rr = t[1];
  S->rr_ref = t[1];
  S->rz_ref = t[0];
  S->rz_min = t[0];
  S->stall = 0;
  S->bb = t[2];
  S->tol2 = S->rtol2 * t[2];
  S->alpha = 0.0;
  S->beta = 0.0;
  S->it = 0;
  S->nrepl = 0;
  S->replace = 0;
  S->status = (!(t[1] == t[1]) || !isfinite(t[1])) ? ST_NAN
              : (t[1] <= S->tol2)                  ? ST_CONV
                                                   : ST_RUN;
}

// The pending step and the next direction of one row (FP64 form):
// z' = z - alpha w, p' = z' + beta p, with explicit fused multiply-adds so
// that a row and its neighbours evaluate it identically.
__device__ __forceinline__ double dir64(double z, double w, double p, double alpha,
                                        double beta, double &zn) {
  zn = fma(-alpha, w, z);
  return fma(beta, p, zn);
}

// x += alpha p; z' = z - alpha w; p' = z' + beta p; q = A p'; w' = D^-1 q;
// partials of r'.z', p'.q, r'.w' (= q.z'), q.w', r'.r' with r' = D z'
__global__ void __launch_bounds__(NT)
    k_step(int npad, const int *__restrict__ sp, const int *__restrict__ sc,
           const double *__restrict__ sv, const double *__restrict__ diag,
           double *z0, double *z1, double *w0, double *w1,
           double *p0, double *p1, double *__restrict__ x, double *part, const CGScalars *S) {
  if (S->status != ST_RUN) return;
  const double alpha = S->alpha, beta = S->beta;
  const bool odd = S->it & 1;
  const double *__restrict__ z = odd ? z1 : z0;
  const double *__restrict__ w = odd ? w1 : w0;
  const double *__restrict__ p = odd ? p1 : p0;
  const int i = my_row();
  double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < npad) {
    const double pi = p[i];
    double zi;
    const double pni = dir64(z[i], w[i], pi, alpha, beta, zi);
    x[i] = fma(alpha, pi, x[i]);
    const double di = diag[i];
    auto nb = [&](int c) { return double3{z[c], w[c], p[c]}; };
    auto pn = [&](double3 v) {
      double unused;
      return dir64(v.x, v.y, v.z, alpha, beta, unused);
    };
    // 32-bit columns, non-temporal like the FP64 values (16-bit offsets
    // measured slower here: 115 vs 97 us at 2048^2)
    const double q = di * pni + sell_dot<double>(i, sp, sc, sv, nb, pn);
    const double wi = di != 0.0 ? q / di : 0.0;
    (odd ? z0 : z1)[i] = zi;
    (odd ? p0 : p1)[i] = pni;
    (odd ? w0 : w1)[i] = wi;
    const double ri = di * zi;
    t[0] = ri * zi;
    t[1] = pni * q;
    t[2] = ri * wi;
    t[3] = q * wi;
    t[4] = ri * ri;
  }
  put_partials<5>(t, part);
}

__global__ void __launch_bounds__(RT) s_step(int G, const double *part, CGScalars *S,
                                             int maxit) {
  if (S->status != ST_RUN) return;
  double t[5];
  sum_partials<5>(part, G, t);
  if (threadIdx.x != 0) return;
  const double rz = t[0], pq = t[1], zq = t[2], qw = t[3], rr = t[4];
  S->it += 1;
  if (!(rr == rr) || !isfinite(rr) || !(rz == rz) || pq != pq) {
    S->status = ST_NAN;
    return;
  }
  S->rz = rz;
  S->rr = rr;
  if (pq > 0.0) {  // the pending step, also kept when the solve stops here
    const double alpha = rz / pq;
    S->alpha = alpha;
    S->beta = rz > 0.0 ? fmax(rz - 2.0 * alpha * zq + alpha * alpha * qw, 0.0) / rz : 0.0;
  }
  if (rr <= S->tol2) S->status = ST_CONV;
  else if (S->it >= maxit) S->status = ST_MAXIT;
  else if (!(pq > 0.0)) S->status = ST_BREAK;  // not SPD
}

// An FP64 solve whose recursive residual met the tolerance but whose true
// residual did not: the true residual replaces the recursive one,
// z = D^-1 (b - A x) in the current buffer, keeping the direction and the
// pending step; partials of r.z and r.r (x: the solution taken out of the
// vector buffer)
__global__ void __launch_bounds__(NT)
    k_restart(int n, int npad, const int *__restrict__ sp, const int *__restrict__ sc,
              const double *__restrict__ sv, const double *__restrict__ diag,
              const double *__restrict__ x, double *z0,
              double *z1, double *part, const CGScalars *S) {
  const double *__restrict__ b = S->b;
  double *__restrict__ z = (S->it & 1) ? z1 : z0;
  const int i = my_row();
  double t[2] = {0.0, 0.0};
  if (i < npad) {
    const double ax = diag[i] * x[i] + sell_dot<double>(i, sp, sc, sv, [&](int c) { return x[c]; });
    const double ri = (i < n ? b[i] : 0.0) - ax;
    const double di = diag[i];
    const double zi = di != 0.0 ? ri / di : 0.0;
    z[i] = zi;
    t[0] = ri * zi;
    t[1] = ri * ri;
  }
  put_partials<2>(t, part);
}

__global__ void __launch_bounds__(RT) s_restart(int G, const double *part, CGScalars *S,
                                                int maxit) {
  double t[2];
  sum_partials<2>(part, G, t);
  if (threadIdx.x != 0) return;
  S->rz = t[0];
  S->rr = t[1];
  S->nrepl += 1;
  if (!(t[1] == t[1]) || !isfinite(t[1])) S->status = ST_NAN;
  else if (t[1] <= S->tol2) S->status = ST_CONV;
  else S->status = S->it >= maxit ? ST_MAXIT : ST_RUN;
}

// ---- mixed precision, scaled form: plain CG on S A S (S = diag^-1/2,
// unit diagonal) with FP32 storage, the solution x = x64 + S y

// x0 = b on identity rows; r = S r0 (FP32); y = 0; p = q = 0; partials of
// the scaled r.r, the true ||r0||^2 and b.b
__global__ void __launch_bounds__(NT)
    k_minit(int n, int npad, const int *__restrict__ row_bd, const double *__restrict__ b,
            const double *__restrict__ dsq, double *__restrict__ x, float *__restrict__ ys,
            float *__restrict__ r, float *__restrict__ p, float *__restrict__ q, double *part) {
  const int i = my_row();
  double t[3] = {0.0, 0.0, 0.0};
  if (i < npad) {
    const double bi = i < n ? b[i] : 0.0;
    const bool ident = i >= n || row_bd[i] >= 0;
    const double ri = ident ? 0.0 : bi;
    x[i] = (ident && i < n) ? bi : 0.0;
    const float rv = (float)(ri * dsq[i]);
    ys[i] = 0.0f;
    r[i] = rv;
    p[i] = 0.0f;
    q[i] = 0.0f;
    t[0] = (double)rv * (double)rv;
    t[1] = ri * ri;
    t[2] = bi * bi;
  }
  put_partials<3>(t, part);
}

// The pending step and the next direction of one row (scaled form):
// r' = r - alpha q, p' = r' + beta p rounded as it is stored, with explicit
// fused multiply-adds, so q' = Â p' is the product with the stored p'.
__device__ __forceinline__ double dir32(float r, float q, float p, double alpha, double beta,
                                        double &rn) {
  rn = fma(-alpha, (double)q, (double)r);
  return (double)(float)fma(beta, (double)p, rn);
}

// y += alpha p; r' = r - alpha q; p' = r' + beta p; q' = (S A S) p', formed
// as Â r' + beta q (Â p = q), so a neighbour contributes its r' = r - alpha q
// (two gathered values); partials of r'.r', p'.q', r'.q', q'.q'
__global__ void __launch_bounds__(NT)
    k_mstep(int npad, const int *__restrict__ sp, const int *__restrict__ sc,
            const int16_t *__restrict__ sd, const float *__restrict__ sv, float *r0,
            float *r1, float *q0, float *q1,
            float *p0, float *p1, float *__restrict__ ys, double *part, const CGScalars *S) {
  if (S->status != ST_RUN || S->replace) return;
  const double alpha = S->alpha, beta = S->beta;
  const bool odd = S->it & 1;
  const float *__restrict__ r = odd ? r1 : r0;
  const float *__restrict__ q = odd ? q1 : q0;
  const float *__restrict__ p = odd ? p1 : p0;
  const int i = my_row();
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < npad) {
    const float pi = p[i], qi = q[i];
    double ri;
    const double pni = dir32(r[i], qi, pi, alpha, beta, ri);
    ys[i] = (float)fma(alpha, (double)pi, (double)ys[i]);
    auto nb = [&](int c) { return make_float2(r[c], q[c]); };
    auto rn = [&](float2 v) { return fma(-alpha, (double)v.y, (double)v.x); };
    // 16-bit column offsets when the layout has them (sd: uniform)
    const double a = fma(beta, (double)qi,
                         ri + (sd ? sell_dot16<float>(i, sp, sd, sv, nb, rn)
                                  : sell_dot<float>(i, sp, sc, sv, nb, rn)));
    const float rv = (float)ri, pv = (float)pni, qv = (float)a;
    (odd ? r0 : r1)[i] = rv;
    (odd ? p0 : p1)[i] = pv;
    (odd ? q0 : q1)[i] = qv;
    t[0] = (double)rv * (double)rv;
    t[1] = (double)pv * (double)qv;
    t[2] = (double)rv * (double)qv;
    t[3] = (double)qv * (double)qv;
  }
  put_partials<4>(t, part);
}

__global__ void __launch_bounds__(RT) s_mstep(int G, const double *part, CGScalars *S,
                                              int maxit) {
  if (S->status != ST_RUN || S->replace) return;
  double t[4];
  sum_partials<4>(part, G, t);
  if (threadIdx.x != 0) return;
  const double rz = t[0], pq = t[1], rq = t[2], qq = t[3];
  S->it += 1;
  if (!(rz == rz) || !isfinite(rz) || pq != pq) {
    S->status = ST_NAN;
    return;
  }
  S->rz = rz;
  if (rz < 0.98 * S->rz_min) {
    S->rz_min = rz;
    S->stall = 0;
  } else {
    S->stall += 1;
  }
  // a scaled residual that looks converged -- below the tolerance times the
  // scaled-to-true ratio of the last replacement -- freezes the solve until
  // the next replacement point, where the FP64 residual decides
  if (rz * S->rr_ref <= S->tol2 * S->rz_ref) S->replace = 1;
  else if (S->it >= maxit) S->status = ST_MAXIT;
  if (!(pq > 0.0)) {  // not SPD (or p = 0 with r = 0: the replacement decides)
    if (!S->replace) S->status = ST_BREAK;
    S->alpha = 0.0;
    S->beta = 0.0;
    return;
  }
  const double alpha = rz / pq;
  S->alpha = alpha;
  S->beta = rz > 0.0 ? fmax(rz - 2.0 * alpha * rq + alpha * alpha * qq, 0.0) / rz : 0.0;
}

// A replacement at a replacement point: when frozen (converged-looking),
// when the scaled residual fell by delta since the last one, or when it
// stalled.  The point's kernels read the same, unchanged scalars.
__device__ __forceinline__ bool replace_due(const CGScalars *S) {
  return S->status == ST_RUN &&
         (S->replace || S->rz <= S->delta2 * S->rz_ref || S->stall >= STALL_ITERS);
}

// when a replacement is due (or forced at the end): x += S y, y = 0 (each
// row reads and writes only its own entries)
__global__ void __launch_bounds__(NT)
    k_mfold(int npad, double *__restrict__ x, float *__restrict__ ys,
            const double *__restrict__ dsq, const CGScalars *S, int force) {
  if (!force && !replace_due(S)) return;
  const int i = my_row();
  if (i < npad) {
    x[i] += dsq[i] * (double)ys[i];
    ys[i] = 0.0f;
  }
}

// when a replacement is due: the true residual b - A x in FP64 (FP64
// matrix), handed to the FP32 iteration as r = S (b - A x) in the current
// buffer; the direction is kept and its product q = Â p recomputed (the step
// carries q by recurrence); partials of the scaled r.r and the true
// ||b - A x||^2
__global__ void __launch_bounds__(NT)
    k_mreplace(int n, int npad, const int *__restrict__ sp, const int *__restrict__ sc,
               const double *__restrict__ sv, const double *__restrict__ diag,
               const float *__restrict__ sv32, const double *__restrict__ x,
               const double *__restrict__ dsq, float *r0, float *r1, const float *p0,
               const float *p1, float *q0, float *q1, double *part, const CGScalars *S) {
  if (!replace_due(S)) return;
  const double *__restrict__ b = S->b;
  const bool odd = S->it & 1;
  float *__restrict__ r = odd ? r1 : r0;
  const float *__restrict__ p = odd ? p1 : p0;
  const int i = my_row();
  double t[2] = {0.0, 0.0};
  if (i < npad) {
    const double ax = diag[i] * x[i] + sell_dot<double>(i, sp, sc, sv, [&](int c) { return x[c]; });
    const double ri = (i < n ? b[i] : 0.0) - ax;
    const float rf = (float)(ri * dsq[i]);
    r[i] = rf;
    (odd ? q1 : q0)[i] =
        (float)((double)p[i] + sell_dot<float>(i, sp, sc, sv32, [&](int c) { return (double)p[c]; }));
    t[0] = (double)rf * (double)rf;
    t[1] = ri * ri;
  }
  put_partials<2>(t, part);
}

__global__ void __launch_bounds__(RT) s_replace(int G, const double *part, CGScalars *S,
                                                 int maxit) {
  if (!replace_due(S)) return;
  double t[2];
  sum_partials<2>(part, G, t);
  __syncthreads();  // every thread has read the decision's inputs
  if (threadIdx.x != 0) return;
  S->rz = t[0];
  S->rr = t[1];
  S->rr_ref = t[1];
  S->rz_ref = t[0];
  S->rz_min = t[0];
  S->stall = 0;
  S->replace = 0;
  S->nrepl += 1;
  if (!(t[1] == t[1]) || !isfinite(t[1])) S->status = ST_NAN;
  else if (t[1] <= S->tol2) S->status = ST_CONV;
  else if (S->it >= maxit) S->status = ST_MAXIT;
}

// partials of ||b - A x||^2 in FP64 (final check; summed on the host)
__global__ void __launch_bounds__(NT)
    k_true_residual(int n, int npad, const int *__restrict__ sp, const int *__restrict__ sc,
                    const double *__restrict__ sv, const double *__restrict__ diag,
                    const double *__restrict__ b, const double *__restrict__ x, double *part) {
  const int i = my_row();
  double v[1] = {0.0};
  if (i < npad) {
    const double ax = diag[i] * x[i] + sell_dot<double>(i, sp, sc, sv, [&](int c) { return x[c]; });
    const double ri = (i < n ? b[i] : 0.0) - ax;
    v[0] = ri * ri;
  }
  put_partials<1>(v, part);
}

template <typename VT>
__global__ void k_fill(VT *a, int n, VT v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace

static CGWork *ensure_work(System &s) {
  if (s.work && s.work->npad == s.npad) return s.work;
  if (s.work) cg_work_free(s.work);
  CGWork *w = new CGWork();
  try {
    w->npad = s.npad;
    int nb = (s.npad + NT - 1) / NT;
    nb = (nb + 7) / 8 * 8;  // XCD mapping needs a multiple of 8
    w->nblocks = nb;
    w->x.alloc(s.npad);
    for (int k = 0; k < 2; k++) {
      w->z64[k].alloc(s.npad);
      w->w64[k].alloc(s.npad);
      w->p64[k].alloc(s.npad);
      w->r32[k].alloc(s.npad);
      w->p32[k].alloc(s.npad);
      w->q32[k].alloc(s.npad);
    }
    w->ys32.alloc(s.npad);
    w->dsq.alloc(s.npad);
    w->sv32s.alloc(std::max(1, s.slice_ptr.back()));
    w->part.alloc((size_t)5 * nb);
    w->h_part.resize(nb);
    w->sc.alloc(1);
    HIPCHECK(hipMemset(w->sc.p, 0, sizeof(CGScalars)));
    HIPCHECK(hipHostMalloc((void **)&w->h_sc, sizeof(CGScalars), hipHostMallocDefault));
  } catch (...) {
    delete w;
    throw;
  }
  s.work = w;
  return w;
}

namespace {

// One step kernel of the given precision (it reads the current vector
// buffer, which the scalars name, and writes the other one).
void launch_step(System &s, CGWork *w, bool mixed, hipStream_t st) {
  const int nb = w->nblocks, np = s.npad;
  const int *sp = s.d_slice_ptr.p, *scol = s.d_sell_col.p;
  if (mixed)
    k_mstep<<<nb, NT, 0, st>>>(np, sp, scol, s.d_sell_d16.p, w->sv32s.p, w->r32[0].p,
                               w->r32[1].p,
                               w->q32[0].p, w->q32[1].p, w->p32[0].p, w->p32[1].p, w->ys32.p,
                               w->part.p, w->sc.p);
  else
    k_step<<<nb, NT, 0, st>>>(np, sp, scol, s.d_sell_val.p, s.d_diag.p,
                              w->z64[0].p, w->z64[1].p, w->w64[0].p, w->w64[1].p, w->p64[0].p,
                              w->p64[1].p, w->x.p, w->part.p, w->sc.p);
}

}  // namespace

// Instrumentation: the solver's step kernel (the SpMV with the iteration's
// vector updates) launched `repeat` times on the system's work vectors
// (r or z = 1, everything else 0, alpha = beta = 0, status running): every
// launch reads the same buffers, runs its full body and writes the same
// values to the other buffers.
void spmv_repeat(System &s, int precision, int repeat, hipStream_t st) {
  CGWork *w = ensure_work(s);
  CGScalars init{};
  HIPCHECK(hipMemcpyAsync(w->sc.p, &init, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));  // init lives on this stack frame
  const unsigned fb = (unsigned)((s.npad + 255) / 256);
  const bool mixed = precision == ADMFEM_PREC_MIXED;
  if (mixed) {  // the scaled-form kernel of the mixed solve
    batch_scale(s, s.d_sell_val.p, s.d_diag.p, w->dsq.p, w->sv32s.p, st);
    k_fill<float><<<fb, 256, 0, st>>>(w->r32[0].p, s.npad, 1.0f);
    k_fill<float><<<fb, 256, 0, st>>>(w->q32[0].p, s.npad, 0.0f);
    k_fill<float><<<fb, 256, 0, st>>>(w->p32[0].p, s.npad, 0.0f);
    k_fill<float><<<fb, 256, 0, st>>>(w->ys32.p, s.npad, 0.0f);
  } else {
    k_fill<double><<<fb, 256, 0, st>>>(w->z64[0].p, s.npad, 1.0);
    k_fill<double><<<fb, 256, 0, st>>>(w->w64[0].p, s.npad, 0.0);
    k_fill<double><<<fb, 256, 0, st>>>(w->p64[0].p, s.npad, 0.0);
    k_fill<double><<<fb, 256, 0, st>>>(w->x.p, s.npad, 0.0);
  }
  for (int k = 0; k < repeat; k++) launch_step(s, w, mixed, st);
  HIPCHECK(hipGetLastError());
}

namespace {

constexpr int GRAPH_ITERS = 8;  // a multiple of the replacement interval (8)

// One iteration with index `it` (the mixed form has a replacement point
// after every eighth iteration and after the last one allowed).
void enqueue_iteration(System &s, CGWork *w, bool mixed, int max_iter, int it,
                       hipStream_t st) {
  const int nb = w->nblocks, np = s.npad;
  CGScalars *S = w->sc.p;
  double *part = w->part.p;
  launch_step(s, w, mixed, st);
  if (mixed) {
    s_mstep<<<1, RT, 0, st>>>(nb, part, S, max_iter);
    if ((it & 7) == 7 || it + 1 == max_iter) {
      k_mfold<<<nb, NT, 0, st>>>(np, w->x.p, w->ys32.p, w->dsq.p, S, 0);
      k_mreplace<<<nb, NT, 0, st>>>(s.n, np, s.d_slice_ptr.p, s.d_sell_col.p, s.d_sell_val.p,
                                    s.d_diag.p, w->sv32s.p, w->x.p, w->dsq.p, w->r32[0].p,
                                    w->r32[1].p, w->p32[0].p, w->p32[1].p, w->q32[0].p,
                                    w->q32[1].p, part, S);
      s_replace<<<1, RT, 0, st>>>(nb, part, S, max_iter);
    }
  } else {
    s_step<<<1, RT, 0, st>>>(nb, part, S, max_iter);
  }
}

// Iterations it .. it + GRAPH_ITERS - 1 for an `it` divisible by
// GRAPH_ITERS with it + GRAPH_ITERS <= max_iter: the nodes are those of
// iterations 0 .. GRAPH_ITERS - 1 (the same replacement points; the last
// iteration allowed, if among them, is a replacement point anyway).
void launch_graph(System &s, CGWork *w, bool mixed, int max_iter, hipStream_t st) {
  const std::vector<const void *> key = {
      (const void *)(intptr_t)mixed, (const void *)(intptr_t)max_iter, s.d_slice_ptr.p,
      s.d_sell_col.p, s.d_sell_val.p, s.d_diag.p, (const void *)(intptr_t)s.n};
  if (!w->graph || w->gkey != key) {
    if (w->graph) {
      HIPCHECK(hipGraphExecDestroy(w->graph));
      w->graph = nullptr;
    }
    if (!w->cap) HIPCHECK(hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking));
    hipGraph_t g;
    HIPCHECK(hipStreamBeginCapture(w->cap, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < GRAPH_ITERS; k++) enqueue_iteration(s, w, mixed, max_iter, k, w->cap);
    HIPCHECK(hipStreamEndCapture(w->cap, &g));
    hipGraphExec_t ex;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
    w->graph = ex;
    w->gkey = key;
  }
  HIPCHECK(hipGraphLaunch(w->graph, st));
}

}  // namespace

void cg_solve(System &s, const double *b, double *x_out, const admfem_solve_opts &o,
              admfem_solve_stats &stats, hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  CGWork *w = ensure_work(s);
  const int nb = w->nblocks, np = s.npad;
  const bool mixed = o.precision == ADMFEM_PREC_MIXED;
  REQUIRE(o.precision == ADMFEM_PREC_FP64 || mixed, ADMFEM_ERR_ARG, "bad precision");
  REQUIRE(o.rtol > 0 && o.max_iter > 0, ADMFEM_ERR_ARG, "bad rtol/max_iter");
  CGScalars init{};
  init.rtol2 = o.rtol * o.rtol;
  const double delta = o.replace_delta > 0 ? o.replace_delta : REPLACE_DELTA;
  init.delta2 = delta * delta;
  init.mixed = mixed ? 1 : 0;
  init.b = b;
  HIPCHECK(hipMemcpyAsync(w->sc.p, &init, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  CGScalars *S = w->sc.p;
  double *part = w->part.p;
  const int *sp = s.d_slice_ptr.p, *scol = s.d_sell_col.p;
  if (mixed) {  // scaled form: S A S from the FP64 matrix, then the FP32 iteration
    batch_scale(s, s.d_sell_val.p, s.d_diag.p, w->dsq.p, w->sv32s.p, st);
    k_minit<<<nb, NT, 0, st>>>(s.n, np, s.d_row_bd.p, b, w->dsq.p, w->x.p, w->ys32.p,
                               w->r32[0].p, w->p32[0].p, w->q32[0].p, part);
  } else {
    k_init<<<nb, NT, 0, st>>>(s.n, np, s.d_row_bd.p, b, s.d_diag.p, w->x.p, w->z64[0].p,
                              w->w64[0].p, w->p64[0].p, part);
  }
  s_init<<<1, RT, 0, st>>>(nb, part, S);
  HIPCHECK(hipGetLastError());
  int chunk = o.check_every > 0 ? o.check_every : 2 * GRAPH_ITERS;
  int it = 0;  // iterations enqueued
  double rr_prev = -1;
  int it_prev = 0;
  double rr_true = 0.0;  // ||b - A x||^2 of the returned x
  for (int restarts = 0;; restarts++) {
    for (;;) {
      const int stop = std::min(it + chunk, o.max_iter);
      while (it < stop) {
        if (it % GRAPH_ITERS == 0 && it + GRAPH_ITERS <= stop) {
          launch_graph(s, w, mixed, o.max_iter, st);
          it += GRAPH_ITERS;
        } else {
          enqueue_iteration(s, w, mixed, o.max_iter, it, st);
          it++;
        }
      }
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpyAsync(w->h_sc, S, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      const CGScalars &h = *w->h_sc;
      if (h.status != ST_RUN || it >= o.max_iter) break;
      // next chunk: the iterations left, estimated from the observed rate of
      // the monitored residual (mixed: the scaled one, against the tolerance
      // times the scaled-to-true ratio of the last replacement)
      const double mon = mixed ? h.rz : h.rr;
      const double target = mixed ? h.tol2 * h.rz_ref / h.rr_ref : h.tol2;
      int next = 16;
      if (rr_prev > 0 && mon > 0 && h.it > it_prev && mon < rr_prev) {
        double rate = std::log(mon / rr_prev) / (h.it - it_prev);
        double left = std::log(target / mon) / rate;
        if (std::isfinite(left)) next = (int)std::min(1024.0, std::max(8.0, 0.9 * left));
      }
      rr_prev = mon;
      it_prev = h.it;
      chunk = o.check_every > 0 ? o.check_every
                                : std::max(GRAPH_ITERS, (next + GRAPH_ITERS / 2) / GRAPH_ITERS *
                                                            GRAPH_ITERS);
    }
    const bool converged = w->h_sc->status == ST_CONV;
    if (mixed && !converged)  // stopped without a final replacement
      k_mfold<<<nb, NT, 0, st>>>(np, w->x.p, w->ys32.p, w->dsq.p, S, 1);
    k_true_residual<<<nb, NT, 0, st>>>(s.n, np, sp, scol, s.d_sell_val.p, s.d_diag.p, b,
                                       w->x.p, part);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(w->h_part.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                            st));
    HIPCHECK(hipStreamSynchronize(st));
    rr_true = 0.0;
    for (int k = 0; k < nb; k++) rr_true += w->h_part[k];
    // FP64: the recursive residual decides only when the true one agrees;
    // otherwise it is replaced by the true one and the iteration goes on
    if (mixed || !converged || rr_true <= w->h_sc->tol2 || it >= o.max_iter || restarts >= 8)
      break;
    k_restart<<<nb, NT, 0, st>>>(s.n, np, sp, scol, s.d_sell_val.p, s.d_diag.p,
                                 w->x.p, w->z64[0].p, w->z64[1].p, part, S);
    s_restart<<<1, RT, 0, st>>>(nb, part, S, o.max_iter);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(w->h_sc, S, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (w->h_sc->status != ST_RUN) {  // x unchanged: its true residual is the restart's
      rr_true = w->h_sc->rr;
      break;
    }
    rr_prev = -1;
    it = w->h_sc->it;  // the steps run so far (later ones were skipped)
  }
  if (x_out != w->x.p)
    HIPCHECK(hipMemcpyAsync(x_out, w->x.p, sizeof(double) * s.n, hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipMemcpyAsync(w->h_sc, S, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  const CGScalars &h = *w->h_sc;
  stats.iterations = h.it;
  stats.replacements = h.nrepl;
  stats.rel_residual = h.bb > 0 ? std::sqrt(rr_true / h.bb) : 0.0;
  switch (h.status) {
    case ST_CONV: stats.status = ADMFEM_OK; break;
    case ST_NAN: stats.status = ADMFEM_ERR_NAN; break;
    case ST_BREAK: stats.status = ADMFEM_ERR_BREAKDOWN; break;
    default: stats.status = ADMFEM_ERR_NOT_CONVERGED; break;
  }
  stats.seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace admfem
