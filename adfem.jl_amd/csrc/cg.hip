// Jacobi-preconditioned conjugate gradients on the SELL-64 form of the
// Dirichlet-eliminated system -- the GPU replacement of the reference's
// direct sparse solve K\F (ADCME SparseSolver, Eigen SparseLU [ext]).
//
// Two precisions share the kernels:
//  * FP64: matrix, vectors and reductions in FP64.
//  * mixed: matrix and the inner vectors (r, z, p, q) stored in FP32 (FP64
//    arithmetic), the solution correction accumulated in FP32; whenever the
//    FP32 residual has fallen by `replace_delta` (or looks converged) the
//    correction is folded into the FP64 solution and the residual is
//    replaced by the true FP64 residual b - A x, computed with the FP64
//    matrix ("reliable updates", van der Vorst & Ye 2000; Clark et al.
//    2010).  Convergence is declared only on a true FP64 residual.
//
// Per iteration:
//   spmv:    p = z + beta p_old (own row, and on the fly for each gathered
//            neighbour), q = A p, per-block partials of p.q
//   scalar:  one block sums the partials in fixed order -> alpha
//   update:  x += alpha p, r -= alpha q, z = D^-1 r, partials of r.z, r.r
//   scalar:  -> beta, convergence / replacement decision
// (mixed, every eighth iteration, when flagged: fold x, replace r, scalar).
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

namespace admfem {

enum { ST_RUN = 0, ST_CONV = 1, ST_MAXIT = 2, ST_NAN = 3, ST_BREAK = 4 };

struct CGScalars {
  double rtol2, delta2;  // set by the host
  double bb, tol2;       // ||b||^2, rtol^2 ||b||^2
  double rz, rz_prev;    // r.z of the current / previous residual
  double alpha, beta;
  double rr;             // ||r||^2 of the current residual (recursive)
  double rr_ref;         // mixed: true ||b - A x||^2 at the last replacement
  double rz_ref;         // mixed: scaled r.r at the last replacement
  double rz_min;         // mixed: smallest scaled r.r since then
  int mixed;
  int status;            // ST_*
  int it;                // completed iterations
  int replace;           // mixed: replacement pending
  int nrepl;
  int stall;             // mixed: iterations without a new rz_min
};


constexpr int NT = 256;   // threads per wide block = 4 slices of 64 rows
constexpr int RT = 1024;  // threads of the single-block scalar kernels

struct CGWork {
  int npad = 0, nblocks = 0;
  DevBuf<double> x, r64, z64, pa64, pb64, q64;
  DevBuf<float> xs32, r32, pa32, pb32, q32;
  DevBuf<double> dsq;    // mixed: 1 / sqrt(a_ii)
  DevBuf<float> sv32s;   // mixed: FP32 scaled off-diagonals a_ij dsq_i dsq_j
  DevBuf<double> part;  // 3 x nblocks block partials
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

// One SELL row, FP64 arithmetic: sum_k v_k * f(col_k).  The first 8 entries
// are unrolled so every column/value load issues before the dependent
// gathers; wider rows (high-valence nodes) finish in a loop.
template <typename VT, class F>
__device__ __forceinline__ double sell_dot(int i, const int *__restrict__ sp,
                                           const int *__restrict__ sc,
                                           const VT *__restrict__ sv, F f) {
  const int sl = i >> 6;
  const int s0 = sp[sl];
  const int off = s0 + (i & 63);
  const int wfull = (sp[sl + 1] - s0) >> 6;
  const int w = min(wfull, 8);
  int c[8];
  VT v[8];
  // FP64 values: the matrix is streamed once per product, so its loads are
  // non-temporal (measured on the FP64 SpMV: 99 -> 82 us at 2048^2; FP32
  // values lose with them, 49 -> 56 us, and keep ordinary loads)
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (k < w) {
      if constexpr (sizeof(VT) == 8) {
        c[k] = __builtin_nontemporal_load(sc + off + 64 * k);
        v[k] = __builtin_nontemporal_load(sv + off + 64 * k);
      } else {
        c[k] = sc[off + 64 * k];
        v[k] = sv[off + 64 * k];
      }
    }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (k < w) s += (double)v[k] * f(c[k]);
  for (int k = 8; k < wfull; k++) s += (double)sv[off + 64 * k] * f(sc[off + 64 * k]);
  return s;
}

// x0 = b on identity (boundary) rows, 0 elsewhere; r0 = b - A x0 (0 on those
// rows, b on the rest: interior rows of the eliminated system have no
// boundary columns); z0 = D^-1 r0; p_old = 0 (with beta = 0 the first spmv
// forms p = z0); partials of r.z, r.r, b.b.
template <typename VT>
__global__ void __launch_bounds__(NT)
    k_init(int n, int npad, const int *__restrict__ row_bd, const double *__restrict__ b,
           const VT *__restrict__ dinv, double *__restrict__ x, float *__restrict__ xs,
           VT *__restrict__ r, VT *__restrict__ z, VT *__restrict__ po, double *part) {
  const int i = my_row();
  double v[3] = {0.0, 0.0, 0.0};
  if (i < npad) {
    const double bi = i < n ? b[i] : 0.0;
    const bool ident = i >= n || row_bd[i] >= 0;
    const double ri = ident ? 0.0 : bi;
    x[i] = (ident && i < n) ? bi : 0.0;
    if (xs) xs[i] = 0.0f;
    const VT rv = (VT)ri;
    const VT zv = (VT)((double)rv * (double)dinv[i]);
    r[i] = rv;
    z[i] = zv;
    po[i] = (VT)0;
    v[0] = (double)rv * (double)zv;
    v[1] = ri * ri;
    v[2] = bi * bi;
  }
  put_partials<3>(v, part);
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
  S->beta = 0.0;
  S->it = 0;
  S->nrepl = 0;
  S->replace = 0;
  S->status = (!(t[1] == t[1]) || !isfinite(t[1])) ? ST_NAN
              : (t[1] <= S->tol2)                  ? ST_CONV
                                                   : ST_RUN;
}

// p = z + beta p_old; q = A p; partials of p.q
template <typename VT>
__global__ void __launch_bounds__(NT)
    k_spmv(int npad, const int *__restrict__ sp, const int *__restrict__ sc,
           const VT *__restrict__ sv, const VT *__restrict__ diag, const VT *__restrict__ z,
           const VT *__restrict__ po, VT *__restrict__ pn, VT *__restrict__ q, double *part,
           const CGScalars *S) {
  if (S->status != ST_RUN) return;
  const double beta = S->beta;
  const int i = my_row();
  double v[1] = {0.0};
  if (i < npad) {
    const double p_i = (double)z[i] + beta * (double)po[i];
    const double s = (double)diag[i] * p_i +
                     sell_dot<VT>(i, sp, sc, sv, [&](int c) {
                       return (double)z[c] + beta * (double)po[c];
                     });
    const VT pv = (VT)p_i, qv = (VT)s;
    pn[i] = pv;
    q[i] = qv;
    v[0] = (double)pv * (double)qv;
  }
  put_partials<1>(v, part);
}

__global__ void __launch_bounds__(RT) s_alpha(int G, const double *part, CGScalars *S) {
  if (S->status != ST_RUN) return;
  double t[1];
  sum_partials<1>(part, G, t);
  if (threadIdx.x != 0) return;
  if (!(t[0] > 0.0)) {  // not SPD, or NaN
    S->status = (t[0] != t[0]) ? ST_NAN : ST_BREAK;
    return;
  }
  S->alpha = S->rz / t[0];
}

// x += alpha p (FP64 x, or the FP32 correction xs in the mixed solve);
// r -= alpha q; z = D^-1 r; partials of r.z, r.r
template <typename VT, typename XT>
__global__ void __launch_bounds__(NT)
    k_update(int npad, const VT *__restrict__ p, const VT *__restrict__ q,
             XT *__restrict__ x, VT *__restrict__ r, VT *__restrict__ z,
             const VT *__restrict__ dinv, double *part, const CGScalars *S) {
  if (S->status != ST_RUN) return;
  const double alpha = S->alpha;
  const int i = my_row();
  double v[2] = {0.0, 0.0};
  if (i < npad) {
    const double pi = (double)p[i];
    x[i] = (XT)((double)x[i] + alpha * pi);
    const VT rv = (VT)((double)r[i] - alpha * (double)q[i]);
    const VT zv = (VT)((double)rv * (double)dinv[i]);
    r[i] = rv;
    z[i] = zv;
    v[0] = (double)rv * (double)zv;
    v[1] = (double)rv * (double)rv;
  }
  put_partials<2>(v, part);
}

__global__ void __launch_bounds__(RT) s_beta(int G, const double *part, CGScalars *S,
                                              int maxit) {
  if (S->status != ST_RUN) return;
  double t[2];
  sum_partials<2>(part, G, t);
  if (threadIdx.x != 0) return;
  S->rz_prev = S->rz;
  S->rz = t[0];
  S->beta = t[0] / S->rz_prev;
  S->it += 1;
  if (!(t[0] == t[0]) || !isfinite(t[0])) {
    S->status = ST_NAN;
  } else if (S->mixed) {
    // scaled form: r.r of the scaled residual is the monitor; a replacement
    // when it fell by delta, stalled, or looks converged (below the
    // tolerance times the scaled-to-true ratio of the last replacement)
    if (t[0] < 0.98 * S->rz_min) {
      S->rz_min = t[0];
      S->stall = 0;
    } else {
      S->stall += 1;
    }
    if (t[0] * S->rr_ref <= S->tol2 * S->rz_ref || t[0] <= S->delta2 * S->rz_ref ||
        S->stall >= STALL_ITERS)
      S->replace = 1;
    if (S->it >= maxit && !S->replace) S->status = ST_MAXIT;
  } else {
    S->rr = t[1];
    if (t[1] <= S->tol2) S->status = ST_CONV;
    else if (S->it >= maxit) S->status = ST_MAXIT;
  }
}

// ---- mixed precision, scaled form: plain CG on S A S (S = diag^-1/2,
// unit diagonal) with FP32 storage, the solution x = x64 + S y

// x0 = b on identity rows; r = S r0 (FP32); y = 0; p_old = 0; partials of
// the scaled r.r, the true ||r0||^2 and b.b
__global__ void __launch_bounds__(NT)
    k_minit(int n, int npad, const int *__restrict__ row_bd, const double *__restrict__ b,
            const double *__restrict__ dsq, double *__restrict__ x, float *__restrict__ ys,
            float *__restrict__ r, float *__restrict__ po, double *part) {
  const int i = my_row();
  double v[3] = {0.0, 0.0, 0.0};
  if (i < npad) {
    const double bi = i < n ? b[i] : 0.0;
    const bool ident = i >= n || row_bd[i] >= 0;
    const double ri = ident ? 0.0 : bi;
    x[i] = (ident && i < n) ? bi : 0.0;
    ys[i] = 0.0f;
    const float rv = (float)(ri * dsq[i]);
    r[i] = rv;
    po[i] = 0.0f;
    v[0] = (double)rv * (double)rv;
    v[1] = ri * ri;
    v[2] = bi * bi;
  }
  put_partials<3>(v, part);
}

// p = r + beta p_old; q = (S A S) p (unit diagonal); partials of p.q
__global__ void __launch_bounds__(NT)
    k_mspmv(int npad, const int *__restrict__ sp, const int *__restrict__ sc,
            const float *__restrict__ sv, const float *__restrict__ r,
            const float *__restrict__ po, float *__restrict__ pn, float *__restrict__ q,
            double *part, const CGScalars *S) {
  if (S->status != ST_RUN) return;
  const double beta = S->beta;
  const int i = my_row();
  double v[1] = {0.0};
  if (i < npad) {
    const double p_i = (double)r[i] + beta * (double)po[i];
    const double a = p_i + sell_dot<float>(i, sp, sc, sv, [&](int c) {
                       return (double)r[c] + beta * (double)po[c];
                     });
    const float pv = (float)p_i, qv = (float)a;
    pn[i] = pv;
    q[i] = qv;
    v[0] = (double)pv * (double)qv;
  }
  put_partials<1>(v, part);
}

// y += alpha p; r -= alpha q; partials of the scaled r.r
__global__ void __launch_bounds__(NT)
    k_mupdate(int npad, const float *__restrict__ p, const float *__restrict__ q,
              float *__restrict__ ys, float *__restrict__ r, double *part, const CGScalars *S) {
  if (S->status != ST_RUN) return;
  const double alpha = S->alpha;
  const int i = my_row();
  double v[2] = {0.0, 0.0};
  if (i < npad) {
    ys[i] = (float)((double)ys[i] + alpha * (double)p[i]);
    const float rv = (float)((double)r[i] - alpha * (double)q[i]);
    r[i] = rv;
    v[0] = (double)rv * (double)rv;
  }
  put_partials<2>(v, part);
}

// when a replacement is pending (or forced at the end): x += S y, y = 0
__global__ void __launch_bounds__(NT)
    k_mfold(int npad, double *__restrict__ x, float *__restrict__ ys,
            const double *__restrict__ dsq, const CGScalars *S, int force) {
  if (!force && (S->status != ST_RUN || !S->replace)) return;
  const int i = my_row();
  if (i < npad) {
    x[i] += dsq[i] * (double)ys[i];
    ys[i] = 0.0f;
  }
}

// when a replacement is pending: the true residual b - A x in FP64 (FP64
// matrix), handed to the FP32 iteration as r = S (b - A x); partials of the
// scaled r.r and the true ||b - A x||^2
__global__ void __launch_bounds__(NT)
    k_mreplace(int n, int npad, const int *__restrict__ sp, const int *__restrict__ sc,
               const double *__restrict__ sv, const double *__restrict__ diag,
               const double *__restrict__ b, const double *__restrict__ x,
               const double *__restrict__ dsq, float *__restrict__ r, double *part,
               const CGScalars *S) {
  if (S->status != ST_RUN || !S->replace) return;
  const int i = my_row();
  double v[2] = {0.0, 0.0};
  if (i < npad) {
    const double ax = diag[i] * x[i] + sell_dot<double>(i, sp, sc, sv, [&](int c) { return x[c]; });
    const double ri = (i < n ? b[i] : 0.0) - ax;
    const float rf = (float)(ri * dsq[i]);
    r[i] = rf;
    v[0] = (double)rf * (double)rf;
    v[1] = ri * ri;
  }
  put_partials<2>(v, part);
}

__global__ void __launch_bounds__(RT) s_replace(int G, const double *part, CGScalars *S,
                                                 int maxit) {
  if (S->status != ST_RUN || !S->replace) return;
  double t[2];
  sum_partials<2>(part, G, t);
  if (threadIdx.x != 0) return;
  S->rz = t[0];
  S->beta = t[0] / S->rz_prev;
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
    w->r64.alloc(s.npad);
    w->z64.alloc(s.npad);
    w->pa64.alloc(s.npad);
    w->pb64.alloc(s.npad);
    w->q64.alloc(s.npad);
    w->xs32.alloc(s.npad);
    w->r32.alloc(s.npad);
    w->dsq.alloc(s.npad);
    w->sv32s.alloc(std::max(1, s.slice_ptr.back()));
    w->pa32.alloc(s.npad);
    w->pb32.alloc(s.npad);
    w->q32.alloc(s.npad);
    w->part.alloc((size_t)3 * nb);
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

// Instrumentation: the solver's SpMV kernel launched `repeat` times on the
// system's work vectors (z = 1, p_old = 0, status running): it runs its full
// body every launch and changes none of its inputs.
void spmv_repeat(System &s, int precision, int repeat, hipStream_t st) {
  CGWork *w = ensure_work(s);
  const int nb = w->nblocks;
  CGScalars init{};
  HIPCHECK(hipMemcpyAsync(w->sc.p, &init, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));  // init lives on this stack frame
  const unsigned fb = (unsigned)((s.npad + 255) / 256);
  if (precision == ADMFEM_PREC_MIXED) {  // the scaled-form kernel of the mixed solve
    batch_scale(s, s.d_sell_val.p, s.d_diag.p, w->dsq.p, w->sv32s.p, st);
    k_fill<float><<<fb, 256, 0, st>>>(w->r32.p, s.npad, 1.0f);
    k_fill<float><<<fb, 256, 0, st>>>(w->pa32.p, s.npad, 0.0f);
    for (int k = 0; k < repeat; k++)
      k_mspmv<<<nb, NT, 0, st>>>(s.npad, s.d_slice_ptr.p, s.d_sell_col.p, w->sv32s.p,
                                 w->r32.p, w->pa32.p, w->pb32.p, w->q32.p, w->part.p, w->sc.p);
  } else {
    k_fill<double><<<fb, 256, 0, st>>>(w->z64.p, s.npad, 1.0);
    k_fill<double><<<fb, 256, 0, st>>>(w->pa64.p, s.npad, 0.0);
    for (int k = 0; k < repeat; k++)
      k_spmv<double><<<nb, NT, 0, st>>>(s.npad, s.d_slice_ptr.p, s.d_sell_col.p,
                                        s.d_sell_val.p, s.d_diag.p, w->z64.p, w->pa64.p,
                                        w->pb64.p, w->q64.p, w->part.p, w->sc.p);
  }
  HIPCHECK(hipGetLastError());
}

namespace {

constexpr int GRAPH_ITERS = 8;  // a multiple of the replacement interval (8)

// One iteration with index `it` (its parity picks the direction buffers;
// the mixed form checks for a replacement after every eighth iteration and
// after the last one allowed).
void enqueue_iteration(System &s, CGWork *w, bool mixed, const double *b, int max_iter, int it,
                       hipStream_t st) {
  const int nb = w->nblocks, np = s.npad;
  const int *sp = s.d_slice_ptr.p, *scol = s.d_sell_col.p;
  CGScalars *S = w->sc.p;
  double *part = w->part.p;
  const bool odd = it & 1;
  if (mixed) {
    float *po = odd ? w->pb32.p : w->pa32.p, *pn = odd ? w->pa32.p : w->pb32.p;
    k_mspmv<<<nb, NT, 0, st>>>(np, sp, scol, w->sv32s.p, w->r32.p, po, pn, w->q32.p, part, S);
    s_alpha<<<1, RT, 0, st>>>(nb, part, S);
    k_mupdate<<<nb, NT, 0, st>>>(np, pn, w->q32.p, w->xs32.p, w->r32.p, part, S);
    s_beta<<<1, RT, 0, st>>>(nb, part, S, max_iter);
    if ((it & 7) == 7 || it + 1 == max_iter) {  // replacement check every 8
      k_mfold<<<nb, NT, 0, st>>>(np, w->x.p, w->xs32.p, w->dsq.p, S, 0);
      k_mreplace<<<nb, NT, 0, st>>>(s.n, np, sp, scol, s.d_sell_val.p, s.d_diag.p, b, w->x.p,
                                    w->dsq.p, w->r32.p, part, S);
      s_replace<<<1, RT, 0, st>>>(nb, part, S, max_iter);
    }
  } else {
    double *po = odd ? w->pb64.p : w->pa64.p, *pn = odd ? w->pa64.p : w->pb64.p;
    k_spmv<double><<<nb, NT, 0, st>>>(np, sp, scol, s.d_sell_val.p, s.d_diag.p, w->z64.p, po,
                                      pn, w->q64.p, part, S);
    s_alpha<<<1, RT, 0, st>>>(nb, part, S);
    k_update<double, double><<<nb, NT, 0, st>>>(np, pn, w->q64.p, w->x.p, w->r64.p, w->z64.p,
                                                s.d_dinv.p, part, S);
    s_beta<<<1, RT, 0, st>>>(nb, part, S, max_iter);
  }
}

// Iterations it .. it + GRAPH_ITERS - 1 for an `it` divisible by
// GRAPH_ITERS with it + GRAPH_ITERS <= max_iter: the nodes are those of
// iterations 0 .. GRAPH_ITERS - 1 (same parities, and the last one allowed,
// if among them, is a replacement point anyway).
void launch_graph(System &s, CGWork *w, bool mixed, const double *b, int max_iter,
                  hipStream_t st) {
  const std::vector<const void *> key = {
      (const void *)(intptr_t)mixed, (const void *)(intptr_t)max_iter,
      mixed ? (const void *)b : nullptr, s.d_slice_ptr.p, s.d_sell_col.p, s.d_sell_val.p,
      s.d_diag.p, s.d_dinv.p, (const void *)(intptr_t)s.n};
  if (!w->graph || w->gkey != key) {
    if (w->graph) {
      HIPCHECK(hipGraphExecDestroy(w->graph));
      w->graph = nullptr;
    }
    if (!w->cap) HIPCHECK(hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking));
    hipGraph_t g;
    HIPCHECK(hipStreamBeginCapture(w->cap, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < GRAPH_ITERS; k++) enqueue_iteration(s, w, mixed, b, max_iter, k, w->cap);
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
  HIPCHECK(hipMemcpyAsync(w->sc.p, &init, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  CGScalars *S = w->sc.p;
  double *part = w->part.p;
  const int *sp = s.d_slice_ptr.p, *scol = s.d_sell_col.p;
  if (mixed) {  // scaled form: S A S from the FP64 matrix, then the FP32 iteration
    batch_scale(s, s.d_sell_val.p, s.d_diag.p, w->dsq.p, w->sv32s.p, st);
    k_minit<<<nb, NT, 0, st>>>(s.n, np, s.d_row_bd.p, b, w->dsq.p, w->x.p, w->xs32.p,
                               w->r32.p, w->pa32.p, part);
  } else {
    k_init<double><<<nb, NT, 0, st>>>(s.n, np, s.d_row_bd.p, b, s.d_dinv.p, w->x.p, nullptr,
                                      w->r64.p, w->z64.p, w->pa64.p, part);
  }
  s_init<<<1, RT, 0, st>>>(nb, part, S);
  HIPCHECK(hipGetLastError());
  int chunk = o.check_every > 0 ? o.check_every : 2 * GRAPH_ITERS;
  int it = 0;  // iterations enqueued
  double rr_prev = -1;
  int it_prev = 0;
  for (;;) {
    const int stop = std::min(it + chunk, o.max_iter);
    while (it < stop) {
      if (it % GRAPH_ITERS == 0 && it + GRAPH_ITERS <= stop) {
        launch_graph(s, w, mixed, b, o.max_iter, st);
        it += GRAPH_ITERS;
      } else {
        enqueue_iteration(s, w, mixed, b, o.max_iter, it, st);
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
    k_mfold<<<nb, NT, 0, st>>>(np, w->x.p, w->xs32.p, w->dsq.p, S, 1);
  k_true_residual<<<nb, NT, 0, st>>>(s.n, np, sp, scol, s.d_sell_val.p, s.d_diag.p, b, w->x.p,
                                     part);
  HIPCHECK(hipGetLastError());
  if (x_out != w->x.p)
    HIPCHECK(hipMemcpyAsync(x_out, w->x.p, sizeof(double) * s.n, hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipMemcpyAsync(w->h_part.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                          st));
  HIPCHECK(hipMemcpyAsync(w->h_sc, S, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  const CGScalars &h = *w->h_sc;
  double rr_true = 0.0;
  for (int k = 0; k < nb; k++) rr_true += w->h_part[k];
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
