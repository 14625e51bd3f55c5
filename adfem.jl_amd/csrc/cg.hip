// Jacobi-preconditioned conjugate gradients on the SELL-64 form of the
// Dirichlet-eliminated system -- the GPU replacement of the reference's
// direct sparse solve K\F (ADCME SparseSolver, Eigen SparseLU [ext]).
//
// Two precisions share the kernels:
//  * FP64: matrix, vectors and reductions in FP64.
//  * mixed: matrix and the inner vectors (r, p, q) in FP32; the solution is
//    accumulated in FP64 every iteration, and whenever the FP32 residual has
//    fallen by `replace_delta` (or looks converged) it is replaced by the
//    true FP64 residual b - A x computed with the FP64 matrix ("reliable
//    updates", van der Vorst & Ye 2000; Clark et al. 2010).  Convergence is
//    only declared on a true FP64 residual.
//
// Per iteration: spmv (q = A p, p.q), update (x, r, z = D^-1 r, r.z, r.r),
// [replace], pupdate (p = z + beta p).  Global reductions are deterministic:
// per-block partials, then the last block to finish (integer ticket) sums
// them in block order and updates the scalar state.  Every kernel first reads
// the solver status and returns at once when the solve has finished, so the
// host can enqueue iterations in chunks and poll the status between chunks.
#include <chrono>
#include <cmath>

#include "geom.h"

namespace admfem {

enum { ST_RUN = 0, ST_CONV = 1, ST_MAXIT = 2, ST_NAN = 3, ST_BREAK = 4 };

struct CGScalars {
  double rz, rz_prev, pq, alpha, beta, rr, bb, tol2, rr_ref, delta2, rr_true;
  int it, maxit, status, replace, nrepl, mixed;
  unsigned tick[4];
};

struct CGWork {
  int npad = 0, nblocks = 0;
  DevBuf<double> x, r64, p64, q64, g64, lam64;
  DevBuf<float> r32, p32, q32;
  DevBuf<double> part;
  DevBuf<CGScalars> sc;
  CGScalars *h_sc = nullptr;  // pinned
  ~CGWork() {
    if (h_sc) (void)hipHostFree(h_sc);
  }
};

void cg_work_free(CGWork *w) { delete w; }

namespace {

constexpr int NT = 256;  // threads per block = 4 slices of 64 rows

// Blocks are dispatched round-robin over the 8 XCDs; map them so that each
// XCD works on one contiguous eighth of the rows (its L2 then holds that
// eighth's x-gather window instead of all of x).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Sum NV values over the grid: per-block partial into part[i*G + block], the
// last block to arrive sums all partials in block order into tot.  Returns
// true in the last block only.
template <int NV>
__device__ bool grid_reduce(const double (&v)[NV], double *part, unsigned *tick,
                            double (&tot)[NV]) {
  __shared__ double sh[NV][NT / 64];
  __shared__ int last;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = wave_sum(v[i]);
    if (lane == 0) sh[i][wv] = s;
  }
  __syncthreads();
  const int G = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < NT / 64; k++) s += sh[i][k];
      part[i * G + blockIdx.x] = s;
    }
    __threadfence();
    unsigned t = atomicAdd(tick, 1u);
    last = (t == (unsigned)G - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  __shared__ double sh2[NT / 64];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double a = 0.0;
    for (int k = threadIdx.x; k < G; k += NT) a += __builtin_nontemporal_load(&part[i * G + k]);
    tot[i] = block_sum<NT>(a, sh2);
  }
  if (threadIdx.x == 0) *tick = 0u;
  return true;
}

// rows [lo, hi) of this block: contiguous chunk of whole slices
struct RowRange {
  int lo, hi;
};
__device__ __forceinline__ RowRange block_rows(int nslices) {
  const int nb = gridDim.x;
  const int lb = xcd_block(blockIdx.x, nb);
  const int per = (nslices + nb - 1) / nb;
  int s0 = min(lb * per, nslices), s1 = min(s0 + per, nslices);
  return {s0 * 64, s1 * 64};
}

template <typename VT>
__device__ __forceinline__ VT sell_row(int r, const int *__restrict__ sp,
                                       const int *__restrict__ sc,
                                       const VT *__restrict__ sv,
                                       const VT *__restrict__ diag,
                                       const VT *__restrict__ p) {
  const int sl = r >> 6, lane = r & 63;
  const int off = sp[sl] + lane;
  const int w = (sp[sl + 1] - sp[sl]) >> 6;
  VT s = diag[r] * p[r];
  for (int k = 0; k < w; k++) s += sv[off + 64 * k] * p[sc[off + 64 * k]];
  return s;
}

// x0 = b on identity (boundary) rows, 0 elsewhere; r0 = b - A x0, which is 0
// on those rows and b on the rest (interior rows of the eliminated system have
// no boundary columns); z0 = D^-1 r0, p0 = z0.
template <typename VT>
__global__ void __launch_bounds__(NT)
    k_init(int nslices, int n, const int *__restrict__ row_bd,
           const double *__restrict__ b, const VT *__restrict__ dinv,
           double *__restrict__ x, VT *__restrict__ r, VT *__restrict__ p,
           double *part, CGScalars *S) {
  RowRange R = block_rows(nslices);
  double rz = 0, rr = 0, bb = 0;
  for (int i = R.lo + threadIdx.x; i < R.hi; i += NT) {
    double bi = i < n ? b[i] : 0.0;
    bool ident = i >= n || row_bd[i] >= 0;
    double ri = ident ? 0.0 : bi;
    x[i] = ident && i < n ? bi : 0.0;
    VT rv = (VT)ri;
    VT z = rv * dinv[i];
    r[i] = rv;
    p[i] = z;
    rz += (double)rv * (double)z;
    rr += ri * ri;
    bb += bi * bi;
  }
  double v[3] = {rz, rr, bb}, t[3];
  if (grid_reduce<3>(v, part, &S->tick[0], t) && threadIdx.x == 0) {
    S->rz = t[0];
    S->rr = t[1];
    S->rr_ref = t[1];
    S->rr_true = t[1];
    S->bb = t[2];
    S->tol2 *= t[2];  // host stored rtol^2
    S->it = 0;
    S->nrepl = 0;
    S->replace = 0;
    S->status = (t[1] <= S->tol2) ? ST_CONV : ST_RUN;
  }
}

template <typename VT>
__global__ void __launch_bounds__(NT)
    k_spmv(int nslices, const int *__restrict__ sp, const int *__restrict__ sc,
           const VT *__restrict__ sv, const VT *__restrict__ diag,
           const VT *__restrict__ p, VT *__restrict__ q, double *part,
           CGScalars *S) {
  if (S->status != ST_RUN) return;
  RowRange R = block_rows(nslices);
  double pq = 0;
  for (int i = R.lo + threadIdx.x; i < R.hi; i += NT) {
    VT s = sell_row<VT>(i, sp, sc, sv, diag, p);
    q[i] = s;
    pq += (double)p[i] * (double)s;
  }
  double v[1] = {pq}, t[1];
  if (grid_reduce<1>(v, part, &S->tick[1], t) && threadIdx.x == 0) {
    S->pq = t[0];
    if (!(t[0] > 0.0)) {
      S->status = (t[0] != t[0]) ? ST_NAN : ST_BREAK;
    } else {
      S->alpha = S->rz / t[0];
    }
  }
}

template <typename VT>
__global__ void __launch_bounds__(NT)
    k_update(int nslices, const VT *__restrict__ p, const VT *__restrict__ q,
             double *__restrict__ x, VT *__restrict__ r,
             const VT *__restrict__ dinv, double *part, CGScalars *S) {
  if (S->status != ST_RUN) return;
  RowRange R = block_rows(nslices);
  const double alpha = S->alpha;
  const VT a = (VT)alpha;
  double rz = 0, rr = 0;
  for (int i = R.lo + threadIdx.x; i < R.hi; i += NT) {
    VT pi = p[i];
    x[i] += alpha * (double)pi;
    VT ri = r[i] - a * q[i];
    r[i] = ri;
    VT z = ri * dinv[i];
    rz += (double)ri * (double)z;
    rr += (double)ri * (double)ri;
  }
  double v[2] = {rz, rr}, t[2];
  if (grid_reduce<2>(v, part, &S->tick[2], t) && threadIdx.x == 0) {
    S->rz_prev = S->rz;
    S->rz = t[0];
    S->beta = t[0] / S->rz_prev;
    S->rr = t[1];
    S->it += 1;
    if (!(t[1] == t[1]) || !isfinite(t[1])) {
      S->status = ST_NAN;
    } else if (S->mixed) {
      if (t[1] <= S->tol2 || t[1] <= S->delta2 * S->rr_ref) S->replace = 1;
      if (S->it >= S->maxit && !S->replace) S->status = ST_MAXIT;
    } else {
      if (t[1] <= S->tol2) S->status = ST_CONV;
      else if (S->it >= S->maxit) S->status = ST_MAXIT;
    }
  }
}

// Mixed precision: r = b - A x in FP64 (FP64 matrix), handed to the FP32
// inner iteration.
__global__ void __launch_bounds__(NT)
    k_replace(int nslices, int n, const int *__restrict__ sp,
              const int *__restrict__ sc, const double *__restrict__ sv,
              const double *__restrict__ diag, const double *__restrict__ b,
              const double *__restrict__ x, float *__restrict__ r32,
              const float *__restrict__ dinv32, double *part, CGScalars *S) {
  if (S->status != ST_RUN || !S->replace) return;
  RowRange R = block_rows(nslices);
  double rz = 0, rr = 0;
  for (int i = R.lo + threadIdx.x; i < R.hi; i += NT) {
    double ri = (i < n ? b[i] : 0.0) - sell_row<double>(i, sp, sc, sv, diag, x);
    float rf = (float)ri;
    r32[i] = rf;
    float z = rf * dinv32[i];
    rz += (double)rf * (double)z;
    rr += ri * ri;
  }
  double v[2] = {rz, rr}, t[2];
  if (grid_reduce<2>(v, part, &S->tick[3], t) && threadIdx.x == 0) {
    S->rz = t[0];
    S->beta = t[0] / S->rz_prev;
    S->rr = t[1];
    S->rr_ref = t[1];
    S->rr_true = t[1];
    S->replace = 0;
    S->nrepl += 1;
    if (!(t[1] == t[1]) || !isfinite(t[1])) S->status = ST_NAN;
    else if (t[1] <= S->tol2) S->status = ST_CONV;
    else if (S->it >= S->maxit) S->status = ST_MAXIT;
  }
}

template <typename VT>
__global__ void __launch_bounds__(NT)
    k_pupdate(int nslices, const VT *__restrict__ r, const VT *__restrict__ dinv,
              VT *__restrict__ p, const CGScalars *S) {
  if (S->status != ST_RUN) return;
  RowRange R = block_rows(nslices);
  const VT beta = (VT)S->beta;
  for (int i = R.lo + threadIdx.x; i < R.hi; i += NT) p[i] = r[i] * dinv[i] + beta * p[i];
}

// ||b - A x||^2 in FP64 (final check), into S->rr_true.
__global__ void __launch_bounds__(NT)
    k_true_residual(int nslices, int n, const int *__restrict__ sp,
                    const int *__restrict__ sc, const double *__restrict__ sv,
                    const double *__restrict__ diag, const double *__restrict__ b,
                    const double *__restrict__ x, double *part, CGScalars *S) {
  RowRange R = block_rows(nslices);
  double rr = 0;
  for (int i = R.lo + threadIdx.x; i < R.hi; i += NT) {
    double ri = (i < n ? b[i] : 0.0) - sell_row<double>(i, sp, sc, sv, diag, x);
    rr += ri * ri;
  }
  double v[1] = {rr}, t[1];
  if (grid_reduce<1>(v, part, &S->tick[0], t) && threadIdx.x == 0) S->rr_true = t[0];
}

}  // namespace

static CGWork *ensure_work(System &s) {
  if (s.work && s.work->npad == s.npad) return s.work;
  if (s.work) cg_work_free(s.work);
  CGWork *w = new CGWork();
  try {
    w->npad = s.npad;
    // block count: a multiple of 8 (XCD mapping), at most 2048, at least one
    // slice per block
    int nb = std::min(2048, (s.nslices + 3) / 4);
    nb = std::max(8, (nb + 7) / 8 * 8);
    w->nblocks = nb;
    w->x.alloc(s.npad);
    w->r64.alloc(s.npad);
    w->p64.alloc(s.npad);
    w->q64.alloc(s.npad);
    w->g64.alloc(s.npad);
    w->lam64.alloc(s.npad);
    w->r32.alloc(s.npad);
    w->p32.alloc(s.npad);
    w->q32.alloc(s.npad);
    w->part.alloc(4 * (size_t)nb);
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

void cg_solve(System &s, const double *b, double *x_out,
              const admfem_solve_opts &o, admfem_solve_stats &stats,
              hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  CGWork *w = ensure_work(s);
  const int nb = w->nblocks, ns = s.nslices;
  const bool mixed = o.precision == ADMFEM_PREC_MIXED;
  REQUIRE(o.precision == ADMFEM_PREC_FP64 || mixed, ADMFEM_ERR_ARG, "bad precision");
  REQUIRE(o.rtol > 0 && o.max_iter > 0, ADMFEM_ERR_ARG, "bad rtol/max_iter");
  CGScalars init{};
  init.tol2 = o.rtol * o.rtol;  // scaled by ||b||^2 in k_init
  init.maxit = o.max_iter;
  init.mixed = mixed ? 1 : 0;
  double delta = o.replace_delta > 0 ? o.replace_delta : 1e-2;
  init.delta2 = delta * delta;
  HIPCHECK(hipMemcpyAsync(w->sc.p, &init, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  CGScalars *S = w->sc.p;
  if (mixed) {
    k_init<float><<<nb, NT, 0, st>>>(ns, s.n, s.d_row_bd.p, b, s.d_dinv32.p, w->x.p,
                                     w->r32.p, w->p32.p, w->part.p, S);
  } else {
    k_init<double><<<nb, NT, 0, st>>>(ns, s.n, s.d_row_bd.p, b, s.d_dinv.p, w->x.p,
                                      w->r64.p, w->p64.p, w->part.p, S);
  }
  HIPCHECK(hipGetLastError());
  int chunk = o.check_every > 0 ? o.check_every : 16;
  int issued = 0;
  double rr_last = -1;
  int it_last = 0;
  for (;;) {
    for (int k = 0; k < chunk; k++) {
      if (mixed) {
        k_spmv<float><<<nb, NT, 0, st>>>(ns, s.d_slice_ptr.p, s.d_sell_col.p,
                                         s.d_sell_val32.p, s.d_diag32.p, w->p32.p,
                                         w->q32.p, w->part.p, S);
        k_update<float><<<nb, NT, 0, st>>>(ns, w->p32.p, w->q32.p, w->x.p, w->r32.p,
                                           s.d_dinv32.p, w->part.p, S);
        k_replace<<<nb, NT, 0, st>>>(ns, s.n, s.d_slice_ptr.p, s.d_sell_col.p,
                                     s.d_sell_val.p, s.d_diag.p, b, w->x.p, w->r32.p,
                                     s.d_dinv32.p, w->part.p, S);
        k_pupdate<float><<<nb, NT, 0, st>>>(ns, w->r32.p, s.d_dinv32.p, w->p32.p, S);
      } else {
        k_spmv<double><<<nb, NT, 0, st>>>(ns, s.d_slice_ptr.p, s.d_sell_col.p,
                                          s.d_sell_val.p, s.d_diag.p, w->p64.p,
                                          w->q64.p, w->part.p, S);
        k_update<double><<<nb, NT, 0, st>>>(ns, w->p64.p, w->q64.p, w->x.p, w->r64.p,
                                            s.d_dinv.p, w->part.p, S);
        k_pupdate<double><<<nb, NT, 0, st>>>(ns, w->r64.p, s.d_dinv.p, w->p64.p, S);
      }
    }
    HIPCHECK(hipGetLastError());
    issued += chunk;
    HIPCHECK(hipMemcpyAsync(w->h_sc, S, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const CGScalars &h = *w->h_sc;
    if (h.status != ST_RUN) break;
    if (issued > o.max_iter + 64) break;  // safety net; kernels stop at maxit
    // next chunk: estimate the iterations left from the observed rate
    int next = 16;
    if (rr_last > 0 && h.rr > 0 && h.it > it_last && h.rr < rr_last) {
      double rate = std::log(h.rr / rr_last) / (h.it - it_last);
      double left = std::log(h.tol2 / h.rr) / rate;
      if (std::isfinite(left)) next = (int)std::min(512.0, std::max(8.0, left));
    }
    rr_last = h.rr;
    it_last = h.it;
    chunk = o.check_every > 0 ? o.check_every : next;
  }
  // final true residual in FP64
  k_true_residual<<<nb, NT, 0, st>>>(ns, s.n, s.d_slice_ptr.p, s.d_sell_col.p,
                                     s.d_sell_val.p, s.d_diag.p, b, w->x.p, w->part.p, S);
  HIPCHECK(hipGetLastError());
  if (x_out != w->x.p)
    HIPCHECK(hipMemcpyAsync(x_out, w->x.p, sizeof(double) * s.n, hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipMemcpyAsync(w->h_sc, S, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  const CGScalars &h = *w->h_sc;
  stats.iterations = h.it;
  stats.replacements = h.nrepl;
  stats.rel_residual = h.bb > 0 ? std::sqrt(h.rr_true / h.bb) : 0.0;
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
