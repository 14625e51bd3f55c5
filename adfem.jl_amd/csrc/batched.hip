// Batched mixed-precision Jacobi-PCG over many independent systems at once,
// for the serving engine (server.cpp), plus on-device synthetic inputs for
// the serving workload.
//
// One launch serves every active system: a block map assigns each block of
// the grid to (system, block-within-system); per-system descriptors hold the
// system's matrix and vectors, per-system scalars its CG state.  Each
// system's reductions run over its own partials in a fixed order, and its
// replacement schedule depends only on its own iteration count, so a system
// produces bit-identical results alone or in any mix.
//
// Jacobi-PCG on A x = b is run in its symmetrically scaled form: plain CG on
// Â y = S r with Â = S A S, S = diag(1/sqrt(a_ii)) (unit diagonal), x = S y
// -- the same iterates in exact arithmetic, but the FP32 inner iteration
// reads no diagonal, no inverse diagonal and keeps no separate z = D^-1 r.
//
// One CG iteration is one wide kernel and one single-block-per-system
// kernel (FP32 storage, FP64 arithmetic):
//   bk_step  applies the pending step and forms the next direction,
//              r' = r - alpha q,  y += alpha p,  p' = r' + beta p,
//              q' = Â p' = Â r' + beta q
//            (a neighbour contributes its r' = r - alpha q, formed on the fly
//            from its r and q), and the partials of r'.r', p'.q', r'.q', q'.q'
//   bs_step  alpha' = r'.r' / p'.q' and beta' = |r' - alpha' q'|^2 / r'.r',
//            the numerator expanded as r'.r' - 2 alpha' r'.q' + alpha'^2 q'.q'
//            (exact for the stored vectors up to rounding, one step ahead,
//            so no second reduction is needed before p'' is formed);
//            a converged-looking residual freezes the system
// so r.r is reduced directly every iteration (never carried by a
// recurrence), and the iteration is standard CG's (Hestenes-Stiefel) with
// both reductions of an iteration done in one pass.  Every 8th iteration of
// a chunk (a replacement point):
//   bk_replace (frozen systems, residual fell by delta since the last
//              replacement, or stalled): x' = x + S y in FP64, the true
//              residual b - A x' with the FP64 matrix, r = S (b - A x'), y' = 0,
//              and q = Â p afresh for the kept direction (the step carries q
//              by recurrence)
//   bs_replace true-residual convergence test, swap the x / y buffers
// The pending alpha and beta carry over a replacement (the direction is
// kept; reliable updates, van der Vorst & Ye 2000).  Systems start at chunk
// boundaries, so a system's replacement points fall on the same iterations
// of its own count whatever else is in the batch.  No atomics, no fences
// (see cg.hip).
#include <cmath>

#include "batched.h"
#include "sell.h"
#include "geom.h"

namespace admfem {

namespace {

constexpr int NT = BATCH_NT;
constexpr int ST_RUN = 0, ST_CONV = 1, ST_MAXIT = 2, ST_NAN = 3, ST_BREAK = 4;

// This block's sums of NV values, stored to part[j * G + lb].
template <int NV>
__device__ __forceinline__ void put_partials(const double (&v)[NV], double *part, int G,
                                             int lb) {
  __shared__ double sh[NV][NT / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double t[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) t[j] = wave_sum(v[j]);
  __syncthreads();  // sh may still be read by a previous call
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) sh[j][wv] = t[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) {
      double a = 0.0;
#pragma unroll
      for (int k = 0; k < NT / 64; k++) a += sh[j][k];
      global_mut(part)[j * G + lb] = a;
    }
  }
}

// Sums of NV partial arrays part[j * G .. j * G + G) over one block (fixed
// order), valid in every thread.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *part_, int G, double (&o)[NV]) {
  __shared__ double sh[NV][NT / 64];
  const gptr<double> part = global_ptr(part_);
  double a[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) a[j] = 0.0;
  // part[k], part[k + NT], ... added in that order, their loads issued 4 at
  // a time so that they are in flight together
  int k = threadIdx.x;
  for (; k + 3 * NT < G; k += 4 * NT) {
    double v[NV][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int j = 0; j < NV; j++) v[j][u] = part[j * G + k + u * NT];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int j = 0; j < NV; j++) a[j] += v[j][u];
  }
  for (; k < G; k += NT) {
#pragma unroll
    for (int j = 0; j < NV; j++) a[j] += part[j * G + k];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; j++) a[j] = wave_sum(a[j]);
  __syncthreads();  // sh may still be read by a previous call
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) sh[j][wv] = a[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; j++) {
    o[j] = 0.0;
#pragma unroll
    for (int k = 0; k < NT / 64; k++) o[j] += sh[j][k];
  }
}

// --------------------------------------------------------------- CG kernels
// init of ONE system (grid = its nblk): x0 = b on identity rows, 0 elsewhere;
// r0 = S (b - A x0) (b on the other rows, 0 on identity rows: their columns
// are eliminated); p = q = 0 and y = 0 (with alpha = beta = 0 the first step
// forms p = r0, q = Â r0); partials of the scaled r.r, ||b - A x0||^2 and b.b
__global__ void __launch_bounds__(NT)
    bk_init(const BSys *__restrict__ D, int sys) {
  const BSys &d = D[sys];
  const int i = blockIdx.x * NT + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < d.npad) {
    const double bi = i < d.n ? d.b[i] : 0.0;
    const bool ident = i >= d.n || d.row_bd[i] >= 0;
    const double ri = ident ? 0.0 : bi;
    d.x[0][i] = (ident && i < d.n) ? bi : 0.0;
    d.ys[0][i] = 0.0f;
    const float rv = (float)(ri * d.dsq[i]);
    d.r[0][i] = rv;
    d.p[0][i] = 0.0f;
    d.q[0][i] = 0.0f;
    v[0] = (double)rv * (double)rv;
    v[1] = ri * ri;
    v[2] = bi * bi;
  }
  put_partials<3>(v, d.part, d.nblk, blockIdx.x);
}

__global__ void __launch_bounds__(NT) bs_init(const BSys *__restrict__ D, BScal *S, int sys,
                                              double rtol2, double delta2, int maxit) {
  const BSys &d = D[sys];
  double t[3];
  sum_partials<3>(d.part, d.nblk, t);
  if (threadIdx.x != 0) return;
  const double rz = t[0], rr = t[1], bb = t[2];
  BScal &s = S[sys];
  s.rtol2 = rtol2;
  s.delta2 = delta2;
  s.maxit = maxit;
  s.bb = bb;
  s.tol2 = rtol2 * bb;
  s.rz = rz;
  s.rz_ref = rz;
  s.rz_min = rz;
  s.rr = rr;
  s.rr_ref = rr;
  s.stall = 0;
  s.alpha = 0.0;
  s.beta = 0.0;
  s.it = 0;
  s.replace = 0;
  s.nrepl = 0;
  s.sel = 0;
  s.status = (!(rr == rr) || !isfinite(rr)) ? ST_NAN : (rr <= s.tol2 ? ST_CONV : ST_RUN);
}

// The pending step's residual and the next direction of one row:
// r' = r - alpha q and p' = r' + beta p in FP64, p' rounded as it is stored.
// Rows and their neighbours evaluate it with the same explicit fused
// multiply-adds, so q' = Â p' is the product with the stored p' exactly.
struct Step {
  double r, p;
};
__device__ __forceinline__ Step step_of(float r, float q, float p, double alpha, double beta) {
  const double rn = fma(-alpha, (double)q, (double)r);
  return Step{rn, (double)(float)fma(beta, (double)p, rn)};
}

__global__ void __launch_bounds__(NT)
    bk_step(const int2 *__restrict__ map, const BSys *__restrict__ D, const BScal *S) {
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;
  const BScal &s = S[m.x];
  if (s.status != ST_RUN || s.replace) return;
  const BSys &d = D[m.x];
  const int i = m.y * NT + threadIdx.x;
  const double alpha = s.alpha, beta = s.beta;
  const int cur = s.it & 1;
  const gptr<float> r = global_ptr(d.r[cur]);
  const gptr<float> p = global_ptr(d.p[cur]);
  const gptr<float> q = global_ptr(d.q[cur]);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < d.npad) {
    const float pi = p[i], qi = q[i];
    const Step own = step_of(r[i], qi, pi, alpha, beta);
    const gmptr<float> ys = global_mut(d.ys[s.sel]);
    ys[i] = (float)fma(alpha, (double)pi, (double)ys[i]);
    // q' = Â p' = Â r' + beta Â p = Â r' + beta q: the neighbours contribute
    // their r' = r - alpha q (two gathered values each, not three)
    auto nb = [&](int c) { return make_float2(r[c], q[c]); };
    auto rn = [&](float2 v) { return fma(-alpha, (double)v.y, (double)v.x); };
    const double a =
        fma(beta, (double)qi,
            own.r + (d.sd ? sell_dot16<float>(i, d.sp, d.sd, d.sv32, nb, rn)
                          : sell_dot<float>(i, d.sp, d.sc, d.sv32, nb, rn)));
    const float rv = (float)own.r, pv = (float)own.p, qv = (float)a;
    global_mut(d.r[cur ^ 1])[i] = rv;
    global_mut(d.p[cur ^ 1])[i] = pv;
    global_mut(d.q[cur ^ 1])[i] = qv;
    v[0] = (double)rv * (double)rv;
    v[1] = (double)pv * (double)qv;
    v[2] = (double)rv * (double)qv;
    v[3] = (double)qv * (double)qv;
  }
  put_partials<4>(v, d.part, d.nblk, m.y);
}

__global__ void __launch_bounds__(NT)
    bs_step(const int *__restrict__ list, const BSys *__restrict__ D, BScal *S) {
  const int sys = list[blockIdx.x];
  BScal &s = S[sys];
  if (s.status != ST_RUN || s.replace) return;
  double t[4];
  sum_partials<4>(D[sys].part, D[sys].nblk, t);
  __syncthreads();  // every thread has read the scalars it tests
  if (threadIdx.x != 0) return;
  const double rz = t[0], pq = t[1], rq = t[2], qq = t[3];
  s.it += 1;
  if (!(rz == rz) || !isfinite(rz) || pq != pq) {
    s.status = ST_NAN;
    return;
  }
  s.rz = rz;
  if (rz < 0.98 * s.rz_min) {
    s.rz_min = rz;
    s.stall = 0;
  } else {
    s.stall += 1;
  }
  // a scaled residual that looks converged -- below the tolerance times the
  // scaled-to-true ratio of the last replacement -- freezes the system until
  // the next replacement point, where the FP64 residual decides
  if (rz * s.rr_ref <= s.tol2 * s.rz_ref) s.replace = 1;
  else if (s.it >= s.maxit) s.status = ST_MAXIT;
  if (!(pq > 0.0)) {  // not SPD (or p = 0 with r = 0: the replacement decides)
    if (!s.replace) s.status = ST_BREAK;
    s.alpha = 0.0;
    s.beta = 0.0;
    return;
  }
  const double alpha = rz / pq;
  const double next = fmax(rz - 2.0 * alpha * rq + alpha * alpha * qq, 0.0);
  s.alpha = alpha;
  s.beta = rz > 0.0 ? next / rz : 0.0;
}

// A system is replaced at a replacement point when it is frozen (converged
// looking), its residual fell by delta since the last replacement, or its
// FP32 residual has stalled.  Every kernel of the point reads the same,
// unchanged scalars: one decision.
__device__ __forceinline__ bool replace_due(const BScal &s) {
  return s.status == ST_RUN &&
         (s.replace || s.rz <= s.delta2 * s.rz_ref || s.stall >= STALL_ITERS);
}

// due systems: x' = x + S y into the other buffer, the true residual
// b - A64 x', r = S (b - A64 x'), y' = 0
__global__ void __launch_bounds__(NT)
    bk_replace(const int2 *__restrict__ map, const BSys *__restrict__ D, const BScal *S) {
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;
  const BScal &s = S[m.x];
  if (!replace_due(s)) return;
  const BSys &d = D[m.x];
  const int i = m.y * NT + threadIdx.x;
  const gptr<double> x = global_ptr(d.x[s.sel]);
  const gptr<float> ys = global_ptr(d.ys[s.sel]);
  const gptr<double> dsq = global_ptr(d.dsq);
  double v[2] = {0.0, 0.0};
  if (i < d.npad) {
    const double xi = x[i] + dsq[i] * (double)ys[i];
    // the true residual's product (FP64) and q = Â p afresh for the kept
    // direction (FP32 matrix; the step updates q by recurrence), one pass
    const gptr<float> p = global_ptr(d.p[s.it & 1]);
    struct XDY {
      double x, d;
      float y;
    };
    double sx, sq;
    sell_dot2(
        i, d.sp, d.sc, d.sv64, d.sv32, [&](int c) { return XDY{x[c], dsq[c], ys[c]}; },
        [](XDY t) { return t.x + t.d * (double)t.y; }, [&](int c) { return p[c]; },
        [](float t) { return (double)t; }, sx, sq);
    const double ax = global_ptr(d.diag64)[i] * xi + sx;
    global_mut(d.x[s.sel ^ 1])[i] = xi;
    global_mut(d.ys[s.sel ^ 1])[i] = 0.0f;
    const double ri = (i < d.n ? global_ptr(d.b)[i] : 0.0) - ax;
    const float rf = (float)(ri * dsq[i]);
    global_mut(d.r[s.it & 1])[i] = rf;
    global_mut(d.q[s.it & 1])[i] = (float)((double)p[i] + sq);
    v[0] = (double)rf * (double)rf;
    v[1] = ri * ri;
  }
  put_partials<2>(v, d.part, d.nblk, m.y);
}

__global__ void __launch_bounds__(NT)
    bs_replace(const int *__restrict__ list, const BSys *__restrict__ D, BScal *S) {
  const int sys = list[blockIdx.x];
  BScal &s = S[sys];
  if (!replace_due(s)) return;
  double t[2];
  sum_partials<2>(D[sys].part, D[sys].nblk, t);
  __syncthreads();  // every thread has read the decision's inputs
  if (threadIdx.x != 0) return;
  const double rz = t[0], rr = t[1];
  s.rz = rz;
  s.rr = rr;
  s.rr_ref = rr;
  s.rz_ref = rz;
  s.rz_min = rz;
  s.stall = 0;
  s.replace = 0;
  s.nrepl += 1;
  s.sel ^= 1;
  if (!(rr == rr) || !isfinite(rr)) s.status = ST_NAN;
  else if (rr <= s.tol2) s.status = ST_CONV;
  else if (s.it >= s.maxit) s.status = ST_MAXIT;
}

// solution of a finished system: out = x + S y (y is zero after a final
// replacement)
__global__ void bk_final(const BSys *__restrict__ D, const BScal *S, int sys,
                         double *__restrict__ out) {
  const BSys &d = D[sys];
  const int sel = S[sys].sel;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.npad) out[i] = d.x[sel][i] + d.dsq[i] * (double)d.ys[sel][i];
}

// Scaling of one assembled system: dsq = 1/sqrt(a_ii) (FP64) and the FP32
// scaled off-diagonals a_ij dsq_i dsq_j in the SELL layout.
__global__ void bk_scale(int npad, const int *__restrict__ sp, const int *__restrict__ sc,
                         const double *__restrict__ sv64, const double *__restrict__ diag64,
                         double *__restrict__ dsq, float *__restrict__ sv32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const double si = 1.0 / sqrt(diag64[i]);
  dsq[i] = si;
  const int s0 = sp[i >> 6], w = (sp[(i >> 6) + 1] - s0) >> 6, off = s0 + (i & 63);
  for (int k = 0; k < w; k++) {
    const int e = off + 64 * k;
    sv32[e] = (float)(sv64[e] * si * (1.0 / sqrt(diag64[sc[e]])));
  }
}

// --------------------------------------------------------------- synthesis
__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unif(uint64_t seed, uint64_t k) {  // [0, 1)
  return (double)(mix64(seed ^ mix64(k)) >> 11) * (1.0 / 9007199254740992.0);
}

// coordinates (internal order): base + U[-a, a]^2 on non-boundary nodes
__global__ void syn_coords(int n, const double *__restrict__ bx, const double *__restrict__ by,
                           const int *__restrict__ row_bd, const int *__restrict__ vrow,
                           const int *__restrict__ iperm, uint64_t seed, double amp,
                           double *__restrict__ x, double *__restrict__ y,
                           double2 *__restrict__ xy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double dx = 0, dy = 0;
  if (row_bd[vrow ? vrow[r] : r] < 0) {  // vertices on Dirichlet rows stay put
    const uint64_t u = (uint64_t)iperm[r];  // jitter keyed by the user node id
    dx = amp * (2.0 * unif(seed, 2 * u) - 1.0);
    dy = amp * (2.0 * unif(seed, 2 * u + 1) - 1.0);
  }
  x[r] = bx[r] + dx;
  y[r] = by[r] + dy;
  xy[r] = make_double2(bx[r] + dx, by[r] + dy);
}

// per Gauss point: kappa = exp(GRF) and f ~ U[-1, 1]
__global__ void syn_gauss(int nelem, QuadRule rule, const int *__restrict__ conn,
                          const double *__restrict__ x, const double *__restrict__ y,
                          SynField F, uint64_t fseed, double *__restrict__ kappa,
                          double *__restrict__ f) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  double ex[3], ey[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int v = conn[(size_t)k * nelem + e];
    ex[k] = x[v];
    ey[k] = y[v];
  }
  for (int k = 0; k < rule.nq; k++) {
    const double h0 = 1 - rule.xi[k] - rule.eta[k], h1 = rule.xi[k], h2 = rule.eta[k];
    const double gx = ex[0] * h0 + ex[1] * h1 + ex[2] * h2;
    const double gy = ey[0] * h0 + ey[1] * h1 + ey[2] * h2;
    double g = 0.0;
    for (int j = 0; j < F.nmodes; j++) g += cos(F.wx[j] * gx + F.wy[j] * gy + F.phi[j]);
    const size_t gi = (size_t)rule.nq * e + k;
    kappa[gi] = exp(F.scale * g);
    f[gi] = 2.0 * unif(fseed, gi) - 1.0;
  }
}

// heat initial state (internal order): a GRF at the dofs' nodes (an edge
// dof: its edge's midpoint, ends[r]; ends null: row r is vertex r)
__global__ void syn_nodal(int n, const double *__restrict__ x, const double *__restrict__ y,
                          const int2 *__restrict__ ends, SynField F, double *__restrict__ u) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double px = x[r], py = y[r];
  if (ends) {
    const int2 e = ends[r];
    px = 0.5 * (x[e.x] + x[e.y]);
    py = 0.5 * (y[e.x] + y[e.y]);
  }
  double g = 0.0;
  for (int j = 0; j < F.nmodes; j++) g += cos(F.wx[j] * px + F.wy[j] * py + F.phi[j]);
  u[r] = F.scale * g;
}

// The block map of a batch: grid block b -> (system slot, block of that
// system).  The logical list is the systems' blocks in order (system k's
// start at pref[k]); XCD x = b % 8 (blocks are dispatched round-robin over
// the 8 XCDs) runs one contiguous eighth of it, entry j of its share at grid
// block 8 j + x, so neighbouring blocks of a system gather from the same
// L2.  Grid blocks past a share get system -1 (no work).
__global__ void expand_map(int nsys, const int *__restrict__ slot, const int *__restrict__ pref,
                           int nmap, int gmap, int2 *__restrict__ map) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= gmap) return;
  const int x = b & 7, j = b >> 3, base = nmap >> 3, rem = nmap & 7;
  if (j >= base + (x < rem ? 1 : 0)) {
    map[b] = make_int2(-1, 0);
    return;
  }
  const int L = x * base + min(x, rem) + j;
  int lo = 0, hi = nsys - 1;  // the last system starting at or before L
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= L) lo = mid;
    else hi = mid - 1;
  }
  map[b] = make_int2(slot[lo], L - pref[lo]);
}

// user-order interleaved coordinates -> internal-order x, y and xy
__global__ void perm_coords(int n, const int *__restrict__ iperm, const double *__restrict__ c,
                            double *__restrict__ x, double *__restrict__ y,
                            double2 *__restrict__ xy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int u = iperm[r];
  const double cx = c[2 * (size_t)u], cy = c[2 * (size_t)u + 1];
  x[r] = cx;
  y[r] = cy;
  xy[r] = make_double2(cx, cy);
}

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_perm_coords(int n, const int *iperm, const double *coords_user, double *x,
                        double *y, double2 *xy, hipStream_t st) {
  perm_coords<<<nb(n, 256), 256, 0, st>>>(n, iperm, coords_user, x, y, xy);
  HIPCHECK(hipGetLastError());
}

void batch_map(int nsys, const int *slot, const int *pref, int nmap, int gmap, int2 *map,
               hipStream_t st) {
  expand_map<<<nb(gmap, 256), 256, 0, st>>>(nsys, slot, pref, nmap, gmap, map);
  HIPCHECK(hipGetLastError());
}

void batch_init(const BSys *D, BScal *S, int sys, int nblk, double rtol2, double delta2,
                int maxit, hipStream_t st) {
  bk_init<<<nblk, NT, 0, st>>>(D, sys);
  bs_init<<<1, NT, 0, st>>>(D, S, sys, rtol2, delta2, maxit);
  HIPCHECK(hipGetLastError());
}

void batch_iterations(const int2 *map, int nmap, const int *list, int nlist, const BSys *D,
                      BScal *S, int iters, hipStream_t st) {
  REQUIRE(iters % REPLACE_EVERY == 0, ADMFEM_ERR_ARG, "chunk must be a multiple of 8");
  for (int k = 0; k < iters; k++) {
    bk_step<<<nmap, NT, 0, st>>>(map, D, S);
    bs_step<<<nlist, NT, 0, st>>>(list, D, S);
    if (k % REPLACE_EVERY == REPLACE_EVERY - 1) {  // replacement point
      bk_replace<<<nmap, NT, 0, st>>>(map, D, S);
      bs_replace<<<nlist, NT, 0, st>>>(list, D, S);
    }
  }
  HIPCHECK(hipGetLastError());
}

void batch_spmv_repeat(const int2 *map, int nmap, const BSys *D, const BScal *S, int repeat,
                       hipStream_t st) {
  for (int k = 0; k < repeat; k++) bk_step<<<nmap, NT, 0, st>>>(map, D, S);
  HIPCHECK(hipGetLastError());
}

void batch_scale(const System &s, const double *sv64, const double *diag64, double *dsq,
                 float *sv32, hipStream_t st) {
  bk_scale<<<nb(s.npad, 256), 256, 0, st>>>(s.npad, s.d_slice_ptr.p, s.d_sell_col.p, sv64,
                                            diag64, dsq, sv32);
  HIPCHECK(hipGetLastError());
}

void batch_final(const BSys *D, const BScal *S, int sys, int npad, double *out,
                 hipStream_t st) {
  bk_final<<<nb(npad, 256), 256, 0, st>>>(D, S, sys, out);
  HIPCHECK(hipGetLastError());
}

void synth_inputs(const System &s, const double *base_x, const double *base_y, uint64_t seed,
                  double amp, const SynField &kf, const SynField *uf, double *x, double *y,
                  double2 *xy, double *kappa, double *f, double *u_prev_int, hipStream_t st) {
  const Mesh &m = *s.mesh;
  const bool p2 = s.degree == 2;
  syn_coords<<<nb(m.nnode, 256), 256, 0, st>>>(m.nnode, base_x, base_y, s.d_row_bd.p,
                                                p2 ? s.d_vrow.p : nullptr, m.d_iperm.p, seed,
                                                amp, x, y, xy);
  syn_gauss<<<nb(m.nelem, 256), 256, 0, st>>>(m.nelem, m.rule, m.conn.p, x, y, kf,
                                               seed ^ 0x5A5A5A5A5A5A5A5Aull, kappa, f);
  if (uf && u_prev_int)
    syn_nodal<<<nb(s.n, 256), 256, 0, st>>>(s.n, x, y, p2 ? s.d_dof_ends.p : nullptr, *uf,
                                             u_prev_int);
  HIPCHECK(hipGetLastError());
}

}  // namespace admfem
