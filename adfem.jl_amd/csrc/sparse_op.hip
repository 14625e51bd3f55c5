// The general sparse solve u = A \ rhs and its gradient, for a matrix given
// as COO triplets -- ADCME's SparseSolver op (the `\` of a SparseTensor,
// README.md:54,75) and SparseSolverGrad: lambda = A^-T dL/du,
// dL/dvv_k = -lambda[i_k] * u[j_k], dL/drhs = lambda.
//
// Duplicates are summed in triplet order (Eigen's setFromTriplets order);
// the matrix must be symmetric positive definite, since the solve is the
// library's Jacobi-PCG (then A^-T = A^-1).
//
// The pattern is built on the device: a stable radix sort of the keys
// row * n + col (equal keys keep their triplet order), one slot per run of
// equal keys.  The last pattern is kept (per device): a call whose indices
// equal it -- checked exactly on the device -- reuses its pattern, solver
// layout and CG workspace and only sums the values again, as an optimisation
// loop or the gradient after its forward solve does.  Calls are serialised
// by a lock (the kept pattern and its workspace are shared).
//
// The solver works in its own row order: reverse Cuthill-McKee of the
// pattern when that narrows its bandwidth (the caller's numbering may be
// anything -- a randomly numbered mesh makes every SpMV gather at random),
// else the caller's.  The right-hand side is scattered into that order and
// the solution back; the matrix values are the same sums, stored at their
// rows' new places.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <numeric>
#include <utility>

#include "internal.h"

namespace admfem {

System *system_from_device_csr(int n, DevBuf<int> &rp, DevBuf<int> &col, int64_t nnz);

namespace {

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// key = row * n + col, value = triplet index; bad indices raise the flag
// (and get key 0, never used: the call fails)
__global__ void coo_keys(int64_t nnz, const long long *__restrict__ ind, int n,
                         unsigned long long *__restrict__ key, int *__restrict__ val,
                         int *__restrict__ err) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const long long i = ind[2 * k], j = ind[2 * k + 1];
  const bool ok = i >= 0 && i < n && j >= 0 && j < n;
  if (!ok) *err = 1;
  key[k] = ok ? (unsigned long long)i * (unsigned long long)n + (unsigned long long)j : 0ull;
  val[k] = (int)k;
}

__global__ void coo_fresh(int64_t nnz, const unsigned long long *__restrict__ skey,
                          int *__restrict__ fresh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nnz) fresh[t] = (t == 0 || skey[t] != skey[t - 1]) ? 1 : 0;
}

// slot s starts at sorted position start[s]; its row and column
__global__ void coo_slots(int64_t nnz, const unsigned long long *__restrict__ skey,
                          const int *__restrict__ fresh, const int *__restrict__ pos, int n,
                          int *__restrict__ start, int *__restrict__ srow,
                          int *__restrict__ scol) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz || !fresh[t]) return;
  const int s = pos[t];
  start[s] = (int)t;
  srow[s] = (int)(skey[t] / (unsigned long long)n);
  scol[s] = (int)(skey[t] % (unsigned long long)n);
}

// the slots' keys in the solver's row order: inv[row] * n + inv[col]
__global__ void reorder_keys(int nslot, int n, const int *__restrict__ srow,
                             const int *__restrict__ scol, const int *__restrict__ inv,
                             unsigned long long *__restrict__ key, int *__restrict__ val) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  key[s] = (unsigned long long)inv[srow[s]] * (unsigned long long)n + (unsigned long long)inv[scol[s]];
  val[s] = s;
}

// sorted position i holds slot sval[i]: its new row and column, and the
// slot's place in the reordered pattern
__global__ void reorder_out(int nslot, int n, const unsigned long long *__restrict__ skey,
                            const int *__restrict__ sval, int *__restrict__ nrow,
                            int *__restrict__ ncol, int *__restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslot) return;
  nrow[i] = (int)(skey[i] / (unsigned long long)n);
  ncol[i] = (int)(skey[i] % (unsigned long long)n);
  pos[sval[i]] = i;
}

// any index differing from the kept pattern's raises the flag (plain stores
// of the same value: no atomics needed)
__global__ void coo_differs(int64_t m, const long long *__restrict__ a,
                            const long long *__restrict__ b, int *__restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m && a[k] != b[k]) *flag = 1;
}

__global__ void set_int(int *__restrict__ a, int v) { *a = v; }

// val[pos[slot]] (pos null: val[slot]) = sum of vv over the slot's
// triplets, in triplet order
__global__ void coo_sum(int nslot, const int *__restrict__ start, const int *__restrict__ trip,
                        const int *__restrict__ pos, const double *__restrict__ vv,
                        double *__restrict__ val) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  double a = 0.0;
  for (int t = start[s]; t < start[s + 1]; t++) a += vv[trip[t]];
  val[pos ? pos[s] : s] = a;
}

__global__ void coo_grad(int64_t nnz, const long long *__restrict__ ind,
                         const double *__restrict__ lam, const double *__restrict__ u,
                         double *__restrict__ gvv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) gvv[k] = -lam[ind[2 * k]] * u[ind[2 * k + 1]];
}

// A COO matrix's pattern, ready to receive values and solve.
struct CooPattern {
  int device = -1, n = 0, nslot = 0;
  int64_t nnz = -1;
  DevBuf<long long> ind;  // the indices it was built from
  DevBuf<int> start;      // nslot + 1 sorted positions
  DevBuf<int> trip;       // triplet index at each sorted position
  // solver row order (empty when it is the caller's): order[k] = caller's
  // row at position k, inv its inverse, pos[slot] = the slot's place in the
  // reordered CSR; b, x: right-hand side and solution in that order
  bool reordered = false;
  DevBuf<int> order, inv, pos;
  DevBuf<double> b, x;
  DevBuf<int> flag;  // set when a call's indices differ from `ind`
  System *s = nullptr;
};

// The solver's row order for the device CSR pattern (rp, col; slot s at
// row srow[s]): reverse Cuthill-McKee if the caller's order is not banded
// already (bandwidth at most 5x the widest level of a breadth-first search
// from the first row coupled to another row) and RCM narrows the bandwidth;
// then rp and col become the reordered pattern, and the order, its inverse
// and each slot's new place are kept in p.  The search does not start at
// row 0: after Dirichlet elimination that is usually a boundary node holding
// only its diagonal, whose one-node "level" would make every order look
// unbanded.
void reorder(CooPattern &p, int n, int nslot, const int *srow, DevBuf<int> &rp,
             DevBuf<int> &col, hipStream_t st) {
  const int64_t bw_id = device_bandwidth(n, rp.p, col.p, nullptr, st);
  const int start = device_first_coupled(n, rp.p, col.p, st);
  if (bw_id <= 5 * device_bfs_max_level(n, rp.p, col.p, start, st)) return;
  DevBuf<int> order, inv;
  rcm(n, rp.p, col.p, order, inv, st);
  if (device_bandwidth(n, rp.p, col.p, inv.p, st) >= bw_id) return;
  DevBuf<unsigned long long> key(nslot), skey(nslot);
  DevBuf<int> val(nslot), sval(nslot), nrow(nslot), ncol(nslot);
  p.pos.alloc(nslot);
  reorder_keys<<<nb(nslot, 256), 256, 0, st>>>(nslot, n, srow, col.p, inv.p, key.p, val.p);
  HIPCHECK(hipGetLastError());
  const unsigned long long nn = (unsigned long long)n * (unsigned long long)n;
  int bits = 1;
  while (bits < 64 && (nn - 1) >> bits) bits++;
  size_t tmp = 0;
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key.p, skey.p, val.p, sval.p, nslot,
                                              0, bits, st));
  DevBuf<char> t(std::max<size_t>(tmp, 1));
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, key.p, skey.p, val.p, sval.p, nslot, 0,
                                              bits, st));
  reorder_out<<<nb(nslot, 256), 256, 0, st>>>(nslot, n, skey.p, sval.p, nrow.p, ncol.p, p.pos.p);
  HIPCHECK(hipGetLastError());
  csr_row_ptr(nrow.p, nslot, n, rp.p, st);
  col.swap(ncol);
  p.order.swap(order);
  p.inv.swap(inv);
  p.b.alloc(n);
  p.x.alloc(n);
  p.reordered = true;
  HIPCHECK(hipStreamSynchronize(st));  // temporaries go out of scope
}

std::mutex g_lock;
CooPattern *g_kept = nullptr;  // kept until the process ends (never freed at exit)

bool same_pattern(CooPattern &p, const long long *ind, int64_t nnz, int n, int dev,
                  hipStream_t st) {
  if (p.device != dev || p.n != n || p.nnz != nnz) return false;
  if (nnz == 0) return true;
  set_int<<<1, 1, 0, st>>>(p.flag.p, 0);
  coo_differs<<<nb(2 * nnz, 256), 256, 0, st>>>(2 * nnz, ind, p.ind.p, p.flag.p);
  HIPCHECK(hipGetLastError());
  int h = 1;
  HIPCHECK(hipMemcpyAsync(&h, p.flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return h == 0;
}

void build_pattern(CooPattern &p, const long long *ind, int64_t nnz, int n, int dev,
                   hipStream_t st) {
  p.device = dev;
  p.n = n;
  p.nnz = nnz;
  p.flag.alloc(1);
  const int64_t m = std::max<int64_t>(nnz, 1);
  DevBuf<unsigned long long> key(m), skey(m);
  DevBuf<int> val(m), fresh(m), pos(m), err(1);
  p.trip.alloc(m);
  set_int<<<1, 1, 0, st>>>(err.p, 0);
  int nslot = 0;
  DevBuf<int> rp(n + 1), col;
  if (nnz) {
    coo_keys<<<nb(nnz, 256), 256, 0, st>>>(nnz, ind, n, key.p, val.p, err.p);
    HIPCHECK(hipGetLastError());
    int herr = 0;
    HIPCHECK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    REQUIRE(!herr, ADMFEM_ERR_ARG, "triplet index out of range [0, n)");
    const unsigned long long nn = (unsigned long long)n * (unsigned long long)n;
    int bits = 1;
    while (bits < 64 && (nn - 1) >> bits) bits++;
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key.p, skey.p, val.p, p.trip.p,
                                                (int)nnz, 0, bits, st));
    DevBuf<char> t(std::max<size_t>(tmp, 1));
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, key.p, skey.p, val.p, p.trip.p,
                                                (int)nnz, 0, bits, st));
    coo_fresh<<<nb(nnz, 256), 256, 0, st>>>(nnz, skey.p, fresh.p);
    tmp = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, fresh.p, pos.p, (int)nnz, st));
    DevBuf<char> t2(std::max<size_t>(tmp, 1));
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(t2.p, tmp, fresh.p, pos.p, (int)nnz, st));
    int last[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(&last[0], pos.p + nnz - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&last[1], fresh.p + nnz - 1, sizeof(int), hipMemcpyDeviceToHost,
                            st));
    HIPCHECK(hipStreamSynchronize(st));
    nslot = last[0] + last[1];
    p.start.alloc(nslot + 1);
    DevBuf<int> srow(nslot), scol(nslot);
    coo_slots<<<nb(nnz, 256), 256, 0, st>>>(nnz, skey.p, fresh.p, pos.p, n, p.start.p, srow.p,
                                            scol.p);
    HIPCHECK(hipGetLastError());
    const int end = (int)nnz;
    HIPCHECK(hipMemcpyAsync(p.start.p + nslot, &end, sizeof(int), hipMemcpyHostToDevice, st));
    csr_row_ptr(srow.p, nslot, n, rp.p, st);
    col.swap(scol);
    p.nslot = nslot;
    reorder(p, n, nslot, srow.p, rp, col, st);
  } else {
    p.start.alloc(1);
    const int zero = 0;
    HIPCHECK(hipMemcpyAsync(p.start.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemsetAsync(rp.p, 0, sizeof(int) * (n + 1), st));
    p.nslot = 0;
  }
  p.s = system_from_device_csr(n, rp, col, nslot);
  p.ind.alloc(std::max<int64_t>(2 * nnz, 1));
  if (nnz)
    HIPCHECK(hipMemcpyAsync(p.ind.p, ind, sizeof(long long) * 2 * nnz, hipMemcpyDeviceToDevice,
                            st));
  HIPCHECK(hipStreamSynchronize(st));  // temporaries go out of scope
}

// The kept pattern for these indices (rebuilt when they differ), with the
// values summed into its solver format.  Caller holds g_lock.
System &pattern_with_values(const long long *ind, const double *vv, int64_t nnz, int n,
                            hipStream_t st) {
  REQUIRE(n > 0 && nnz >= 0 && nnz < INT32_MAX, ADMFEM_ERR_ARG, "bad sizes");
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (!g_kept || !same_pattern(*g_kept, ind, nnz, n, dev, st)) {
    if (g_kept) {
      system_free(g_kept->s);
      delete g_kept;
      g_kept = nullptr;
    }
    CooPattern *p = new CooPattern();
    try {
      build_pattern(*p, ind, nnz, n, dev, st);
    } catch (...) {
      system_free(p->s);
      delete p;
      throw;
    }
    g_kept = p;
  }
  CooPattern &p = *g_kept;
  if (p.nslot)
    coo_sum<<<nb(p.nslot, 256), 256, 0, st>>>(p.nslot, p.start.p, p.trip.p,
                                               p.reordered ? p.pos.p : nullptr, vv, p.s->val.p);
  HIPCHECK(hipGetLastError());
  launch_to_sell(*p.s, st);
  return *p.s;
}

// x = A^-1 b for the kept pattern p, b and x in the caller's order
void solve_kept(CooPattern &p, const double *b, double *x, const admfem_solve_opts &o,
                admfem_solve_stats &stats, hipStream_t st) {
  if (!p.reordered) return cg_solve(*p.s, b, x, o, stats, st);
  launch_scatter(p.inv.p, b, p.b.p, p.n, st);  // b_int[inv[i]] = b[i]
  cg_solve(*p.s, p.b.p, p.x.p, o, stats, st);
  launch_scatter(p.order.p, p.x.p, x, p.n, st);  // x[order[k]] = x_int[k]
  HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace

void sparse_solve(const long long *ind, const double *vv, int64_t nnz, int n, const double *rhs,
                  double *u, const admfem_solve_opts &o, admfem_solve_stats &stats,
                  hipStream_t st) {
  std::lock_guard<std::mutex> g(g_lock);
  pattern_with_values(ind, vv, nnz, n, st);
  solve_kept(*g_kept, rhs, u, o, stats, st);
}

void sparse_solve_info(int64_t info[4]) {
  std::lock_guard<std::mutex> g(g_lock);
  info[0] = g_kept ? g_kept->n : 0;
  info[1] = g_kept ? g_kept->nnz : 0;
  info[2] = g_kept ? g_kept->nslot : 0;
  info[3] = g_kept && g_kept->reordered ? 1 : 0;
}

void sparse_solve_grad(const long long *ind, const double *vv, int64_t nnz, int n,
                       const double *u, const double *grad_u, double *grad_vv,
                       double *grad_rhs, const admfem_solve_opts &o, admfem_solve_stats &stats,
                       hipStream_t st) {
  std::lock_guard<std::mutex> g(g_lock);
  pattern_with_values(ind, vv, nnz, n, st);
  solve_kept(*g_kept, grad_u, grad_rhs, o, stats, st);  // lambda = A^-1 dL/du = dL/drhs
  if (nnz) coo_grad<<<nb(nnz, 256), 256, 0, st>>>(nnz, ind, grad_rhs, u, grad_vv);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace admfem
