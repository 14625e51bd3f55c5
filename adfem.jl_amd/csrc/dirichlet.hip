// ImposeDirichlet and ImposeDirichletGrad on COO triplets
// (deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:27-93), on the device.
//
// Forward: triplets with both ends off the boundary are kept in input order
// (stable compaction by prefix sum), then (b,b,1.0) for every distinct
// boundary node in ascending order.  The RHS lifting rhs[i] -= v*bdval[j]
// (i interior, j boundary) is done per row in triplet order -- the lifted
// triplets are stably sorted by row and each row's run is summed serially --
// so orhs is bitwise the reference's.  Duplicate bd entries: the last one
// wins, as std::map assignment does.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <numeric>

#include "geom.h"

namespace admfem {

namespace {

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void fill_int(int *a, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// bdidx[node] = index into bd/bdval of the entry that wins for node
__global__ void set_bdidx(const int *__restrict__ keys, const int *__restrict__ win,
                          int nu, int *__restrict__ bdidx) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nu) bdidx[keys[t]] = win[t];
}

// keep flag (both ends interior) and lift flag (i interior, j boundary)
__global__ void classify(const long long *__restrict__ ind, int64_t sN, int N,
                         const int *__restrict__ bdidx, int *__restrict__ keep,
                         int *__restrict__ lift, int *__restrict__ err) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sN) return;
  long long i = ind[2 * k], j = ind[2 * k + 1];
  if (i < 0 || i >= N || j < 0 || j >= N) {  // never index with a bad id
    keep[k] = 0;
    lift[k] = 0;
    *err = 1;
    return;
  }
  bool bi = bdidx[i] >= 0, bj = bdidx[j] >= 0;
  keep[k] = (!bi && !bj) ? 1 : 0;
  lift[k] = (!bi && bj) ? 1 : 0;
}

__global__ void scatter_kept(const long long *__restrict__ ind,
                             const double *__restrict__ vv, int64_t sN,
                             const int *__restrict__ keep,
                             const int64_t *__restrict__ pos,
                             long long *__restrict__ oind, double *__restrict__ ov,
                             bool ind16) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sN || !keep[k]) return;
  store_triplet(oind, ov, pos[k], ind[2 * k], ind[2 * k + 1], vv[k], ind16);
}

// lifted triplet list: key = row (forward) or column (backward), value = k
__global__ void scatter_lift(const long long *__restrict__ ind, int64_t sN,
                             const int *__restrict__ lift,
                             const int64_t *__restrict__ lpos, int use_col,
                             int *__restrict__ lkey, int64_t *__restrict__ lval) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sN || !lift[k]) return;
  int64_t z = lpos[k];
  lkey[z] = (int)ind[2 * k + use_col];
  lval[z] = k;
}

__global__ void append_diag(const int *__restrict__ keys, int nu, int64_t base,
                            long long *__restrict__ oind, double *__restrict__ ov) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nu) return;
  oind[2 * (base + t)] = keys[t];
  oind[2 * (base + t) + 1] = keys[t];
  ov[base + t] = 1.0;
}

__global__ void copy_rhs(const double *__restrict__ rhs, double *__restrict__ orhs, int N) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) orhs[i] = rhs[i];
}

// one thread per run of equal keys in the sorted lift list: serial update in
// triplet order (the reference's loop order for that row)
__global__ void lift_rows(const int *__restrict__ skey, const int64_t *__restrict__ sval,
                          int64_t nl, const long long *__restrict__ ind,
                          const double *__restrict__ vv,
                          const int *__restrict__ bdidx,
                          const double *__restrict__ bdval,
                          double *__restrict__ orhs) {
#pragma clang fp contract(off)
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl || (t > 0 && skey[t] == skey[t - 1])) return;
  int row = skey[t];
  double acc = orhs[row];
  for (int64_t u = t; u < nl && skey[u] == row; u++) {
    int64_t k = sval[u];
    acc = acc - vv[k] * bdval[bdidx[ind[2 * k + 1]]];
  }
  orhs[row] = acc;
}

__global__ void set_bd_rhs(const int *__restrict__ keys, const int *__restrict__ win,
                           int nu, const double *__restrict__ bdval,
                           double *__restrict__ orhs) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nu) orhs[keys[t]] = bdval[win[t]];
}

// backward: grad_vv, grad_rhs
__global__ void grad_vv_rhs(const long long *__restrict__ ind, int64_t sN,
                            const int *__restrict__ keep,
                            const int64_t *__restrict__ pos,
                            const int *__restrict__ bdidx,
                            const double *__restrict__ bdval,
                            const double *__restrict__ gov,
                            const double *__restrict__ gorhs,
                            double *__restrict__ gvv) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sN) return;
  long long i = ind[2 * k], j = ind[2 * k + 1];
  bool bi = bdidx[i] >= 0, bj = bdidx[j] >= 0;
  double g = keep[k] ? gov[pos[k]] : 0.0;
  if (!bi && bj) g -= bdval[bdidx[j]] * gorhs[i];
  gvv[k] = g;
}

__global__ void grad_rhs_k(const int *__restrict__ bdidx, const double *__restrict__ gorhs,
                           double *__restrict__ grhs, int N) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) grhs[i] = bdidx[i] >= 0 ? 0.0 : gorhs[i];
}

__global__ void zero_d(double *a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// grad_bdval[win(j)] -= v_k * grad_orhs[i_k] over lifted triplets grouped by
// column j (serial per column in triplet order), then += grad_orhs[j].
__global__ void grad_bdval_k(const int *__restrict__ skey, const int64_t *__restrict__ sval,
                             int64_t nl, const long long *__restrict__ ind,
                             const double *__restrict__ vv,
                             const int *__restrict__ bdidx,
                             const double *__restrict__ gorhs,
                             double *__restrict__ gbd) {
#pragma clang fp contract(off)
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl || (t > 0 && skey[t] == skey[t - 1])) return;
  int j = skey[t];
  int w = bdidx[j];
  double acc = gbd[w];
  for (int64_t u = t; u < nl && skey[u] == j; u++) {
    int64_t k = sval[u];
    acc = acc - vv[k] * gorhs[ind[2 * k]];
  }
  gbd[w] = acc;
}

__global__ void grad_bdval_diag(const int *__restrict__ keys, const int *__restrict__ win,
                                int nu, const double *__restrict__ gorhs,
                                double *__restrict__ gbd) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nu) gbd[win[t]] += gorhs[keys[t]];
}

// Host side: distinct boundary nodes (0-based, ascending) and the index of
// the winning (last) bd entry of each.
void bd_keys(const long long *bd_dev, int bdN, int N, hipStream_t st,
             std::vector<int> &keys, std::vector<int> &win) {
  std::vector<long long> bd(bdN);
  if (bdN)
    HIPCHECK(hipMemcpyAsync(bd.data(), bd_dev, sizeof(long long) * bdN,
                            hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  std::vector<int> idx(bdN);
  std::iota(idx.begin(), idx.end(), 0);
  for (int i = 0; i < bdN; i++)
    REQUIRE(bd[i] >= 1 && bd[i] <= N, ADMFEM_ERR_ARG, "bd entry out of range (1-based)");
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return bd[a] < bd[b]; });
  keys.clear();
  win.clear();
  for (int t = 0; t < bdN; t++) {
    int key = (int)(bd[idx[t]] - 1);
    if (!keys.empty() && keys.back() == key) {
      win.back() = idx[t];  // later entry wins
    } else {
      keys.push_back(key);
      win.push_back(idx[t]);
    }
  }
}

// Per-device workspace of the op, grown on demand and kept until the process
// ends (never freed at exit): a call allocates nothing once it has seen its
// sizes.  Calls on one device are serialised by its lock.
struct DirWork {
  std::mutex mu;
  DevBuf<int> keys, win, bdidx, keep, lift, err, lkey, skey;
  DevBuf<int64_t> pos, lpos, lval, sval, cnt;
  DevBuf<char> tmp;
};

template <class T>
void ensure(DevBuf<T> &b, size_t n) {
  if (b.n < n) b.alloc(std::max(n, b.n + b.n / 4));
}

DirWork &dir_work() {
  static std::mutex mu;
  static DirWork *w[64] = {};
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  REQUIRE(dev >= 0 && dev < 64, ADMFEM_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> g(mu);
  if (!w[dev]) w[dev] = new DirWork();
  return *w[dev];
}

struct DirPrep {
  DirWork &w;
  int nu = 0;
  int64_t nkept = 0, nlift = 0;
};

void exclusive_scan(DirWork &w, const int *in, int64_t *out, int64_t n, hipStream_t st) {
  size_t tmp = 0;
  HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, st));
  ensure(w.tmp, std::max<size_t>(tmp, 1));
  HIPCHECK(hipcub::DeviceScan::ExclusiveSum(w.tmp.p, tmp, in, out, (int)n, st));
}

__global__ void set_i(int *__restrict__ a, int v) { *a = v; }

void prepare(DirPrep &d, const long long *ind, int64_t sN, const long long *bd, int bdN,
             int N, hipStream_t st) {
  REQUIRE(sN >= 0 && sN < (int64_t)INT32_MAX, ADMFEM_ERR_ARG, "triplet count out of range");
  DirWork &w = d.w;
  std::vector<int> keys, win;
  bd_keys(bd, bdN, N, st, keys, win);
  d.nu = (int)keys.size();
  ensure(w.keys, std::max<size_t>(keys.size(), 1));
  ensure(w.win, std::max<size_t>(win.size(), 1));
  if (d.nu) {
    HIPCHECK(hipMemcpyAsync(w.keys.p, keys.data(), sizeof(int) * d.nu, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(w.win.p, win.data(), sizeof(int) * d.nu, hipMemcpyHostToDevice, st));
  }
  ensure(w.bdidx, N);
  fill_int<<<nblk(N, 256), 256, 0, st>>>(w.bdidx.p, N, -1);
  if (d.nu) set_bdidx<<<nblk(d.nu, 256), 256, 0, st>>>(w.keys.p, w.win.p, d.nu, w.bdidx.p);
  ensure(w.keep, sN + 1);
  ensure(w.lift, sN + 1);
  ensure(w.pos, sN + 1);
  ensure(w.lpos, sN + 1);
  ensure(w.err, 1);
  set_i<<<1, 1, 0, st>>>(w.err.p, 0);
  set_i<<<1, 1, 0, st>>>(w.keep.p + sN, 0);  // the scans' last entry: their totals
  set_i<<<1, 1, 0, st>>>(w.lift.p + sN, 0);
  if (sN)
    classify<<<nblk(sN, 256), 256, 0, st>>>(ind, sN, N, w.bdidx.p, w.keep.p, w.lift.p, w.err.p);
  HIPCHECK(hipGetLastError());
  exclusive_scan(w, w.keep.p, w.pos.p, sN + 1, st);
  exclusive_scan(w, w.lift.p, w.lpos.p, sN + 1, st);
  int herr = 0;
  HIPCHECK(hipMemcpyAsync(&herr, w.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(&d.nkept, w.pos.p + sN, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(&d.nlift, w.lpos.p + sN, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  REQUIRE(!herr, ADMFEM_ERR_ARG, "triplet index out of range [0, N)");
}

// stably sorted (key, triplet) list of the lifted triplets, in w.skey / w.sval
void sorted_lift(const DirPrep &d, const long long *ind, int64_t sN, int use_col,
                 hipStream_t st) {
  DirWork &w = d.w;
  const int64_t nl = d.nlift;
  ensure(w.lkey, std::max<int64_t>(nl, 1));
  ensure(w.lval, std::max<int64_t>(nl, 1));
  ensure(w.skey, std::max<int64_t>(nl, 1));
  ensure(w.sval, std::max<int64_t>(nl, 1));
  if (nl == 0) return;
  scatter_lift<<<nblk(sN, 256), 256, 0, st>>>(ind, sN, w.lift.p, w.lpos.p, use_col, w.lkey.p,
                                              w.lval.p);
  size_t tmp = 0;
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, w.lkey.p, w.skey.p, w.lval.p,
                                              w.sval.p, (int)nl, 0, 32, st));
  ensure(w.tmp, std::max<size_t>(tmp, 1));
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(w.tmp.p, tmp, w.lkey.p, w.skey.p, w.lval.p,
                                              w.sval.p, (int)nl, 0, 32, st));
}

}  // namespace

int64_t run_impose_dirichlet(const long long *ind, const double *vv, int64_t sN,
                             const long long *bd, const double *bdval, int bdN,
                             const double *rhs, int N, long long *oind, double *ov,
                             double *orhs, hipStream_t st) {
  REQUIRE(N > 0 && bdN >= 0, ADMFEM_ERR_ARG, "bad sizes");
  DirWork &w = dir_work();
  std::lock_guard<std::mutex> g(w.mu);
  DirPrep d{w};
  prepare(d, ind, sN, bd, bdN, N, st);
  if (sN)
    scatter_kept<<<nblk(sN, 256), 256, 0, st>>>(ind, vv, sN, w.keep.p, w.pos.p, oind, ov,
                                                ((uintptr_t)oind & 15) == 0);
  if (d.nu) append_diag<<<nblk(d.nu, 256), 256, 0, st>>>(w.keys.p, d.nu, d.nkept, oind, ov);
  copy_rhs<<<nblk(N, 256), 256, 0, st>>>(rhs, orhs, N);
  sorted_lift(d, ind, sN, 0, st);
  if (d.nlift)
    lift_rows<<<nblk(d.nlift, 256), 256, 0, st>>>(w.skey.p, w.sval.p, d.nlift, ind, vv,
                                                  w.bdidx.p, bdval, orhs);
  if (d.nu) set_bd_rhs<<<nblk(d.nu, 256), 256, 0, st>>>(w.keys.p, w.win.p, d.nu, bdval, orhs);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
  return d.nkept + d.nu;
}

void run_impose_dirichlet_grad(const double *gov, const double *gorhs,
                               const long long *ind, const double *vv, int64_t sN,
                               const long long *bd, const double *bdval, int bdN, int N,
                               double *gvv, double *grhs, double *gbd, hipStream_t st) {
  REQUIRE(N > 0 && bdN >= 0, ADMFEM_ERR_ARG, "bad sizes");
  DirWork &w = dir_work();
  std::lock_guard<std::mutex> g(w.mu);
  DirPrep d{w};
  prepare(d, ind, sN, bd, bdN, N, st);
  if (sN)
    grad_vv_rhs<<<nblk(sN, 256), 256, 0, st>>>(ind, sN, w.keep.p, w.pos.p, w.bdidx.p, bdval,
                                               gov, gorhs, gvv);
  grad_rhs_k<<<nblk(N, 256), 256, 0, st>>>(w.bdidx.p, gorhs, grhs, N);
  if (bdN) zero_d<<<nblk(bdN, 256), 256, 0, st>>>(gbd, bdN);
  sorted_lift(d, ind, sN, 1, st);
  if (d.nlift)
    grad_bdval_k<<<nblk(d.nlift, 256), 256, 0, st>>>(w.skey.p, w.sval.p, d.nlift, ind, vv,
                                                     w.bdidx.p, gorhs, gbd);
  if (d.nu)
    grad_bdval_diag<<<nblk(d.nu, 256), 256, 0, st>>>(w.keys.p, w.win.p, d.nu, gorhs, gbd);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace admfem
