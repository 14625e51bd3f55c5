// ImposeDirichlet and ImposeDirichletGrad on COO triplets
// (deps/MFEM/ImposeDirichlet/ImposeDirichlet.h:27-93), on the device.
//
// Forward: triplets with both ends off the boundary are kept in input order
// (stable compaction), then (b,b,1.0) for every distinct
// boundary node in ascending order.  The RHS lifting rhs[i] -= v*bdval[j]
// (i interior, j boundary) is done per row in triplet order -- the lifted
// triplets are stably sorted by row and each row's run is summed serially --
// so orhs is bitwise the reference's.  Duplicate bd entries: the last one
// wins, as std::map assignment does.
//
// The compaction is a two-pass tiled select with no per-triplet workspace:
// pass 1 counts each tile's kept and lifted triplets (the tile's indices
// read once), a scan of the tile counts gives each tile its output offsets,
// and pass 2 re-reads the tile, ranks its triplets in input order with wave
// ballots and writes them out (the backward pass recomputes the same ranks
// to find each kept triplet's output slot).
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

// Tiles of the triplet list: DIR_ITEMS rows of DIR_NT triplets, item u of
// thread t being triplet tile * DIR_TILE + u * DIR_NT + t (coalesced loads;
// input order is u-major, then thread).
constexpr int DIR_NT = 256;
constexpr int DIR_ITEMS = 8;
constexpr int DIR_TILE = DIR_NT * DIR_ITEMS;
constexpr int DIR_WAVES = DIR_NT / 64;

__device__ __forceinline__ void load_pair(const long long *__restrict__ ind, int64_t k,
                                          bool ind16, long long &i, long long &j) {
  if (ind16) {
    const ll2_t v = reinterpret_cast<const ll2_t *>(ind)[k];
    i = v.x;
    j = v.y;
  } else {
    i = ind[2 * k];
    j = ind[2 * k + 1];
  }
}

// The tile's triplets: indices (checked to lie in [0, N), so int), keep flag
// (both ends interior) and lift flag (i interior, j boundary).  An index
// outside [0, N) flags err and neither.
struct TileItems {
  int i[DIR_ITEMS], j[DIR_ITEMS];
  bool keep[DIR_ITEMS], lift[DIR_ITEMS];
  __device__ __forceinline__ void load(const long long *__restrict__ ind, int64_t sN, int N,
                                       const int *__restrict__ bdidx, bool ind16,
                                       int *__restrict__ err) {
    const int64_t base = (int64_t)blockIdx.x * DIR_TILE + threadIdx.x;
    long long a[DIR_ITEMS], b[DIR_ITEMS];
#pragma unroll
    for (int u = 0; u < DIR_ITEMS; u++) {
      a[u] = 0;
      b[u] = 0;
      const int64_t k = base + (int64_t)u * DIR_NT;
      if (k < sN) load_pair(ind, k, ind16, a[u], b[u]);
    }
#pragma unroll
    for (int u = 0; u < DIR_ITEMS; u++) {
      keep[u] = false;
      lift[u] = false;
      i[u] = 0;
      j[u] = 0;
      const int64_t k = base + (int64_t)u * DIR_NT;
      if (k >= sN) continue;
      if (a[u] < 0 || a[u] >= N || b[u] < 0 || b[u] >= N) {  // never index with a bad id
        *err = 1;
        continue;
      }
      i[u] = (int)a[u];
      j[u] = (int)b[u];
      const bool bi = bdidx[i[u]] >= 0, bj = bdidx[j[u]] >= 0;
      keep[u] = !bi && !bj;
      lift[u] = !bi && bj;
    }
  }
};

// pass 1: the kept and lifted triplets of each tile
__global__ void __launch_bounds__(DIR_NT)
    tile_count(const long long *__restrict__ ind, int64_t sN, int N,
               const int *__restrict__ bdidx, bool ind16, int *__restrict__ tkeep,
               int *__restrict__ tlift, int *__restrict__ err) {
  __shared__ int sk[DIR_WAVES], sl[DIR_WAVES];
  TileItems T;
  T.load(ind, sN, N, bdidx, ind16, err);
  int nk = 0, nl = 0;
#pragma unroll
  for (int u = 0; u < DIR_ITEMS; u++) {
    nk += __popcll(__ballot(T.keep[u]));
    nl += __popcll(__ballot(T.lift[u]));
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sk[wave] = nk;
    sl[wave] = nl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int w = 0; w < DIR_WAVES; w++) {
      a += sk[w];
      b += sl[w];
    }
    tkeep[blockIdx.x] = a;
    tlift[blockIdx.x] = b;
  }
}

// Ranks of the tile's kept and lifted triplets in input order: rk[u] / rl[u]
// = the tile's kept / lifted triplets before item u of this thread.
struct TileRanks {
  int rk[DIR_ITEMS], rl[DIR_ITEMS];
  __device__ __forceinline__ void compute(const TileItems &T) {
    __shared__ int wk[DIR_ITEMS][DIR_WAVES], wl[DIR_ITEMS][DIR_WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int u = 0; u < DIR_ITEMS; u++) {
      const unsigned long long mk = __ballot(T.keep[u]), ml = __ballot(T.lift[u]);
      rk[u] = __popcll(mk & below);
      rl[u] = __popcll(ml & below);
      if (lane == 0) {
        wk[u][wave] = __popcll(mk);
        wl[u][wave] = __popcll(ml);
      }
    }
    __syncthreads();
    int ak = 0, al = 0;  // kept / lifted before (u, wave) in input order
#pragma unroll
    for (int u = 0; u < DIR_ITEMS; u++) {
      int bk = 0, bl = 0;
      for (int w = 0; w < DIR_WAVES; w++) {
        if (w == wave) {
          bk = ak;
          bl = al;
        }
        ak += wk[u][w];
        al += wl[u][w];
      }
      rk[u] += bk;
      rl[u] += bl;
    }
  }
};

// pass 2, forward: kept triplets to their output slots (tile offset + rank),
// lifted triplets to the lift list as (row, triplet index)
__global__ void __launch_bounds__(DIR_NT)
    tile_place(const long long *__restrict__ ind, const double *__restrict__ vv, int64_t sN,
               int N, const int *__restrict__ bdidx, bool ind16,
               const int64_t *__restrict__ tkpos, const int64_t *__restrict__ tlpos,
               long long *__restrict__ oind, double *__restrict__ ov, bool oind16,
               int *__restrict__ lkey, int64_t *__restrict__ lval, int *__restrict__ err) {
  TileItems T;
  T.load(ind, sN, N, bdidx, ind16, err);
  TileRanks R;
  R.compute(T);
  const int64_t base = (int64_t)blockIdx.x * DIR_TILE + threadIdx.x;
  const int64_t k0 = tkpos[blockIdx.x], l0 = tlpos[blockIdx.x];
#pragma unroll
  for (int u = 0; u < DIR_ITEMS; u++) {
    const int64_t k = base + (int64_t)u * DIR_NT;
    if (T.keep[u]) store_triplet(oind, ov, k0 + R.rk[u], T.i[u], T.j[u], vv[k], oind16);
    if (T.lift[u]) {
      lkey[l0 + R.rl[u]] = T.i[u];
      lval[l0 + R.rl[u]] = k;
    }
  }
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

// pass 2, backward: grad_vv (a kept triplet's is grad_ov at its output slot;
// a lifted one's also gets -bdval[j] grad_orhs[i]) and the lift list as
// (column, triplet index)
__global__ void __launch_bounds__(DIR_NT)
    tile_grad(const long long *__restrict__ ind, int64_t sN, int N,
              const int *__restrict__ bdidx, bool ind16, const int64_t *__restrict__ tkpos,
              const int64_t *__restrict__ tlpos, const double *__restrict__ bdval,
              const double *__restrict__ gov, const double *__restrict__ gorhs,
              double *__restrict__ gvv, int *__restrict__ lkey, int64_t *__restrict__ lval,
              int *__restrict__ err) {
  TileItems T;
  T.load(ind, sN, N, bdidx, ind16, err);
  TileRanks R;
  R.compute(T);
  const int64_t base = (int64_t)blockIdx.x * DIR_TILE + threadIdx.x;
  const int64_t k0 = tkpos[blockIdx.x], l0 = tlpos[blockIdx.x];
#pragma unroll
  for (int u = 0; u < DIR_ITEMS; u++) {
    const int64_t k = base + (int64_t)u * DIR_NT;
    if (k >= sN) continue;
    double g = T.keep[u] ? gov[k0 + R.rk[u]] : 0.0;
    if (T.lift[u]) {
      g -= bdval[bdidx[T.j[u]]] * gorhs[T.i[u]];
      lkey[l0 + R.rl[u]] = T.j[u];
      lval[l0 + R.rl[u]] = k;
    }
    gvv[k] = g;
  }
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
  DevBuf<int> keys, win, bdidx, tkeep, tlift, err, lkey, skey;
  DevBuf<int64_t> tkpos, tlpos, lval, sval;
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
  int nu = 0, ntiles = 0;
  bool ind16 = false;
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
  d.ntiles = (int)((sN + DIR_TILE - 1) / DIR_TILE);
  d.ind16 = ((uintptr_t)ind & 15) == 0;
  const int nt = d.ntiles;
  ensure(w.tkeep, nt + 1);
  ensure(w.tlift, nt + 1);
  ensure(w.tkpos, nt + 1);
  ensure(w.tlpos, nt + 1);
  ensure(w.err, 1);
  set_i<<<1, 1, 0, st>>>(w.err.p, 0);
  set_i<<<1, 1, 0, st>>>(w.tkeep.p + nt, 0);  // the scans' last entry: their totals
  set_i<<<1, 1, 0, st>>>(w.tlift.p + nt, 0);
  if (nt)
    tile_count<<<nt, DIR_NT, 0, st>>>(ind, sN, N, w.bdidx.p, d.ind16, w.tkeep.p, w.tlift.p,
                                      w.err.p);
  HIPCHECK(hipGetLastError());
  exclusive_scan(w, w.tkeep.p, w.tkpos.p, nt + 1, st);
  exclusive_scan(w, w.tlift.p, w.tlpos.p, nt + 1, st);
  int herr = 0;
  HIPCHECK(hipMemcpyAsync(&herr, w.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(&d.nkept, w.tkpos.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipMemcpyAsync(&d.nlift, w.tlpos.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  REQUIRE(!herr, ADMFEM_ERR_ARG, "triplet index out of range [0, N)");
}

// room for the lift list (written by tile_place / tile_grad into w.lkey /
// w.lval in input order)
void lift_room(const DirPrep &d) {
  DirWork &w = d.w;
  const int64_t nl = d.nlift;
  ensure(w.lkey, std::max<int64_t>(nl, 1));
  ensure(w.lval, std::max<int64_t>(nl, 1));
  ensure(w.skey, std::max<int64_t>(nl, 1));
  ensure(w.sval, std::max<int64_t>(nl, 1));
}

// the lift list stably sorted by key, into w.skey / w.sval
void sort_lift(const DirPrep &d, hipStream_t st) {
  DirWork &w = d.w;
  const int64_t nl = d.nlift;
  if (nl == 0) return;
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
  lift_room(d);
  if (d.ntiles)
    tile_place<<<d.ntiles, DIR_NT, 0, st>>>(ind, vv, sN, N, w.bdidx.p, d.ind16, w.tkpos.p,
                                            w.tlpos.p, oind, ov, ((uintptr_t)oind & 15) == 0,
                                            w.lkey.p, w.lval.p, w.err.p);
  if (d.nu) append_diag<<<nblk(d.nu, 256), 256, 0, st>>>(w.keys.p, d.nu, d.nkept, oind, ov);
  copy_rhs<<<nblk(N, 256), 256, 0, st>>>(rhs, orhs, N);
  sort_lift(d, st);
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
  lift_room(d);
  if (d.ntiles)
    tile_grad<<<d.ntiles, DIR_NT, 0, st>>>(ind, sN, N, w.bdidx.p, d.ind16, w.tkpos.p,
                                           w.tlpos.p, bdval, gov, gorhs, gvv, w.lkey.p,
                                           w.lval.p, w.err.p);
  grad_rhs_k<<<nblk(N, 256), 256, 0, st>>>(w.bdidx.p, gorhs, grhs, N);
  if (bdN) zero_d<<<nblk(bdN, 256), 256, 0, st>>>(gbd, bdN);
  sort_lift(d, st);
  if (d.nlift)
    grad_bdval_k<<<nblk(d.nlift, 256), 256, 0, st>>>(w.skey.p, w.sval.p, d.nlift, ind, vv,
                                                     w.bdidx.p, gorhs, gbd);
  if (d.nu)
    grad_bdval_diag<<<nblk(d.nu, 256), 256, 0, st>>>(w.keys.p, w.win.p, d.nu, gorhs, gbd);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace admfem
