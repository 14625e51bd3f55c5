// Mesh topology on the device: the internal node order, the connectivity in
// that order, node -> element incidences, the pre-Dirichlet pattern and the
// per-incidence slot words -- what deps/MFEM/Common.cpp:20-142 and the mesh
// preprocessing of src/MFEM/MFEM.jl:60-119 leave to be derived for a
// solver -- built with radix sorts, scans and one thread per node, element,
// row or incidence.  The host keeps the node order (perm / iperm); the rest
// stays on the device (the pattern getter copies the pattern back,
// mesh_user_pattern).
//
// Node order: the order of first appearance walking the elements and their
// vertices (nodes of no element last, ascending) when that order is banded
// -- bandwidth at most 5x the widest level of a breadth-first search from
// the first element's first vertex -- else reverse Cuthill-McKee over the
// node graph (rcm: level by level on the device).  Rows then run along fronts: the k-th
// neighbours of consecutive rows are consecutive nodes of the adjacent
// fronts, so the solver's column gathers coalesce within a wavefront, and
// the elements of a block of rows are consecutive in the user's element
// order.
//
// Integer atomics appear only where their result does not depend on the
// order in which they land: atomicMin of first appearances, atomicMax of the
// bandwidth, and the breadth-first search's claims and level counts (which
// node claims a neighbour first varies, the level of every node and each
// level's size do not).
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdlib>

#include "internal.h"

namespace admfem {

std::vector<int> rcm_order(int n, const std::vector<int> &rp, const std::vector<int> &col);

namespace {

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int key_bits(uint64_t maxkey) {
  int b = 1;
  while (b < 64 && (maxkey >> b)) b++;
  return b;
}

template <class K, class V>
void sort_pairs(const K *kin, K *kout, const V *vin, V *vout, int n, int bits, hipStream_t st) {
  size_t tmp = 0;
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, n, 0, bits,
                                              st));
  DevBuf<char> t(std::max<size_t>(tmp, 1));
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, kin, kout, vin, vout, n, 0, bits, st));
  HIPCHECK(hipStreamSynchronize(st));  // before the temporary storage is released
}

void exclusive_sum(const int *in, int *out, int n, hipStream_t st) {
  size_t tmp = 0;
  HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, st));
  DevBuf<char> t(std::max<size_t>(tmp, 1));
  HIPCHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, n, st));
  HIPCHECK(hipStreamSynchronize(st));
}

__global__ void fill_int(int *a, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void iota_int(int *a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (int)i;
}

// key[v] = first position 3e + k at which node v appears
__global__ void first_seen_kernel(int m3, const int *__restrict__ elems, int *__restrict__ key) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m3) atomicMin(&key[elems[t]], t);
}

// perm[iperm[k]] = k
__global__ void invert_kernel(int n, const int *__restrict__ iperm, int *__restrict__ perm) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) perm[iperm[k]] = k;
}

// internal coordinates: xy[r] = coords[iperm[r]]
__global__ void coords_kernel(int n, const int *__restrict__ iperm,
                              const double2 *__restrict__ cu, double *__restrict__ x,
                              double *__restrict__ y, double2 *__restrict__ xy) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double2 c = cu[iperm[r]];
  x[r] = c.x;
  y[r] = c.y;
  xy[r] = c;
}

// connectivity in SoA form, internal and user ids, and each (element,
// vertex) pair's internal row (AoS position 3e + k)
__global__ void conn_kernel(int ne, const int *__restrict__ elems, const int *__restrict__ perm,
                            int *__restrict__ conn, int *__restrict__ conn_user,
                            int *__restrict__ rowkey) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int u = elems[3 * (size_t)e + k], r = perm[u];
    conn[(size_t)k * ne + e] = r;
    conn_user[(size_t)k * ne + e] = u;
    rowkey[3 * (size_t)e + k] = r;
  }
}

// ptr[r] = first position of row r in the sorted row keys (empty rows too)
__global__ void run_start_kernel(int m, int n, const int *__restrict__ srow, int *__restrict__ ptr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > m) return;
  const int prev = i == 0 ? -1 : srow[i - 1];
  const int cur = i == m ? n : srow[i];
  for (int r = prev + 1; r <= cur; r++) ptr[r] = i;
}

__global__ void div3_kernel(int m, const int *__restrict__ t, int *__restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) e[i] = t[i] / 3;
}

// The sorted distinct vertices of row r's elements, in c[j * PAT_NT] (the
// calling thread's column of a block-wide LDS array); -1 when more than
// MAX_ROW.
constexpr int PAT_NT = 64;
__device__ int row_union(int r, const int *__restrict__ ptr, const int *__restrict__ inc_elem,
                         const int *__restrict__ conn, int ne, int *c) {
  int len = 0;
  for (int t = ptr[r]; t < ptr[r + 1]; t++) {
    const int e = inc_elem[t];
    for (int q = 0; q < 3; q++) {
      const int v = conn[(size_t)q * ne + e];
      int pos = 0;
      while (pos < len && c[pos * PAT_NT] < v) pos++;
      if (pos < len && c[pos * PAT_NT] == v) continue;
      if (len == MAX_ROW) return -1;
      for (int j = len; j > pos; j--) c[j * PAT_NT] = c[(j - 1) * PAT_NT];
      c[pos * PAT_NT] = v;
      len++;
    }
  }
  return len;
}

__global__ void __launch_bounds__(PAT_NT)
    pattern_count_kernel(int n, const int *__restrict__ ptr, const int *__restrict__ inc_elem,
                         const int *__restrict__ conn, int ne, int *__restrict__ cnt,
                         int *__restrict__ err) {
  __shared__ int lc[MAX_ROW * PAT_NT];
  const int r = blockIdx.x * PAT_NT + threadIdx.x;
  if (r >= n) return;
  const int len = row_union(r, ptr, inc_elem, conn, ne, lc + threadIdx.x);
  if (len < 0) *err = 1;
  cnt[r] = len < 0 ? 0 : len;
}

__global__ void __launch_bounds__(PAT_NT)
    pattern_fill_kernel(int n, const int *__restrict__ ptr, const int *__restrict__ inc_elem,
                        const int *__restrict__ conn, int ne, const int *__restrict__ rp,
                        int *__restrict__ col) {
  __shared__ int lc[MAX_ROW * PAT_NT];
  const int r = blockIdx.x * PAT_NT + threadIdx.x;
  if (r >= n) return;
  const int len = row_union(r, ptr, inc_elem, conn, ne, lc + threadIdx.x);
  for (int j = 0; j < len; j++) col[rp[r] + j] = lc[j * PAT_NT + threadIdx.x];
}

// slot word of incidence i (row srow[i], element/vertex tv[i] = 3e + p):
// p, then the position of each of the element's vertices in the row
__global__ void slots_kernel(int m, int ne, const int *__restrict__ srow,
                             const int *__restrict__ tv, const int *__restrict__ conn,
                             const int *__restrict__ rp, const int *__restrict__ col,
                             uint32_t *__restrict__ slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int r = srow[i], t = tv[i], e = t / 3, p = t - 3 * e;
  const int *c = col + rp[r];
  const int len = rp[r + 1] - rp[r];
  uint32_t w = (uint32_t)p;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int v = conn[(size_t)q * ne + e];
    int lo = 0, hi = len;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (c[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    w |= (uint32_t)lo << (2 + SLOT_BITS * q);
  }
  slots[i] = w;
}

// max |pos(r) - pos(c)| over the pattern's entries (pos = inv, or the
// identity when inv is null)
__global__ void bandwidth_kernel(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                 const int *__restrict__ inv, int *__restrict__ bw) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int pr = inv ? inv[r] : r;
  int d = 0;
  for (int k = rp[r]; k < rp[r + 1]; k++) d = max(d, abs((inv ? inv[col[k]] : col[k]) - pr));
  if (d > 0) atomicMax(bw, d);
}

// off-diagonal entries per row
__global__ void degree_kernel(int n, const int *__restrict__ rp, const int *__restrict__ col,
                              int *__restrict__ deg) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int d = 0;
  for (int k = rp[r]; k < rp[r + 1]; k++) d += col[k] != r;
  deg[r] = d;
}

// smallest (degree, id) over the rows with an off-diagonal entry, taken
// from the listed rows (list null: all rows), as degree << 32 | id
__global__ void min_degree_kernel(int m, const int *__restrict__ list, const int *__restrict__ deg,
                                  unsigned long long *__restrict__ best) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int v = list ? list[i] : i;
  if (deg[v] > 0)
    atomicMin(best, (unsigned long long)deg[v] << 32 | (unsigned long long)(unsigned)v);
}

__global__ void first_coupled_kernel(int n, const int *__restrict__ deg, int *__restrict__ first) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && deg[r] > 0) atomicMin(first, r);
}

// One level of a breadth-first search: the frontier is q[off[L], off[L] +
// cnt[L]); its unvisited neighbours are claimed for level L + 1 and appended
// after it.  off[L + 1] is written for the next level's launch.
constexpr int BFS_NT = 256, BFS_BLOCKS = 64;
__global__ void __launch_bounds__(BFS_NT)
    bfs_level_kernel(int L, const int *__restrict__ rp, const int *__restrict__ col,
                     int *__restrict__ level, int *__restrict__ q, int *__restrict__ off,
                     int *__restrict__ cnt) {
  const int o = off[L], c = cnt[L];
  if (blockIdx.x == 0 && threadIdx.x == 0) off[L + 1] = o + c;
  for (int i = blockIdx.x * BFS_NT + threadIdx.x; i < c; i += BFS_BLOCKS * BFS_NT) {
    const int u = q[o + i];
    for (int k = rp[u]; k < rp[u + 1]; k++) {
      const int v = col[k];
      if (level[v] < 0 && atomicCAS(&level[v], -1, L + 1) == -1)
        q[o + c + atomicAdd(&cnt[L + 1], 1)] = v;
    }
  }
}

// Cuthill-McKee positions level by level, in one workgroup: a node of level
// L + 1 follows its parent -- its neighbour of level L placed first -- and
// the parent's children are placed in (degree, id) order, which is the
// order a breadth-first search visiting each node's unvisited neighbours
// sorted by (degree, id) produces.  Level L of the search occupies
// positions [off[L], off[L] + cnt[L]).  On entry pos holds INT_MAX, ord -1
// and ccnt zeros (left zero).  A pattern that is not symmetric raises *bad
// (every index stays in bounds).
constexpr int CM_NT = 1024;
__global__ void __launch_bounds__(CM_NT)
    cm_kernel(int nlev, int start, const int *__restrict__ rp, const int *__restrict__ col,
              const int *__restrict__ deg, const int *__restrict__ level,
              const int *__restrict__ q, const int *__restrict__ off,
              const int *__restrict__ cnt, int *__restrict__ pos, int *__restrict__ ord,
              int *__restrict__ ppos, int *__restrict__ ccnt, int *__restrict__ coff,
              int *__restrict__ bad) {
  __shared__ int part[CM_NT];
  const int tid = threadIdx.x;
  if (tid == 0) {
    pos[start] = 0;
    ord[0] = start;
  }
  __threadfence();
  __syncthreads();
  for (int L = 0; L + 1 < nlev; L++) {
    const int oL = off[L], cL = cnt[L], o1 = off[L + 1], c1 = cnt[L + 1];
    for (int i = tid; i < c1; i += CM_NT) {  // parents
      const int v = q[o1 + i];
      int best = INT_MAX;
      for (int k = rp[v]; k < rp[v + 1]; k++) {
        const int u = col[k];
        if (level[u] == L) best = min(best, pos[u]);
      }
      if (best < oL || best >= oL + cL) {  // the pattern is not symmetric
        *bad = 1;
        best = INT_MAX;
      } else {
        atomicAdd(&ccnt[best - oL], 1);
      }
      ppos[v] = best;
    }
    __threadfence();
    __syncthreads();
    // children offsets: exclusive scan of ccnt over level L's positions
    const int per = (cL + CM_NT - 1) / CM_NT;
    const int b = tid * per, e = min(b + per, cL);
    int sum = 0;
    for (int i = b; i < e; i++) sum += ccnt[i];
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < CM_NT; d <<= 1) {
      const int v = tid >= d ? part[tid - d] : 0;
      __syncthreads();
      part[tid] += v;
      __syncthreads();
    }
    int run = part[tid] - sum;
    for (int i = b; i < e; i++) {
      coff[i] = run;
      run += ccnt[i];
      ccnt[i] = 0;
    }
    __threadfence();
    __syncthreads();
    for (int i = tid; i < c1; i += CM_NT) {  // places
      const int v = q[o1 + i];
      const int pp = ppos[v];
      if (pp == INT_MAX) continue;
      const int par = ord[pp];
      if (par < 0) {
        *bad = 1;
        continue;
      }
      int rank = 0;
      for (int k = rp[par]; k < rp[par + 1]; k++) {
        const int w = col[k];
        if (w != v && level[w] == L + 1 && ppos[w] == pp &&
            (deg[w] < deg[v] || (deg[w] == deg[v] && w < v)))
          rank++;
      }
      const int ps = o1 + coff[pp - oL] + rank;
      if (ps >= o1 + c1) {  // not symmetric: the siblings do not add up
        *bad = 1;
        continue;
      }
      pos[v] = ps;
      ord[ps] = v;
    }
    __threadfence();
    __syncthreads();
  }
}

// The final order: nodes without off-diagonal entries first (ascending),
// then the component reversed (reverse Cuthill-McKee); order[k] = node at
// position k, inv its inverse.
__global__ void rcm_place_kernel(int n, int ncomp, const int *__restrict__ deg,
                                 const int *__restrict__ iso_rank, const int *__restrict__ pos,
                                 int *__restrict__ order, int *__restrict__ inv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int niso = n - ncomp;
  const int k = deg[v] == 0 ? iso_rank[v] : niso + (ncomp - 1 - pos[v]);
  order[k] = v;
  inv[v] = k;
}

__global__ void isolated_flag_kernel(int n, const int *__restrict__ deg, int *__restrict__ f) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) f[v] = deg[v] == 0 ? 1 : 0;
}

// ip[k] = iperm[ord[k]]
__global__ void compose_kernel(int n, const int *__restrict__ iperm, const int *__restrict__ ord,
                               int *__restrict__ ip) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) ip[k] = iperm[ord[k]];
}

// flag when order / inv is not a permutation and its inverse
__global__ void perm_check_kernel(int n, const int *__restrict__ order,
                                  const int *__restrict__ inv, int *__restrict__ bad) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int k = inv[v];
  if (k < 0 || k >= n || order[k] != v) *bad = 1;
}

// Everything that follows from the node order iperm (device, n entries).
void build_order_dependent(Mesh &m, const DevBuf<int> &elems, const DevBuf<double2> &cu,
                           DevBuf<int> &iperm, hipStream_t st) {
  const int n = m.nnode, ne = m.nelem, m3 = 3 * ne;
  m.d_perm.alloc(n);
  invert_kernel<<<nb(n, 256), 256, 0, st>>>(n, iperm.p, m.d_perm.p);
  m.conn.alloc(3 * (size_t)ne);
  m.conn_user.alloc(3 * (size_t)ne);
  DevBuf<int> rowkey(m3), srow(m3), tval(m3), sval(m3);
  conn_kernel<<<nb(ne, 256), 256, 0, st>>>(ne, elems.p, m.d_perm.p, m.conn.p, m.conn_user.p,
                                           rowkey.p);
  iota_int<<<nb(m3, 256), 256, 0, st>>>(tval.p, m3);
  HIPCHECK(hipGetLastError());
  // incidences: stable sort by row keeps each row's elements ascending
  sort_pairs(rowkey.p, srow.p, tval.p, sval.p, m3, key_bits((uint64_t)n), st);
  m.d_n2e_ptr.alloc(n + 1);
  m.d_n2e_elem.alloc(m3);
  run_start_kernel<<<nb(m3 + 1, 256), 256, 0, st>>>(m3, n, srow.p, m.d_n2e_ptr.p);
  div3_kernel<<<nb(m3, 256), 256, 0, st>>>(m3, sval.p, m.d_n2e_elem.p);
  // pattern: per row the distinct vertices of its elements
  DevBuf<int> cnt(n + 1), err(1);
  fill_int<<<1, 1, 0, st>>>(err.p, 1, 0);
  fill_int<<<1, 1, 0, st>>>(cnt.p + n, 1, 0);
  pattern_count_kernel<<<nb(n, PAT_NT), PAT_NT, 0, st>>>(n, m.d_n2e_ptr.p, m.d_n2e_elem.p,
                                                         m.conn.p, ne, cnt.p, err.p);
  HIPCHECK(hipGetLastError());
  m.d_rp.alloc(n + 1);
  exclusive_sum(cnt.p, m.d_rp.p, n + 1, st);
  int h[2] = {0, 0};
  HIPCHECK(hipMemcpy(&h[0], err.p, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&h[1], m.d_rp.p + n, sizeof(int), hipMemcpyDeviceToHost));
  REQUIRE(!h[0], ADMFEM_ERR_UNSUPPORTED, "a node has more than 61 neighbours");
  m.d_col.alloc(std::max(h[1], 1));
  pattern_fill_kernel<<<nb(n, PAT_NT), PAT_NT, 0, st>>>(n, m.d_n2e_ptr.p, m.d_n2e_elem.p,
                                                        m.conn.p, ne, m.d_rp.p, m.d_col.p);
  m.d_slots_full.alloc(m3);
  slots_kernel<<<nb(m3, 256), 256, 0, st>>>(m3, ne, srow.p, sval.p, m.conn.p, m.d_rp.p,
                                            m.d_col.p, m.d_slots_full.p);
  m.x.alloc(n);
  m.y.alloc(n);
  m.xy.alloc(n);
  coords_kernel<<<nb(n, 256), 256, 0, st>>>(n, iperm.p, cu.p, m.x.p, m.y.p, m.xy.p);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
  m.nnz_full = h[1];
}

// ---------------------------------------------------------------- edges
// Element edge t = 3e + j joins vertices (j, j+1 mod 3): key min * n + max
// of its user ids (degenerate edges: n * n, after every real key).
__global__ void edge_key_kernel(int ne, int n, const int *__restrict__ conn_user,
                                unsigned long long *__restrict__ key) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int a = conn_user[(size_t)j * ne + e], b = conn_user[(size_t)((j + 1) % 3) * ne + e];
    key[3 * (size_t)e + j] =
        a == b ? (unsigned long long)n * n
               : (unsigned long long)min(a, b) * n + (unsigned long long)max(a, b);
  }
}

// fresh[i]: sorted position i starts a run of equal (real) keys; nvalid =
// the number of real keys
__global__ void edge_fresh_kernel(int m, unsigned long long bad,
                                  const unsigned long long *__restrict__ skey,
                                  int *__restrict__ fresh, int *__restrict__ nvalid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const bool real = skey[i] != bad;
  fresh[i] = real && (i == 0 || skey[i] != skey[i - 1]) ? 1 : 0;
  if (real && (i + 1 == m || skey[i + 1] == bad)) *nvalid = i + 1;
}

// per run k (starting at sorted position i): its start, and a flag at the
// element edge where the edge first appears (the run's smallest t: the sort
// is stable over ascending t)
__global__ void edge_runs_kernel(int m, const int *__restrict__ fresh,
                                 const int *__restrict__ run, const int *__restrict__ sv,
                                 int *__restrict__ start, int *__restrict__ first_t,
                                 int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !fresh[i]) return;
  const int k = run[i];
  start[k] = i;
  first_t[k] = sv[i];
  flag[sv[i]] = 1;
}

// per run: the edge's MFEM id (rank of its first appearance), its 1-based
// end points, and the boundary flags of an edge used by an odd number of
// elements and of its end points (plain stores of 1)
__global__ void edge_out_kernel(int nruns, int n, const int *__restrict__ start,
                                const int *__restrict__ first_t, const int *__restrict__ rank,
                                const unsigned long long *__restrict__ skey,
                                long long *__restrict__ edges, char *__restrict__ bd_edge,
                                char *__restrict__ bd_node, int *__restrict__ run_id) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nruns) return;
  const int id = rank[first_t[k]];
  run_id[k] = id;
  const unsigned long long key = skey[start[k]];
  const int a = (int)(key / (unsigned long long)n), b = (int)(key % (unsigned long long)n);
  edges[id] = a + 1;
  edges[(size_t)nruns + id] = b + 1;
  const bool odd = ((start[k + 1] - start[k]) & 1) != 0;
  bd_edge[id] = odd ? 1 : 0;
  if (odd) {
    bd_node[a] = 1;
    bd_node[b] = 1;
  }
}

// elem_edge[t] = MFEM id of element edge t (-1: degenerate)
__global__ void elem_edge_kernel(int m, const int *__restrict__ nvalid,
                                 const int *__restrict__ fresh, const int *__restrict__ run,
                                 const int *__restrict__ run_id, const int *__restrict__ sv,
                                 int *__restrict__ elem_edge) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  elem_edge[sv[i]] = i < *nvalid ? run_id[run[i] + fresh[i] - 1] : -1;
}

// ------------------------------------------------------------- systems
// post-Dirichlet row lengths: a boundary row keeps its diagonal, an
// interior row its interior columns (ImposeDirichlet.h:27-51)
__global__ void post_count_kernel(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                  const int *__restrict__ row_bd, int *__restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int c = 1;
  if (row_bd[r] < 0) {
    c = 0;
    for (int k = rp[r]; k < rp[r + 1]; k++) c += row_bd[col[k]] < 0;
  }
  cnt[r] = c;
}

__global__ void post_fill_kernel(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                 const int *__restrict__ row_bd, const int *__restrict__ prp,
                                 int *__restrict__ pcol) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int o = prp[r];
  if (row_bd[r] >= 0) {
    pcol[o] = r;
    return;
  }
  for (int k = rp[r]; k < rp[r + 1]; k++)
    if (row_bd[col[k]] < 0) pcol[o++] = col[k];
}

// SELL-64: slice size = 64 x the widest row's off-diagonal count
__global__ void slice_size_kernel(int n, int nslices, const int *__restrict__ rp,
                                  const int *__restrict__ col, int *__restrict__ size) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl > nslices) return;
  int w = 0;
  if (sl < nslices)
    for (int r = sl * 64; r < min(sl * 64 + 64, n); r++) {
      int c = 0;
      for (int k = rp[r]; k < rp[r + 1]; k++) c += col[k] != r;
      w = max(w, c);
    }
  size[sl] = 64 * w;
}

// off-diagonal columns, column-major inside the slice; padding points at the
// row itself (padding rows: 0)
__global__ void sell_col_kernel(int n, int npad, const int *__restrict__ rp,
                                const int *__restrict__ col, const int *__restrict__ slice_ptr,
                                int *__restrict__ sell_col) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npad) return;
  const int off = slice_ptr[r >> 6] + (r & 63);
  const int width = (slice_ptr[(r >> 6) + 1] - slice_ptr[r >> 6]) >> 6;
  int kk = 0;
  if (r < n)
    for (int k = rp[r]; k < rp[r + 1]; k++)
      if (col[k] != r) sell_col[off + 64 * kk++] = col[k];
  for (; kk < width; kk++) sell_col[off + 64 * kk] = r < n ? r : 0;
}

// SELL columns as 16-bit offsets from the row; *bad = 1 when one does not
// fit (padding rows past n read their own entry: offset 0)
__global__ void sell_d16_kernel(int n, int npad, const int *__restrict__ slice_ptr,
                                const int *__restrict__ sell_col, int16_t *__restrict__ d16,
                                int *__restrict__ bad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= npad) return;
  const int off = slice_ptr[r >> 6] + (r & 63);
  const int width = (slice_ptr[(r >> 6) + 1] - slice_ptr[r >> 6]) >> 6;
  for (int k = 0; k < width; k++) {
    const int d = r < n ? sell_col[off + 64 * k] - r : 0;
    if (d < -32768 || d > 32767) *bad = 1;
    d16[off + 64 * k] = (int16_t)d;
  }
}

// Block-wide sort of a[0, cnt) in LDS (bitonic, padded to a power of two
// with INT_MAX; a holds room for it) and its distinct values: their count,
// written to out[0, count) only when count <= cap.  part: SB_NT + 1 ints.
constexpr int SB_NT = 256;
__device__ int sort_unique(int *a, int cnt, int *out, int cap, int *part) {
  int S = 1;
  while (S < cnt) S <<= 1;
  for (int i = cnt + (int)threadIdx.x; i < S; i += SB_NT) a[i] = INT_MAX;
  __syncthreads();
  for (int k = 2; k <= S; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < S; i += SB_NT) {
        const int l = i ^ j;
        if (l > i) {
          const int x = a[i], y = a[l];
          if ((x > y) == ((i & k) == 0)) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
  const int per = (cnt + SB_NT - 1) / SB_NT;
  const int b = (int)threadIdx.x * per, e = min(b + per, cnt);
  int c = 0;
  for (int i = b; i < e; i++) c += i == 0 || a[i] != a[i - 1];
  part[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int sum = 0;
    for (int t = 0; t < SB_NT; t++) {
      const int v = part[t];
      part[t] = sum;
      sum += v;
    }
    part[SB_NT] = sum;
  }
  __syncthreads();
  const int total = part[SB_NT];
  if (total <= cap) {
    int o = part[threadIdx.x];
    for (int i = b; i < e; i++)
      if (i == 0 || a[i] != a[i - 1]) out[o++] = a[i];
  }
  __syncthreads();
  return total;
}

__device__ __forceinline__ int lower_bound_int(const int *a, int len, int v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Assembly blocks (one 64-row SELL slice each): the distinct elements
// incident to the block's interior rows (ascending), the distinct vertices
// of those elements (ascending internal ids), each element's record
// {element, 3 x 10-bit local vertex index} and a pair word per incidence of
// an interior row (local vertex p, three column codes -- SELL position,
// PAIR_DIAG or SLOT_LIFT -- and the element's index in the block's list).
// COUNT: the list lengths; else the lists and words at the scanned offsets.
constexpr int BLK_CAP_A = 8192;  // LDS candidates: incidences, then 3 x elements
template <bool COUNT>
__global__ void __launch_bounds__(SB_NT)
    asm_blocks_kernel(int n, int ne_mesh, const int *__restrict__ n2e_ptr,
                      const int *__restrict__ n2e_elem, const uint32_t *__restrict__ slots,
                      const int *__restrict__ conn, const int *__restrict__ row_bd,
                      const int *__restrict__ prp, const int *__restrict__ pcol,
                      int *__restrict__ cnt_e, int *__restrict__ cnt_v,
                      const int *__restrict__ eptr, const int *__restrict__ vptr,
                      int2 *__restrict__ brec, int *__restrict__ bvert,
                      uint32_t *__restrict__ pairs, int *__restrict__ err) {
  __shared__ int A[BLK_CAP_A];
  __shared__ int E[PAIR_MAX_ELEMS];
  __shared__ int V[BLK_MAX_VERTS];
  __shared__ int part[SB_NT + 1];
  __shared__ int rofs[AS_ROWS + 1];
  const int b = blockIdx.x, r0 = b * AS_ROWS;
  const int tid = threadIdx.x;
  // the interior rows' incidences, row after row
  if (tid < AS_ROWS) {
    const int r = r0 + tid;
    rofs[tid + 1] = (r < n && row_bd[r] < 0) ? n2e_ptr[r + 1] - n2e_ptr[r] : 0;
  }
  __syncthreads();
  if (tid == 0) {
    rofs[0] = 0;
    for (int k = 0; k < AS_ROWS; k++) rofs[k + 1] += rofs[k];
  }
  __syncthreads();
  const int ninc = rofs[AS_ROWS];
  if (ninc > PAIR_MAX_ELEMS) {  // uniform over the block
    if (tid == 0) *err = 1;
    return;
  }
  if (tid < AS_ROWS && rofs[tid + 1] > rofs[tid]) {
    const int t0 = n2e_ptr[r0 + tid];
    for (int k = rofs[tid]; k < rofs[tid + 1]; k++) A[k] = n2e_elem[t0 + k - rofs[tid]];
  }
  __syncthreads();
  const int ne = sort_unique(A, ninc, E, PAIR_MAX_ELEMS, part);
  // the elements' vertices
  if (3 * ne > BLK_CAP_A) {
    if (tid == 0) *err = 2;
    return;
  }
  for (int i = tid; i < ne; i += SB_NT)
#pragma unroll
    for (int q = 0; q < 3; q++) A[3 * i + q] = conn[(size_t)q * ne_mesh + E[i]];
  __syncthreads();
  const int nv = sort_unique(A, 3 * ne, V, BLK_MAX_VERTS, part);
  if (nv > BLK_MAX_VERTS) {
    if (tid == 0) *err = 3;
    return;
  }
  if (COUNT) {
    if (tid == 0) {
      cnt_e[b] = ne;
      cnt_v[b] = nv;
    }
    return;
  }
  const int e0 = eptr[b], v0 = vptr[b];
  for (int i = tid; i < ne; i += SB_NT) {
    int loc = 0;
#pragma unroll
    for (int q = 0; q < 3; q++)
      loc |= lower_bound_int(V, nv, conn[(size_t)q * ne_mesh + E[i]]) << (BLK_VERT_BITS * q);
    brec[e0 + i] = make_int2(E[i], loc);
  }
  for (int i = tid; i < nv; i += SB_NT) bvert[v0 + i] = V[i];
  if (tid < AS_ROWS) {
    const int r = r0 + tid;
    if (r < n && row_bd[r] < 0) {
      const int *c = pcol + prp[r];
      const int len = prp[r + 1] - prp[r];
      for (int t = n2e_ptr[r]; t < n2e_ptr[r + 1]; t++) {
        const int e = n2e_elem[t];
        uint32_t w = (slots[t] & 3u) | ((uint32_t)lower_bound_int(E, ne, e) << 20);
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const int cq = conn[(size_t)q * ne_mesh + e];
          uint32_t code;
          if (row_bd[cq] >= 0) code = SLOT_LIFT;
          else if (cq == r) code = PAIR_DIAG;
          else {
            const int j = lower_bound_int(c, len, cq);  // the row holds its diagonal
            code = (uint32_t)(cq < r ? j : j - 1);
          }
          w |= code << (2 + SLOT_BITS * q);
        }
        pairs[t] = w;
      }
    }
  }
}

}  // namespace

void csr_row_ptr(const int *srow, int m, int n, int *rp, hipStream_t st) {
  run_start_kernel<<<nb((int64_t)m + 1, 256), 256, 0, st>>>(m, n, srow, rp);
  HIPCHECK(hipGetLastError());
}

int64_t device_bandwidth(int n, const int *rp, const int *col, const int *inv, hipStream_t st) {
  DevBuf<int> bw(1);
  fill_int<<<1, 1, 0, st>>>(bw.p, 1, 0);
  bandwidth_kernel<<<nb(n, 256), 256, 0, st>>>(n, rp, col, inv, bw.p);
  HIPCHECK(hipGetLastError());
  int h = 0;
  HIPCHECK(hipMemcpyAsync(&h, bw.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return h;
}

int device_first_coupled(int n, const int *rp, const int *col, hipStream_t st) {
  DevBuf<int> deg(n), first(1);
  degree_kernel<<<nb(n, 256), 256, 0, st>>>(n, rp, col, deg.p);
  fill_int<<<1, 1, 0, st>>>(first.p, 1, INT_MAX);
  first_coupled_kernel<<<nb(n, 256), 256, 0, st>>>(n, deg.p, first.p);
  HIPCHECK(hipGetLastError());
  int h = 0;
  HIPCHECK(hipMemcpyAsync(&h, first.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return h == INT_MAX ? 0 : h;
}

namespace {

// A breadth-first search from s: level of every node (-1: not reached), the
// visit queue with level L at [off[L], off[L] + cnt[L]), the level sizes.
struct BfsRun {
  DevBuf<int> level, q, off, cnt;
  std::vector<int> size;  // nodes per level, levels 0 .. nlev - 1
};

void run_bfs(int n, const int *rp, const int *col, int s, BfsRun &b, hipStream_t st) {
  b.level.alloc(n);
  b.q.alloc(n);
  b.off.alloc(n + 2);
  b.cnt.alloc(n + 2);
  fill_int<<<nb(n, 256), 256, 0, st>>>(b.level.p, n, -1);
  fill_int<<<nb(n + 2, 256), 256, 0, st>>>(b.cnt.p, n + 2, 0);
  fill_int<<<1, 1, 0, st>>>(b.level.p + s, 1, 0);
  fill_int<<<1, 1, 0, st>>>(b.q.p, 1, s);
  fill_int<<<1, 1, 0, st>>>(b.off.p, 1, 0);
  fill_int<<<1, 1, 0, st>>>(b.cnt.p, 1, 1);
  HIPCHECK(hipGetLastError());
  // levels in batches: a level with no nodes ends the search (every later
  // launch sees an empty frontier)
  constexpr int BATCH = 64;
  int L = 0;
  for (;;) {
    const int end = std::min(L + BATCH, n);
    for (; L < end; L++)
      bfs_level_kernel<<<BFS_BLOCKS, BFS_NT, 0, st>>>(L, rp, col, b.level.p, b.q.p, b.off.p,
                                                      b.cnt.p);
    HIPCHECK(hipGetLastError());
    int last = 0;
    HIPCHECK(hipMemcpyAsync(&last, b.cnt.p + L, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (last == 0 || L >= n) break;
  }
  std::vector<int> h(L + 1);
  HIPCHECK(hipMemcpy(h.data(), b.cnt.p, sizeof(int) * (L + 1), hipMemcpyDeviceToHost));
  b.size.clear();
  for (int v : h) {
    if (v == 0) break;
    b.size.push_back(v);
  }
}

unsigned long long min_degree(int m, const int *list, const int *deg, hipStream_t st) {
  DevBuf<unsigned long long> best(1);
  HIPCHECK(hipMemsetAsync(best.p, 0xff, sizeof(unsigned long long), st));
  if (m > 0) min_degree_kernel<<<nb(m, 256), 256, 0, st>>>(m, list, deg, best.p);
  HIPCHECK(hipGetLastError());
  unsigned long long h = 0;
  HIPCHECK(hipMemcpyAsync(&h, best.p, sizeof(h), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return h;
}

}  // namespace

int64_t device_bfs_max_level(int n, const int *rp, const int *col, int s, hipStream_t st) {
  BfsRun b;
  run_bfs(n, rp, col, s, b, st);
  int64_t w = 0;
  for (int v : b.size) w = std::max<int64_t>(w, v);
  return w;
}

// Reverse Cuthill-McKee order of the graph of a symmetric CSR pattern
// (device rp / col, diagonal entries ignored): nodes without neighbours
// first, ascending, then the rest in reverse Cuthill-McKee order from a
// pseudo-peripheral start (George & Liu: up to four searches, each from a
// least-degree node of the last one's deepest level, while the depth
// grows).  On the device when the rest is one connected component; else
// (or when the pattern turns out not to be symmetric) on the host
// (rcm_order).  order[k] = node at position k; inv its inverse.
void rcm(int n, const int *rp, const int *col, DevBuf<int> &order, DevBuf<int> &inv,
         hipStream_t st) {
  order.alloc(n);
  inv.alloc(n);
  DevBuf<int> deg(n), flag(n + 1), iso(n + 1), bad(1);
  degree_kernel<<<nb(n, 256), 256, 0, st>>>(n, rp, col, deg.p);
  isolated_flag_kernel<<<nb(n, 256), 256, 0, st>>>(n, deg.p, flag.p);
  fill_int<<<1, 1, 0, st>>>(flag.p + n, 1, 0);
  fill_int<<<1, 1, 0, st>>>(bad.p, 1, 0);
  HIPCHECK(hipGetLastError());
  exclusive_sum(flag.p, iso.p, n + 1, st);
  int niso = 0;
  HIPCHECK(hipMemcpy(&niso, iso.p + n, sizeof(int), hipMemcpyDeviceToHost));
  const int ncomp = n - niso;
  DevBuf<int> pos(n);
  bool ok = true;
  if (ncomp > 0) {
    int start = (int)(min_degree(n, nullptr, deg.p, st) & 0xffffffffu);
    BfsRun b;
    int from = -1, ecc = -1;
    for (int rep = 0; rep < 4; rep++) {
      run_bfs(n, rp, col, start, b, st);
      from = start;
      const int maxl = (int)b.size.size() - 1;
      int last_off = 0;
      for (int L = 0; L < maxl; L++) last_off += b.size[L];
      const int cand = (int)(min_degree(b.size[maxl], b.q.p + last_off, deg.p, st) & 0xffffffffu);
      if (maxl <= ecc) break;
      ecc = maxl;
      start = cand;
    }
    if (from != start) run_bfs(n, rp, col, start, b, st);
    int reached = 0;
    for (int v : b.size) reached += v;
    ok = reached == ncomp;
    if (ok) {
      DevBuf<int> ord(n), ppos(n), ccnt(n), coff(n);
      fill_int<<<nb(n, 256), 256, 0, st>>>(ccnt.p, n, 0);
      fill_int<<<nb(n, 256), 256, 0, st>>>(pos.p, n, INT_MAX);
      fill_int<<<nb(n, 256), 256, 0, st>>>(ord.p, n, -1);
      cm_kernel<<<1, CM_NT, 0, st>>>((int)b.size.size(), start, rp, col, deg.p, b.level.p, b.q.p,
                                     b.off.p, b.cnt.p, pos.p, ord.p, ppos.p, ccnt.p, coff.p,
                                     bad.p);
      HIPCHECK(hipGetLastError());
      int hb = 0;
      HIPCHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      ok = hb == 0;
    }
  }
  if (ok) {
    rcm_place_kernel<<<nb(n, 256), 256, 0, st>>>(n, ncomp, deg.p, iso.p, pos.p, order.p, inv.p);
    perm_check_kernel<<<nb(n, 256), 256, 0, st>>>(n, order.p, inv.p, bad.p);
    HIPCHECK(hipGetLastError());
    int hb = 0;
    HIPCHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    ok = hb == 0;
  }
  if (ok) return;
  // host: several components (or a pattern that is not symmetric)
  int nnz = 0;
  HIPCHECK(hipMemcpy(&nnz, rp + n, sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> hrp(n + 1), hcol(nnz), hinv(n);
  HIPCHECK(hipMemcpy(hrp.data(), rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(hcol.data(), col, sizeof(int) * nnz, hipMemcpyDeviceToHost));
  const std::vector<int> ho = rcm_order(n, hrp, hcol);
  for (int k = 0; k < n; k++) hinv[ho[k]] = k;
  order.upload(ho);
  inv.upload(hinv);
}

void mesh_topology(Mesh &m, const double *coords, hipStream_t st) {
  const int n = m.nnode, ne = m.nelem;
  REQUIRE((int64_t)3 * ne < (int64_t)1 << 30, ADMFEM_ERR_UNSUPPORTED, "too many elements");
  const int m3 = 3 * ne;
  DevBuf<int> elems, key(n), skey(n), ids(n), iperm(n);
  elems.upload(m.elems);
  DevBuf<double2> cu(n);
  HIPCHECK(hipMemcpy(cu.p, coords, sizeof(double2) * n, hipMemcpyHostToDevice));
  // element order: nodes sorted by first appearance (stable: unseen nodes,
  // all keyed m3, stay ascending at the end)
  fill_int<<<nb(n, 256), 256, 0, st>>>(key.p, n, m3);
  first_seen_kernel<<<nb(m3, 256), 256, 0, st>>>(m3, elems.p, key.p);
  iota_int<<<nb(n, 256), 256, 0, st>>>(ids.p, n);
  HIPCHECK(hipGetLastError());
  sort_pairs(key.p, skey.p, ids.p, iperm.p, n, key_bits((uint64_t)m3), st);
  build_order_dependent(m, elems, cu, iperm, st);
  // banded? (pattern under the element order; the search starts at the
  // first element's first vertex)
  int s = 0;
  HIPCHECK(hipMemcpy(&s, m.d_perm.p + m.elems[0], sizeof(int), hipMemcpyDeviceToHost));
  const int64_t hbw = device_bandwidth(n, m.d_rp.p, m.d_col.p, nullptr, st);
  if (hbw > 5 * device_bfs_max_level(n, m.d_rp.p, m.d_col.p, s, st)) {
    // reverse Cuthill-McKee over the pattern in the element order
    DevBuf<int> ord(n), inv(n), ip(n);
    rcm(n, m.d_rp.p, m.d_col.p, ord, inv, st);
    compose_kernel<<<nb(n, 256), 256, 0, st>>>(n, iperm.p, ord.p, ip.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    build_order_dependent(m, elems, cu, ip, st);
    iperm.swap(ip);
  }
  m.iperm.resize(n);
  m.perm.resize(n);
  HIPCHECK(hipMemcpy(m.iperm.data(), iperm.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(m.perm.data(), m.d_perm.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  m.d_iperm.alloc(n);
  HIPCHECK(hipMemcpy(m.d_iperm.p, iperm.p, sizeof(int) * n, hipMemcpyDeviceToDevice));
}

// MFEM edge table (Common.cpp:133-140, Mesh::GetVertexToVertexTable /
// DSTable::Push order): an edge gets the next id the first time it is met
// walking the oriented elements and their edges (0,1),(1,2),(2,0); the
// host receives the count, the boundary nodes (ends of edges used by one
// element, src/MFEM/MCore.jl:385-406 and 439-450: set toggling keeps odd
// counts) and, for P2, every element edge's id and the boundary edges; the
// table itself stays on the device until asked for (mesh_edges_host).
void mesh_edges(Mesh &m) {
  if (m.nedges >= 0) return;
  const int n = m.nnode, ne = m.nelem, m3 = 3 * ne;
  hipStream_t st = nullptr;
  const unsigned long long bad = (unsigned long long)n * (unsigned long long)n;
  DevBuf<unsigned long long> key(m3), skey(m3);
  DevBuf<int> tv(m3), sv(m3), fresh(m3), run(m3), nvalid(1);
  edge_key_kernel<<<nb(ne, 256), 256, 0, st>>>(ne, n, m.conn_user.p, key.p);
  iota_int<<<nb(m3, 256), 256, 0, st>>>(tv.p, m3);
  HIPCHECK(hipGetLastError());
  sort_pairs(key.p, skey.p, tv.p, sv.p, m3, key_bits(bad), st);
  key.release();
  fill_int<<<1, 1, 0, st>>>(nvalid.p, 1, 0);
  edge_fresh_kernel<<<nb(m3, 256), 256, 0, st>>>(m3, bad, skey.p, fresh.p, nvalid.p);
  HIPCHECK(hipGetLastError());
  exclusive_sum(fresh.p, run.p, m3, st);
  int h[3] = {0, 0, 0};
  HIPCHECK(hipMemcpy(&h[0], run.p + m3 - 1, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&h[1], fresh.p + m3 - 1, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&h[2], nvalid.p, sizeof(int), hipMemcpyDeviceToHost));
  const int nruns = h[0] + h[1];
  DevBuf<int> start(nruns + 1), first_t(std::max(nruns, 1)), flag(m3), rank(m3),
      run_id(std::max(nruns, 1));
  fill_int<<<nb(m3, 256), 256, 0, st>>>(flag.p, m3, 0);
  fill_int<<<1, 1, 0, st>>>(start.p + nruns, 1, h[2]);
  edge_runs_kernel<<<nb(m3, 256), 256, 0, st>>>(m3, fresh.p, run.p, sv.p, start.p, first_t.p,
                                                flag.p);
  HIPCHECK(hipGetLastError());
  exclusive_sum(flag.p, rank.p, m3, st);
  m.d_edges.alloc(2 * (size_t)std::max(nruns, 1));
  DevBuf<char> bd_edge(std::max(nruns, 1)), bd_node(n);
  HIPCHECK(hipMemset(bd_node.p, 0, n));
  edge_out_kernel<<<nb(nruns, 256), 256, 0, st>>>(nruns, n, start.p, first_t.p, rank.p, skey.p,
                                                  m.d_edges.p, bd_edge.p, bd_node.p, run_id.p);
  HIPCHECK(hipGetLastError());
  m.bd_node.resize(n);
  HIPCHECK(hipMemcpy(m.bd_node.data(), bd_node.p, n, hipMemcpyDeviceToHost));
  if (m.degree == 2) {
    DevBuf<int> ee(m3);
    elem_edge_kernel<<<nb(m3, 256), 256, 0, st>>>(m3, nvalid.p, fresh.p, run.p, run_id.p, sv.p,
                                                  ee.p);
    HIPCHECK(hipGetLastError());
    m.elem_edge.resize(m3);
    m.bd_edge.resize(nruns);
    HIPCHECK(hipMemcpy(m.elem_edge.data(), ee.p, sizeof(int) * m3, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(m.bd_edge.data(), bd_edge.p, nruns, hipMemcpyDeviceToHost));
  }
  m.nedges = nruns;
}

void mesh_edges_host(Mesh &m) {
  mesh_edges(m);
  if ((int64_t)m.edges.size() == 2 * m.nedges) return;
  m.edges.resize(2 * (size_t)m.nedges);
  if (m.nedges)
    HIPCHECK(hipMemcpy(m.edges.data(), m.d_edges.p, sizeof(long long) * m.edges.size(),
                       hipMemcpyDeviceToHost));
}

// SELL-64 layout of the off-diagonal entries of the system's CSR pattern
// (d_rp, d_col on the device; slices of 64 rows, column-major inside a
// slice, width = the slice's widest row; padding entries point at the row
// itself with value 0) and the device buffers of the assembled values and of
// the solver.  The host keeps the slice offsets.
void solver_layout(System *s) {
  hipStream_t st = nullptr;
  const int n = s->n;
  s->nslices = s->npad / 64;
  DevBuf<int> size(s->nslices + 1);
  slice_size_kernel<<<nb(s->nslices + 1, 256), 256, 0, st>>>(n, s->nslices, s->d_rp.p,
                                                             s->d_col.p, size.p);
  HIPCHECK(hipGetLastError());
  s->d_slice_ptr.alloc(s->nslices + 1);
  exclusive_sum(size.p, s->d_slice_ptr.p, s->nslices + 1, st);
  s->slice_ptr.resize(s->nslices + 1);
  HIPCHECK(hipMemcpy(s->slice_ptr.data(), s->d_slice_ptr.p, sizeof(int) * (s->nslices + 1),
                     hipMemcpyDeviceToHost));
  s->sell_wmax = 0;
  for (int sl = 0; sl < s->nslices; sl++)
    s->sell_wmax = std::max(s->sell_wmax, (s->slice_ptr[sl + 1] - s->slice_ptr[sl]) / 64);
  const size_t ns = std::max(1, s->slice_ptr.back());
  s->d_sell_col.alloc(ns);
  sell_col_kernel<<<nb(s->npad, 256), 256, 0, st>>>(n, s->npad, s->d_rp.p, s->d_col.p,
                                                     s->d_slice_ptr.p, s->d_sell_col.p);
  HIPCHECK(hipGetLastError());
  {  // 16-bit column offsets for the FP32 step kernels, when every one fits
    DevBuf<int> bad(1);
    fill_int<<<1, 1, 0, st>>>(bad.p, 1, 0);
    s->d_sell_d16.alloc(ns);
    sell_d16_kernel<<<nb(s->npad, 256), 256, 0, st>>>(n, s->npad, s->d_slice_ptr.p,
                                                      s->d_sell_col.p, s->d_sell_d16.p, bad.p);
    HIPCHECK(hipGetLastError());
    int hb = 0;
    HIPCHECK(hipMemcpy(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost));
    if (hb) {
      DevBuf<int16_t> none;
      s->d_sell_d16.swap(none);
    }
  }
  if (!s->mesh) s->val.alloc(std::max<int64_t>(1, s->nnz));
  s->rhs.alloc(s->npad);
  HIPCHECK(hipMemset(s->rhs.p, 0, s->rhs.bytes()));
  s->d_sell_val.alloc(ns);
  HIPCHECK(hipMemset(s->d_sell_val.p, 0, s->d_sell_val.bytes()));
  s->d_diag.alloc(s->npad);
  HIPCHECK(hipDeviceSynchronize());
}

// The P1 system on the device: post-Dirichlet pattern from the mesh's
// pre-BC pattern and s->row_bd, the solver layout, the assembly blocks.
void system_topology(System *s) {
  hipStream_t st = nullptr;
  const Mesh &m = *s->mesh;
  const int n = s->n, ne = m.nelem;
  s->d_row_bd.upload(s->row_bd);
  DevBuf<int> cnt(n + 1);
  fill_int<<<1, 1, 0, st>>>(cnt.p + n, 1, 0);
  post_count_kernel<<<nb(n, 256), 256, 0, st>>>(n, m.d_rp.p, m.d_col.p, s->d_row_bd.p, cnt.p);
  HIPCHECK(hipGetLastError());
  s->d_rp.alloc(n + 1);
  exclusive_sum(cnt.p, s->d_rp.p, n + 1, st);
  int nnz = 0;
  HIPCHECK(hipMemcpy(&nnz, s->d_rp.p + n, sizeof(int), hipMemcpyDeviceToHost));
  s->nnz = nnz;
  s->d_col.alloc(std::max(nnz, 1));
  post_fill_kernel<<<nb(n, 256), 256, 0, st>>>(n, m.d_rp.p, m.d_col.p, s->d_row_bd.p, s->d_rp.p,
                                               s->d_col.p);
  HIPCHECK(hipGetLastError());
  solver_layout(s);
  REQUIRE(s->sell_wmax < (int)PAIR_DIAG, ADMFEM_ERR_UNSUPPORTED, "row wider than 61");
  // assembly blocks: list lengths, their offsets, then the lists and words
  const int nblk = s->npad / AS_ROWS, m3 = 3 * ne;
  DevBuf<int> ce(nblk + 1), cv(nblk + 1), err(1);
  fill_int<<<1, 1, 0, st>>>(err.p, 1, 0);
  fill_int<<<1, 1, 0, st>>>(ce.p + nblk, 1, 0);
  fill_int<<<1, 1, 0, st>>>(cv.p + nblk, 1, 0);
  asm_blocks_kernel<true><<<nblk, SB_NT, 0, st>>>(
      n, ne, m.d_n2e_ptr.p, m.d_n2e_elem.p, m.d_slots_full.p, m.conn.p, s->d_row_bd.p,
      s->d_rp.p, s->d_col.p, ce.p, cv.p, nullptr, nullptr, nullptr, nullptr, nullptr, err.p);
  HIPCHECK(hipGetLastError());
  int herr = 0;
  HIPCHECK(hipMemcpy(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost));
  REQUIRE(herr == 0, ADMFEM_ERR_UNSUPPORTED,
          herr == 3 ? "assembly block touches more than 1024 vertices"
                    : "assembly block has too many elements");
  std::vector<int> he(nblk + 1), hv(nblk + 1);
  HIPCHECK(hipMemcpy(he.data(), ce.p, sizeof(int) * (nblk + 1), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(hv.data(), cv.p, sizeof(int) * (nblk + 1), hipMemcpyDeviceToHost));
  s->asm_max_e = *std::max_element(he.begin(), he.end());
  s->asm_max_v = *std::max_element(hv.begin(), hv.end());
  s->d_blk_eptr.alloc(nblk + 1);
  s->d_blk_vptr.alloc(nblk + 1);
  exclusive_sum(ce.p, s->d_blk_eptr.p, nblk + 1, st);
  exclusive_sum(cv.p, s->d_blk_vptr.p, nblk + 1, st);
  int tot[2] = {0, 0};
  HIPCHECK(hipMemcpy(&tot[0], s->d_blk_eptr.p + nblk, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&tot[1], s->d_blk_vptr.p + nblk, sizeof(int), hipMemcpyDeviceToHost));
  s->d_blk_rec.alloc(std::max(tot[0], 1));
  s->d_blk_vert.alloc(std::max(tot[1], 1));
  s->d_pairs.alloc(std::max(m3, 1));
  HIPCHECK(hipMemset(s->d_pairs.p, 0, s->d_pairs.bytes()));
  asm_blocks_kernel<false><<<nblk, SB_NT, 0, st>>>(
      n, ne, m.d_n2e_ptr.p, m.d_n2e_elem.p, m.d_slots_full.p, m.conn.p, s->d_row_bd.p,
      s->d_rp.p, s->d_col.p, nullptr, nullptr, s->d_blk_eptr.p, s->d_blk_vptr.p, s->d_blk_rec.p,
      s->d_blk_vert.p, s->d_pairs.p, err.p);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
}

// Host copy of a system's post-Dirichlet pattern (the getters' input).
void system_host_pattern(System &s) {
  if ((int64_t)s.col.size() == s.nnz && (int)s.rp.size() == s.n + 1) return;
  s.rp.resize(s.n + 1);
  s.col.resize(s.nnz);
  HIPCHECK(hipMemcpy(s.rp.data(), s.d_rp.p, sizeof(int) * (s.n + 1), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(s.col.data(), s.d_col.p, sizeof(int) * s.nnz, hipMemcpyDeviceToHost));
}

}  // namespace admfem
