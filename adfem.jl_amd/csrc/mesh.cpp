// Host-side mesh and system set-up: orientation fix, node reordering,
// node->element incidence, CSR patterns and the per-incidence slot maps that
// let the assembly kernels gather without atomics.
//
// Restates the topology half of NNFEM_Mesh::init (deps/MFEM/Common.cpp:20-142):
// orientation (FinalizeTriMesh(1,0,true), Common.cpp:40), the P1 dof map
// (dof[k] = node[k], Common.cpp:63-67) and the edge table (133-140).  The
// per-element geometry (area, weights, Gauss points) is computed on the GPU
// (geometry.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "internal.h"

namespace admfem {

int quad_rule(int order, QuadRule &r) {
  // Reference-triangle rules (area 1/2).  Order 2 is the symmetric 3-point
  // rule (a,a),(a,b),(b,a) with a = 1/6, b = 2/3, weight 1/6 each, in MFEM's
  // listing order (the rule Common.cpp:42-44 takes for the P1 default
  // order 2, src/MFEM/MFEM.jl:73); orders 0/1 are the centroid rule.  The
  // caller resolves the default order (-1) for its degree.
  if (order == 0 || order == 1) {
    r.nq = 1;
    r.xi[0] = r.eta[0] = 1.0 / 3.0;
    r.w[0] = 0.5;
    return 1;
  }
  if (order == 2) {
    const double a = 1.0 / 6.0, b = 1.0 - 2.0 * a;
    r.nq = 3;
    r.xi[0] = a; r.eta[0] = a;
    r.xi[1] = a; r.eta[1] = b;
    r.xi[2] = b; r.eta[2] = a;
    r.w[0] = r.w[1] = r.w[2] = 1.0 / 6.0;
    return 3;
  }
  if (order == 3 || order == 4) {
    // the 6-point degree-4 rule (Strang & Fix; Dunavant 1985): two orbits
    // (a,a),(a,b),(b,a), b = 1-2a, listed orbit by orbit in the order of the
    // 3-point rule (P2's default order 4, src/MFEM/MFEM.jl:74-75)
    const double a[2] = {0.091576213509770743459571463402202508,
                         0.44594849091596488631832925388305199};
    const double w[2] = {0.054975871827660933819163162450105264,
                         0.11169079483900573284750350421656140};
    r.nq = 6;
    for (int o = 0; o < 2; o++) {
      const double b = 1.0 - 2.0 * a[o];
      r.xi[3 * o + 0] = a[o]; r.eta[3 * o + 0] = a[o];
      r.xi[3 * o + 1] = a[o]; r.eta[3 * o + 1] = b;
      r.xi[3 * o + 2] = b;    r.eta[3 * o + 2] = a[o];
      r.w[3 * o] = r.w[3 * o + 1] = r.w[3 * o + 2] = w[o];
    }
    return 6;
  }
  return -1;
}

size_t Mesh::device_bytes() const {
  return x.bytes() + y.bytes() + xy.bytes() + conn.bytes() + conn_user.bytes() +
         d_perm.bytes() + d_iperm.bytes() + d_n2e_ptr.bytes() +
         d_n2e_elem.bytes() + d_slots_full.bytes() + d_rp.bytes() +
         d_col.bytes() + gx.bytes() + gy.bytes() + gw.bytes() + area.bytes() +
         conn_dof.bytes() + d_d2e_ptr.bytes() + d_d2e.bytes();
}

// Reverse Cuthill-McKee order of the node graph (adjacency = pattern rows
// without the diagonal).  Returns order[k] = node placed at position k.
std::vector<int> rcm_order(int n, const std::vector<int> &rp, const std::vector<int> &col) {
  std::vector<int> deg(n);
  for (int i = 0; i < n; i++) deg[i] = rp[i + 1] - rp[i];
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  std::vector<int> level(n, -1);
  std::vector<int> nb;
  // BFS from s over unvisited nodes, marking `mark`; returns visit order
  auto bfs = [&](int s, std::vector<char> &mark, std::vector<int> &out,
                 bool sorted) {
    size_t head = out.size();
    out.push_back(s);
    mark[s] = 1;
    while (head < out.size()) {
      int u = out[head++];
      nb.clear();
      for (int k = rp[u]; k < rp[u + 1]; k++) {
        int v = col[k];
        if (!mark[v]) {
          mark[v] = 1;
          nb.push_back(v);
        }
      }
      if (sorted)
        std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) {
          return deg[a] < deg[b] || (deg[a] == deg[b] && a < b);
        });
      out.insert(out.end(), nb.begin(), nb.end());
    }
  };
  std::vector<int> byDeg(n);
  std::iota(byDeg.begin(), byDeg.end(), 0);
  std::stable_sort(byDeg.begin(), byDeg.end(),
                   [&](int a, int b) { return deg[a] < deg[b]; });
  std::vector<char> tmpmark(n, 0);
  std::vector<int> comp;
  for (int cand : byDeg) {
    if (seen[cand]) continue;
    // pseudo-peripheral start: repeat BFS from a min-degree node of the
    // last level while the eccentricity grows (George & Liu)
    int start = cand;
    int ecc = -1;
    for (int rep = 0; rep < 4; rep++) {
      comp.clear();
      bfs(start, tmpmark, comp, false);
      // levels
      level[start] = 0;
      int maxl = 0;
      for (int u : comp) {
        for (int k = rp[u]; k < rp[u + 1]; k++) {
          int v = col[k];
          if (level[v] < 0) {
            level[v] = level[u] + 1;
            maxl = std::max(maxl, level[v]);
          }
        }
      }
      int best = start;
      for (int u : comp)
        if (level[u] == maxl && (best == start || deg[u] < deg[best])) best = u;
      for (int u : comp) {
        tmpmark[u] = 0;
        level[u] = -1;
      }
      if (maxl <= ecc) break;
      ecc = maxl;
      start = best;
    }
    size_t before = order.size();
    bfs(start, seen, order, true);
    std::reverse(order.begin() + before, order.end());
  }
  return order;
}



// Element order: nodes numbered by first appearance in the element list
// (vertices in their order within each element; nodes of no element last,
// ascending).  order[k] = node placed at position k.
static std::vector<int> element_order(int n, const std::vector<int> &elems) {
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  for (int v : elems)
    if (!seen[v]) {
      seen[v] = 1;
      order.push_back(v);
    }
  for (int v = 0; v < n; v++)
    if (!seen[v]) order.push_back(v);
  return order;
}

// Bandwidth of the node graph under `order`: max |pos(u) - pos(v)| over edges.
static int64_t order_bandwidth(const std::vector<int> &order, const std::vector<int> &rp,
                               const std::vector<int> &col) {
  const int n = (int)order.size();
  std::vector<int> pos(n);
  for (int k = 0; k < n; k++) pos[order[k]] = k;
  int64_t bw = 0;
  for (int u = 0; u < n; u++)
    for (int k = rp[u]; k < rp[u + 1]; k++)
      bw = std::max<int64_t>(bw, std::abs(pos[u] - pos[col[k]]));
  return bw;
}

// Widest level of a breadth-first search of the node graph from s (its
// component only): the yardstick for a banded order, whose bandwidth is a
// small multiple of it (reverse Cuthill-McKee's is at most about twice the
// widest level of its own level structure).
int64_t bfs_max_level(int n, const std::vector<int> &rp, const std::vector<int> &col, int s) {
  std::vector<int> level(n, -1), q;
  q.reserve(n);
  q.push_back(s);
  level[s] = 0;
  std::vector<int64_t> width(1, 1);
  for (size_t h = 0; h < q.size(); h++) {
    const int u = q[h];
    for (int k = rp[u]; k < rp[u + 1]; k++) {
      const int v = col[k];
      if (level[v] < 0) {
        level[v] = level[u] + 1;
        if ((size_t)level[v] >= width.size()) width.push_back(0);
        width[level[v]]++;
        q.push_back(v);
      }
    }
  }
  return *std::max_element(width.begin(), width.end());
}

Mesh *mesh_build(const double *coords, int nnode, const int *elems, int nelem,
                 int order, int degree) {
  REQUIRE(nnode > 0 && nelem > 0, ADMFEM_ERR_ARG, "empty mesh");
  REQUIRE(coords && elems, ADMFEM_ERR_ARG, "null coords/elems");
  REQUIRE(degree == 1 || degree == 2, ADMFEM_ERR_UNSUPPORTED,
          "only degree 1 (P1) and 2 (P2) elements are built");
  if (order == -1) order = degree == 1 ? 2 : 4;  // src/MFEM/MFEM.jl:71-77
  QuadRule rule;
  REQUIRE(quad_rule(order, rule) > 0, ADMFEM_ERR_UNSUPPORTED,
          "integration order " + std::to_string(order) + " not built (orders 0-4 are)");
  Mesh *m = new Mesh();
  try {
    m->nnode = nnode;
    m->nelem = nelem;
    m->order = order;
    m->degree = degree;
    m->nd = degree == 1 ? 3 : 6;
    m->ndof = nnode;
    m->rule = rule;
    m->ngauss = (int64_t)rule.nq * nelem;
    m->coords.assign(coords, coords + 2 * (size_t)nnode);
    m->elems.assign(elems, elems + 3 * (size_t)nelem);
    for (size_t k = 0; k < m->elems.size(); k++)
      REQUIRE(m->elems[k] >= 0 && m->elems[k] < nnode, ADMFEM_ERR_ARG,
              "element vertex id out of range");
    // orientation fix (FinalizeTriMesh(1,0,true), Common.cpp:40): swap the
    // first two vertices of a triangle with negative Jacobian determinant
    for (int e = 0; e < nelem; e++) {
      int *v = &m->elems[3 * (size_t)e];
      double x1 = coords[2 * v[0]], y1 = coords[2 * v[0] + 1];
      double x2 = coords[2 * v[1]], y2 = coords[2 * v[1] + 1];
      double x3 = coords[2 * v[2]], y3 = coords[2 * v[2] + 1];
      double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
      if (det < 0.0) std::swap(v[0], v[1]);
    }
    // node -> element incidence in user numbering (elements ascending)
    std::vector<int> uptr(nnode + 1, 0);
    for (int e = 0; e < nelem; e++)
      for (int k = 0; k < 3; k++) uptr[m->elems[3 * (size_t)e + k] + 1]++;
    for (int i = 0; i < nnode; i++) uptr[i + 1] += uptr[i];
    std::vector<int> uinc(uptr[nnode]);
    {
      std::vector<int> pos(uptr.begin(), uptr.end() - 1);
      for (int e = 0; e < nelem; e++)
        for (int k = 0; k < 3; k++) uinc[pos[m->elems[3 * (size_t)e + k]]++] = e;
    }
    // pre-BC pattern, user numbering: union of incident elements' vertices,
    // built per chunk of nodes on host threads, then concatenated
    std::vector<int> &urp = m->urp, &ucol = m->ucol;
    urp.assign(nnode + 1, 0);
    {
      std::vector<std::vector<int>> part(16);
      auto rows = [&](int64_t b, int64_t e, int c) {
        std::vector<int> &out = part[c];
        out.reserve((size_t)(e - b) * 8);
        std::vector<int> cand;
        for (int i = (int)b; i < (int)e; i++) {
          cand.clear();
          for (int k = uptr[i]; k < uptr[i + 1]; k++) {
            const int *v = &m->elems[3 * (size_t)uinc[k]];
            cand.push_back(v[0]);
            cand.push_back(v[1]);
            cand.push_back(v[2]);
          }
          std::sort(cand.begin(), cand.end());
          cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
          REQUIRE((int)cand.size() <= MAX_ROW, ADMFEM_ERR_UNSUPPORTED,
                  "a node has more than 61 neighbours");
          out.insert(out.end(), cand.begin(), cand.end());
          urp[i + 1] = (int)cand.size();
        }
      };
      parallel_chunks(nnode, PAR_GRAIN, rows);
      for (int i = 0; i < nnode; i++) urp[i + 1] += urp[i];
      ucol.resize(urp[nnode]);
      parallel_chunks(nnode, PAR_GRAIN, [&](int64_t b, int64_t, int c) {
        std::copy(part[c].begin(), part[c].end(), ucol.begin() + urp[b]);
      });
    }
    // internal numbering: the element order when it is banded (bandwidth
    // at most 5x the widest level of a breadth-first search from the first
    // element's first vertex), else reverse Cuthill-McKee over the node
    // graph.  Either way rows run along fronts: the k-th neighbours of
    // consecutive rows are consecutive nodes of the adjacent fronts, so the
    // solver's column gathers coalesce within a wavefront.  The element
    // order also makes the elements of a block of rows consecutive in the
    // user's element order, so the assembly's per-element inputs (kappa, f,
    // given in that order) share cache lines.
    {
      std::vector<int> arp(nnode + 1, 0), acol;
      acol.reserve(ucol.size());
      for (int i = 0; i < nnode; i++) {
        for (int k = urp[i]; k < urp[i + 1]; k++)
          if (ucol[k] != i) acol.push_back(ucol[k]);
        arp[i + 1] = (int)acol.size();
      }
      std::vector<int> eo = element_order(nnode, m->elems);
      const int64_t bw_eo = order_bandwidth(eo, arp, acol);
      if (bw_eo <= 5 * bfs_max_level(nnode, arp, acol, m->elems[0]))
        m->iperm = std::move(eo);
      else
        m->iperm = rcm_order(nnode, arp, acol);
      m->perm.assign(nnode, 0);
      for (int k = 0; k < nnode; k++) m->perm[m->iperm[k]] = k;
    }
    const std::vector<int> &P = m->perm, &IP = m->iperm;
    // internal pattern and incidences
    m->rp.assign(nnode + 1, 0);
    m->col.resize(ucol.size());
    m->n2e_ptr.assign(nnode + 1, 0);
    m->n2e_elem.resize(uinc.size());
    m->n2e_p.resize(uinc.size());
    for (int r = 0; r < nnode; r++) {
      int u = IP[r];
      int len = urp[u + 1] - urp[u];
      m->rp[r + 1] = m->rp[r] + len;
      int inc = uptr[u + 1] - uptr[u];
      m->n2e_ptr[r + 1] = m->n2e_ptr[r] + inc;
    }
    std::vector<uint32_t> slots(uinc.size());
    std::vector<int> conn_int(3 * (size_t)nelem), conn_usr(3 * (size_t)nelem);
    parallel_chunks(nelem, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
      for (int e = (int)b; e < (int)e1; e++)
        for (int k = 0; k < 3; k++) {
          int u = m->elems[3 * (size_t)e + k];
          conn_usr[(size_t)k * nelem + e] = u;
          conn_int[(size_t)k * nelem + e] = P[u];
        }
    });
    // rows are independent: host threads over row chunks
    parallel_chunks(nnode, PAR_GRAIN, [&](int64_t rb, int64_t re, int) {
    for (int r = (int)rb; r < (int)re; r++) {
      int u = IP[r];
      int *c = &m->col[m->rp[r]];
      int len = 0;
      for (int k = urp[u]; k < urp[u + 1]; k++) c[len++] = P[ucol[k]];
      std::sort(c, c + len);
      int base = m->n2e_ptr[r];
      for (int k = uptr[u]; k < uptr[u + 1]; k++) {
        int e = uinc[k];
        int t = base + (k - uptr[u]);
        m->n2e_elem[t] = e;
        int p = -1;
        uint32_t w = 0;
        for (int q = 0; q < 3; q++) {
          int vq = m->elems[3 * (size_t)e + q];
          if (vq == u && p < 0) p = q;
          int cq = P[vq];
          int s = (int)(std::lower_bound(c, c + len, cq) - c);
          w |= (uint32_t)s << (2 + SLOT_BITS * q);
        }
        REQUIRE(p >= 0, ADMFEM_ERR_ARG, "internal incidence error");
        m->n2e_p[t] = (uint8_t)p;
        slots[t] = w | (uint32_t)p;
      }
    }
    });
    // device copies
    std::vector<double> xs(nnode), ys(nnode);
    std::vector<double2> xys(nnode);
    for (int r = 0; r < nnode; r++) {
      xs[r] = coords[2 * (size_t)IP[r]];
      ys[r] = coords[2 * (size_t)IP[r] + 1];
      xys[r] = make_double2(xs[r], ys[r]);
    }
    m->x.upload(xs);
    m->y.upload(ys);
    m->xy.upload(xys);
    m->conn.upload(conn_int);
    m->conn_user.upload(conn_usr);
    m->d_perm.upload(m->perm);
    m->d_iperm.upload(m->iperm);
    m->d_n2e_ptr.upload(m->n2e_ptr);
    m->d_n2e_elem.upload(m->n2e_elem);
    m->d_slots_full.upload(slots);
    m->d_rp.upload(m->rp);
    m->d_col.upload(m->col);
    m->gx.alloc(m->ngauss);
    m->gy.alloc(m->ngauss);
    m->gw.alloc(m->ngauss);
    m->area.alloc(nelem);
    launch_geometry(*m, nullptr);
    if (degree == 2) {
      // P2 dofs (Common.cpp:63-72): the vertices, then nnode + edge id of the
      // element's edges (0,1),(1,2),(2,0); dof -> element incidences in
      // ascending element order for the per-dof gathers
      mesh_edges(*m);
      m->ndof = nnode + m->nedges;
      m->dofs.resize(6 * (size_t)nelem);
      std::vector<int> cd(6 * (size_t)nelem);
      parallel_chunks(nelem, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
        for (int e = (int)b; e < (int)e1; e++)
          for (int k = 0; k < 3; k++) {
            const int v = m->elems[3 * (size_t)e + k];
            const int d = nnode + m->elem_edge[3 * (size_t)e + k];
            m->dofs[6 * (size_t)e + k] = v;
            m->dofs[6 * (size_t)e + 3 + k] = d;
            cd[(size_t)k * nelem + e] = v;
            cd[(size_t)(3 + k) * nelem + e] = d;
          }
      });
      REQUIRE(m->ndof < (int64_t)1 << 31, ADMFEM_ERR_UNSUPPORTED, "too many P2 dofs");
      std::vector<int> dptr(m->ndof + 1, 0), d2e(6 * (size_t)nelem);
      for (size_t k = 0; k < m->dofs.size(); k++) dptr[m->dofs[k] + 1]++;
      for (int64_t i = 0; i < m->ndof; i++) dptr[i + 1] += dptr[i];
      std::vector<int> pos(dptr.begin(), dptr.end() - 1);
      for (int e = 0; e < nelem; e++)
        for (int p = 0; p < 6; p++) d2e[pos[m->dofs[6 * (size_t)e + p]]++] = e * 8 + p;
      REQUIRE(nelem < (1 << 28), ADMFEM_ERR_UNSUPPORTED, "too many P2 elements");
      m->conn_dof.upload(cd);
      m->d_d2e_ptr.upload(dptr);
      m->d_d2e.upload(d2e);
    }
    HIPCHECK(hipStreamSynchronize(nullptr));
  } catch (...) {
    delete m;
    throw;
  }
  return m;
}

// Edge table in MFEM's numbering: an edge gets the next id the first time it
// is met walking the oriented elements and their edges (0,1),(1,2),(2,0)
// (Mesh::GetVertexToVertexTable / DSTable::Push order); output as
// Common.cpp:133-140: [min+1 ...; max+1 ...].  The same pass counts each
// edge's uses for the boundary nodes (edges used by one element,
// src/MFEM/MCore.jl:385-406 and 439-450: set toggling keeps odd counts).
//
// Per node v (host threads over node chunks), walking v's incident elements
// in ascending order: the first sighting of edge (v, w) with v < w is its
// first appearance overall, since every element holding the edge holds v.
// A serial scan of the (element, edge) keys in walk order then numbers them.
void mesh_edges(Mesh &m) {
  if (m.nedges >= 0) return;
  const int n = m.nnode;
  static const int ev[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  std::vector<char> first(3 * (size_t)m.nelem, 0);
  m.bd_node.assign(n, 0);
  parallel_chunks(n, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
    int cnt[MAX_ROW + 1];
    char seen[MAX_ROW + 1];
    for (int v = (int)b; v < (int)e1; v++) {
      const int *c = &m.ucol[m.urp[v]];
      const int len = m.urp[v + 1] - m.urp[v];
      std::fill(cnt, cnt + len, 0);
      std::fill(seen, seen + len, 0);
      const int r = m.perm[v];
      for (int t = m.n2e_ptr[r]; t < m.n2e_ptr[r + 1]; t++) {
        const int el = m.n2e_elem[t];
        const int *vv = &m.elems[3 * (size_t)el];
        for (int j = 0; j < 3; j++) {
          const int a = vv[ev[j][0]], bq = vv[ev[j][1]];
          if (a == bq || (a != v && bq != v)) continue;
          const int w = a == v ? bq : a;
          const int sl = (int)(std::lower_bound(c, c + len, w) - c);
          cnt[sl]++;
          if (v < w && !seen[sl]) {
            seen[sl] = 1;
            first[3 * (size_t)el + j] = 1;
          }
        }
      }
      char bd = 0;
      for (int k = 0; k < len; k++) bd |= (char)(cnt[k] & 1);
      m.bd_node[v] = bd;
    }
  });
  std::vector<long long> lo_of, hi_of;
  lo_of.reserve((size_t)n * 3);
  hi_of.reserve((size_t)n * 3);
  const bool p2 = m.degree == 2;
  std::vector<int> key_id(p2 ? first.size() : 0, -1);
  for (size_t k = 0; k < first.size(); k++)
    if (first[k]) {
      const int *vv = &m.elems[3 * (k / 3)];
      const int a = vv[ev[k % 3][0]], bq = vv[ev[k % 3][1]];
      if (p2) key_id[k] = (int)lo_of.size();
      lo_of.push_back(std::min(a, bq) + 1);
      hi_of.push_back(std::max(a, bq) + 1);
    }
  if (p2) {
    // every element edge's id, and which edges are on the boundary: per
    // lower endpoint v, the first sighting of (v, w) carries the id
    m.elem_edge.assign(first.size(), -1);
    m.bd_edge.assign(lo_of.size(), 0);
    parallel_chunks(n, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
      int sid[MAX_ROW + 1], cnt[MAX_ROW + 1];
      for (int v = (int)b; v < (int)e1; v++) {
        const int *c = &m.ucol[m.urp[v]];
        const int len = m.urp[v + 1] - m.urp[v];
        std::fill(sid, sid + len, -1);
        std::fill(cnt, cnt + len, 0);
        const int r = m.perm[v];
        for (int t = m.n2e_ptr[r]; t < m.n2e_ptr[r + 1]; t++) {
          const int el = m.n2e_elem[t];
          const int *vv = &m.elems[3 * (size_t)el];
          for (int j = 0; j < 3; j++) {
            const int a = vv[ev[j][0]], bq = vv[ev[j][1]];
            if (a == bq || std::min(a, bq) != v) continue;
            const int w = std::max(a, bq);
            const int sl = (int)(std::lower_bound(c, c + len, w) - c);
            if (sid[sl] < 0) sid[sl] = key_id[3 * (size_t)el + j];
            cnt[sl]++;
            m.elem_edge[3 * (size_t)el + j] = sid[sl];
          }
        }
        for (int k = 0; k < len; k++)
          if (sid[k] >= 0) m.bd_edge[sid[k]] = (char)(cnt[k] & 1);
      }
    });
  }
  m.nedges = (int64_t)lo_of.size();
  m.edges.resize(2 * lo_of.size());
  std::copy(lo_of.begin(), lo_of.end(), m.edges.begin());
  std::copy(hi_of.begin(), hi_of.end(), m.edges.begin() + lo_of.size());
}

size_t System::device_bytes() const {
  return d_rp.bytes() + d_col.bytes() + d_row_bd.bytes() + d_blk_eptr.bytes() +
         d_blk_rec.bytes() + d_blk_vptr.bytes() + d_blk_vert.bytes() + d_pairs.bytes() + val.bytes() + rhs.bytes() +
         d_slice_ptr.bytes() + d_sell_col.bytes() +
         d_sell_val.bytes() + d_diag.bytes() + d_perm2.bytes() + d_iperm2.bytes() +
         d_conn2.bytes() + d_inc_ptr.bytes() + d_blk_elem.bytes() + d_pairs2.bytes();
}

// SELL-64 layout of the off-diagonal entries of the system's CSR pattern
// (slices of 64 rows, column-major inside a slice, width = the slice's
// widest row; padding entries point at the row itself with value 0) and the
// device buffers of the assembled values and of the solver.
static void build_solver_layout(System *s) {
  const int n = s->n;
  s->nslices = s->npad / 64;
  s->slice_ptr.assign(s->nslices + 1, 0);
  auto offdiag = [&](int r) {
    int c = 0;
    for (int k = s->rp[r]; k < s->rp[r + 1]; k++) c += s->col[k] != r;
    return c;
  };
  for (int sl = 0; sl < s->nslices; sl++) {
    int wmax = 0;
    for (int l = 0; l < 64; l++) {
      int r = sl * 64 + l;
      if (r >= n) break;
      wmax = std::max(wmax, offdiag(r));
    }
    s->slice_ptr[sl + 1] = s->slice_ptr[sl] + 64 * wmax;
    s->sell_wmax = std::max(s->sell_wmax, wmax);
  }
  std::vector<int> sell_col(std::max(1, s->slice_ptr[s->nslices]));
  for (int sl = 0; sl < s->nslices; sl++) {
    int off = s->slice_ptr[sl];
    int wmax = (s->slice_ptr[sl + 1] - off) / 64;
    for (int l = 0; l < 64; l++) {
      int r = sl * 64 + l;
      int kk = 0;
      if (r < n)
        for (int k = s->rp[r]; k < s->rp[r + 1]; k++)
          if (s->col[k] != r) sell_col[off + kk++ * 64 + l] = s->col[k];
      for (; kk < wmax; kk++) sell_col[off + kk * 64 + l] = r < n ? r : 0;
    }
  }
  s->d_rp.upload(s->rp);
  s->d_col.upload(s->col);
  if (!s->mesh) s->val.alloc(std::max<size_t>(1, s->col.size()));
  s->rhs.alloc(s->npad);
  HIPCHECK(hipMemset(s->rhs.p, 0, s->rhs.bytes()));
  s->d_slice_ptr.upload(s->slice_ptr);
  s->d_sell_col.upload(sell_col);
  s->d_sell_val.alloc(sell_col.size());
  HIPCHECK(hipMemset(s->d_sell_val.p, 0, s->d_sell_val.bytes()));
  s->d_diag.alloc(s->npad);
}

// The fused system of a P2 mesh.  Rows are the ndof dofs (vertices, then
// nnode + edge id, deps/MFEM/Common.cpp:63-72) in an internal order built
// from the mesh's node order: each vertex dof at its node's position,
// followed by the edge dofs whose lower-numbered end it is (ascending other
// end), so the dof graph keeps the node graph's banded fronts.  Post-BC
// pattern and Dirichlet bookkeeping as for P1 (ImposeDirichlet.h:27-51 on
// the P2 triplets); the assembly blocks are 64-row SELL slices with the
// distinct elements incident to their interior rows and one 64-bit pair word
// per (interior row, incident element).
static System *system_build_p2(Mesh *m, const long long *bdnode, int nbd) {
  REQUIRE(m->ndof < (int64_t)1 << 30, ADMFEM_ERR_UNSUPPORTED,
          "too many P2 dofs for the fused system");
  System *s = new System();
  try {
    const int nn = m->nnode, ne = m->nelem, n = (int)m->ndof;
    s->mesh = m;
    s->degree = 2;
    s->n = n;
    s->npad = (n + 63) / 64 * 64;
    s->nbd = nbd;
    // internal dof order: edge dofs bucketed by their lower-numbered end
    {
      std::vector<int> lo(n - nn, -1), hi(n - nn, -1);
      for (int e = 0; e < ne; e++)
        for (int k = 0; k < 3; k++) {
          const int d = m->dofs[6 * (size_t)e + 3 + k] - nn;
          if (lo[d] >= 0) continue;
          const int a = m->perm[m->elems[3 * (size_t)e + k]];
          const int b = m->perm[m->elems[3 * (size_t)e + (k + 1) % 3]];
          lo[d] = std::min(a, b);
          hi[d] = std::max(a, b);
        }
      std::vector<int> bptr(nn + 1, 0), bk(n - nn);
      for (int d = 0; d < n - nn; d++) bptr[lo[d] + 1]++;
      for (int r = 0; r < nn; r++) bptr[r + 1] += bptr[r];
      {
        std::vector<int> pos(bptr.begin(), bptr.end() - 1);
        for (int d = 0; d < n - nn; d++) bk[pos[lo[d]]++] = d;
      }
      s->iperm2.reserve(n);
      for (int r = 0; r < nn; r++) {
        s->iperm2.push_back(m->iperm[r]);
        std::sort(bk.begin() + bptr[r], bk.begin() + bptr[r + 1],
                  [&](int a, int b) { return hi[a] < hi[b]; });
        for (int k = bptr[r]; k < bptr[r + 1]; k++) s->iperm2.push_back(nn + bk[k]);
      }
      s->perm2.assign(n, 0);
      for (int k = 0; k < n; k++) s->perm2[s->iperm2[k]] = k;
    }
    // element dofs as internal rows, row -> (element, local dof) incidences
    // in ascending element order
    std::vector<int> conn2(6 * (size_t)ne), iptr(n + 1, 0), ie(6 * (size_t)ne);
    std::vector<uint8_t> ip(6 * (size_t)ne);
    for (int e = 0; e < ne; e++)
      for (int q = 0; q < 6; q++) {
        const int r = s->perm2[m->dofs[6 * (size_t)e + q]];
        conn2[(size_t)q * ne + e] = r;
        iptr[r + 1]++;
      }
    for (int r = 0; r < n; r++) iptr[r + 1] += iptr[r];
    {
      std::vector<int> pos(iptr.begin(), iptr.end() - 1);
      for (int e = 0; e < ne; e++)
        for (int q = 0; q < 6; q++) {
          const int t = pos[conn2[(size_t)q * ne + e]]++;
          ie[t] = e;
          ip[t] = (uint8_t)q;
        }
    }
    s->row_bd.assign(n, -1);
    for (int i = 0; i < nbd; i++) {
      const long long b = bdnode[i];
      REQUIRE(b >= 1 && b <= n, ADMFEM_ERR_ARG, "bdnode entry out of range (1-based dof ids)");
      s->row_bd[s->perm2[b - 1]] = i;  // later entries win
    }
    for (int r = 0; r < n; r++) s->nbd_unique += s->row_bd[r] >= 0;
    // post-BC pattern: per interior row the interior dofs of its elements,
    // per boundary row its diagonal; lengths, then columns, on host threads
    auto row_cols = [&](int r, std::vector<int> &c) {
      c.clear();
      if (s->row_bd[r] >= 0) {
        c.push_back(r);
        return;
      }
      for (int t = iptr[r]; t < iptr[r + 1]; t++)
        for (int q = 0; q < 6; q++) {
          const int cq = conn2[(size_t)q * ne + ie[t]];
          if (s->row_bd[cq] < 0) c.push_back(cq);
        }
      std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
      REQUIRE((int)c.size() <= MAX_ROW, ADMFEM_ERR_UNSUPPORTED,
              "a P2 dof has more than 61 neighbours");
    };
    s->rp.assign(n + 1, 0);
    parallel_chunks(n, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
      std::vector<int> c;
      for (int r = (int)b; r < (int)e1; r++) {
        row_cols(r, c);
        s->rp[r + 1] = (int)c.size();
      }
    });
    for (int r = 0; r < n; r++) s->rp[r + 1] += s->rp[r];
    s->col.resize(s->rp[n]);
    parallel_chunks(n, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
      std::vector<int> c;
      for (int r = (int)b; r < (int)e1; r++) {
        row_cols(r, c);
        std::copy(c.begin(), c.end(), s->col.begin() + s->rp[r]);
      }
    });
    build_solver_layout(s);
    REQUIRE(s->sell_wmax < (int)PAIR_DIAG, ADMFEM_ERR_UNSUPPORTED, "row wider than 61");
    // assembly blocks (host threads over chunks of blocks, element lists
    // concatenated in block order)
    const int nblk = s->npad / AS_ROWS;
    std::vector<int> beptr(nblk + 1, 0);
    std::vector<uint64_t> pairs(std::max<size_t>(1, ie.size()), 0);
    std::vector<std::vector<int>> part(16);
    std::vector<int> part_max(16, 0);
    parallel_chunks(nblk, 64, [&](int64_t bb, int64_t be, int c) {
      std::vector<int> &bl = part[c];
      std::vector<int> offd;
      for (int b = (int)bb; b < (int)be; b++) {
        const int r0 = b * AS_ROWS, r1 = std::min(r0 + AS_ROWS, n);
        const size_t first = bl.size();
        for (int r = r0; r < r1; r++)
          if (s->row_bd[r] < 0)
            for (int t = iptr[r]; t < iptr[r + 1]; t++) bl.push_back(ie[t]);
        std::sort(bl.begin() + first, bl.end());
        bl.erase(std::unique(bl.begin() + first, bl.end()), bl.end());
        const int nb = (int)(bl.size() - first);
        REQUIRE(nb <= PAIR_MAX_ELEMS, ADMFEM_ERR_UNSUPPORTED,
                "assembly block has too many elements");
        part_max[c] = std::max(part_max[c], nb);
        beptr[b + 1] = nb;
        for (int r = r0; r < r1; r++) {
          if (s->row_bd[r] >= 0) continue;
          offd.clear();
          for (int k = s->rp[r]; k < s->rp[r + 1]; k++)
            if (s->col[k] != r) offd.push_back(s->col[k]);
          for (int t = iptr[r]; t < iptr[r + 1]; t++) {
            const int e = ie[t];
            const uint64_t le = (uint64_t)(std::lower_bound(bl.begin() + first, bl.end(), e) -
                                           (bl.begin() + first));
            uint64_t w = (uint64_t)ip[t] | (le << P2_PAIR_ELEM);
            for (int q = 0; q < 6; q++) {
              const int cq = conn2[(size_t)q * ne + e];
              uint64_t code;
              if (s->row_bd[cq] >= 0) code = SLOT_LIFT;
              else if (cq == r) code = PAIR_DIAG;
              else code = (uint64_t)(std::lower_bound(offd.begin(), offd.end(), cq) - offd.begin());
              w |= code << (P2_PAIR_CODE0 + SLOT_BITS * q);
            }
            pairs[t] = w;
          }
        }
      }
    });
    for (int b = 0; b < nblk; b++) beptr[b + 1] += beptr[b];
    for (int c = 0; c < 16; c++) s->asm_max_e = std::max(s->asm_max_e, part_max[c]);
    std::vector<int> belem(std::max<int>(1, beptr[nblk]), 0);
    parallel_chunks(nblk, 64, [&](int64_t bb, int64_t, int c) {
      std::copy(part[c].begin(), part[c].end(), belem.begin() + beptr[bb]);
    });
    s->d_row_bd.upload(s->row_bd);
    s->d_perm2.upload(s->perm2);
    s->d_iperm2.upload(s->iperm2);
    s->d_conn2.upload(conn2);
    s->d_inc_ptr.upload(iptr);
    s->d_blk_eptr.upload(beptr);
    s->d_blk_elem.upload(belem);
    s->d_pairs2.upload(pairs);
  } catch (...) {
    system_free(s);
    throw;
  }
  return s;
}

System *system_build(Mesh *m, const long long *bdnode, int nbd) {
  REQUIRE(m, ADMFEM_ERR_ARG, "null mesh");
  REQUIRE(nbd >= 0 && (nbd == 0 || bdnode), ADMFEM_ERR_ARG, "bad bdnode");
  if (m->degree == 2) return system_build_p2(m, bdnode, nbd);
  System *s = new System();
  try {
    const int n = m->nnode;
    s->mesh = m;
    s->n = n;
    s->npad = (n + 63) / 64 * 64;
    s->nbd = nbd;
    s->row_bd.assign(n, -1);
    for (int i = 0; i < nbd; i++) {
      long long b = bdnode[i];
      REQUIRE(b >= 1 && b <= n, ADMFEM_ERR_ARG, "bdnode entry out of range (1-based)");
      s->row_bd[m->perm[b - 1]] = i;  // later entries win (std::map assignment)
    }
    for (int r = 0; r < n; r++) s->nbd_unique += s->row_bd[r] >= 0;
    // post-BC pattern: interior-interior pairs + boundary diagonals
    // (ImposeDirichlet.h:27-51 keeps triplets with i,j not on the boundary,
    // then appends (b,b,1))
    s->rp.assign(n + 1, 0);
    s->col.reserve(m->col.size());
    for (int r = 0; r < n; r++) {
      if (s->row_bd[r] >= 0) {
        s->col.push_back(r);
      } else {
        for (int k = m->rp[r]; k < m->rp[r + 1]; k++)
          if (s->row_bd[m->col[k]] < 0) s->col.push_back(m->col[k]);
      }
      s->rp[r + 1] = (int)s->col.size();
    }
    build_solver_layout(s);
    // assembly blocks: elements incident to the block's interior rows, and
    // one pair word per (row, element) incidence; host threads over chunks
    // of blocks, element lists concatenated in block order
    const int nblk = (s->npad + AS_ROWS - 1) / AS_ROWS;
    std::vector<int> beptr(nblk + 1, 0), bvptr(nblk + 1, 0);
    std::vector<uint32_t> pairs(m->n2e_elem.size(), 0);
    std::vector<std::vector<int>> part(16), part_v(16), part_loc(16);
    std::vector<int> part_max(16, 0), part_maxv(16, 0);
    auto blocks = [&](int64_t bb, int64_t be, int c) {
      std::vector<int> &bl = part[c], &bv = part_v[c], &bloc = part_loc[c];
      std::vector<int> offd;
      for (int b = (int)bb; b < (int)be; b++) {
        const int r0 = b * AS_ROWS, r1 = std::min(r0 + AS_ROWS, n);
        const size_t first = bl.size();
        for (int r = r0; r < r1; r++)
          if (s->row_bd[r] < 0)
            for (int t = m->n2e_ptr[r]; t < m->n2e_ptr[r + 1]; t++) bl.push_back(m->n2e_elem[t]);
        std::sort(bl.begin() + first, bl.end());
        bl.erase(std::unique(bl.begin() + first, bl.end()), bl.end());
        const int ne = (int)(bl.size() - first);
        REQUIRE(ne <= PAIR_MAX_ELEMS, ADMFEM_ERR_UNSUPPORTED,
                "assembly block has too many elements");
        part_max[c] = std::max(part_max[c], ne);
        beptr[b + 1] = ne;
        // the block's distinct vertices (ascending internal ids: a few runs
        // of consecutive rows under RCM) and each element's local indices
        const size_t vfirst = bv.size();
        for (size_t k = first; k < bl.size(); k++)
          for (int q = 0; q < 3; q++) bv.push_back(m->perm[m->elems[3 * (size_t)bl[k] + q]]);
        std::sort(bv.begin() + vfirst, bv.end());
        bv.erase(std::unique(bv.begin() + vfirst, bv.end()), bv.end());
        const int nv = (int)(bv.size() - vfirst);
        REQUIRE(nv <= BLK_MAX_VERTS, ADMFEM_ERR_UNSUPPORTED,
                "assembly block touches more than 1024 vertices");
        part_maxv[c] = std::max(part_maxv[c], nv);
        bvptr[b + 1] = nv;
        for (size_t k = first; k < bl.size(); k++) {
          int loc = 0;
          for (int q = 0; q < 3; q++) {
            const int v = m->perm[m->elems[3 * (size_t)bl[k] + q]];
            const int l = (int)(std::lower_bound(bv.begin() + vfirst, bv.end(), v) -
                                (bv.begin() + vfirst));
            loc |= l << (BLK_VERT_BITS * q);
          }
          bloc.push_back(loc);
        }
        for (int r = r0; r < r1; r++) {
          if (s->row_bd[r] >= 0) continue;
          offd.clear();
          for (int k = s->rp[r]; k < s->rp[r + 1]; k++)
            if (s->col[k] != r) offd.push_back(s->col[k]);
          for (int t = m->n2e_ptr[r]; t < m->n2e_ptr[r + 1]; t++) {
            const int e = m->n2e_elem[t];
            const uint32_t lei = (uint32_t)(std::lower_bound(bl.begin() + first, bl.end(), e) -
                                            (bl.begin() + first));
            uint32_t w = m->n2e_p[t] | (lei << 20);
            for (int q = 0; q < 3; q++) {
              const int cq = m->perm[m->elems[3 * (size_t)e + q]];
              uint32_t code;
              if (s->row_bd[cq] >= 0) code = SLOT_LIFT;
              else if (cq == r) code = PAIR_DIAG;
              else code = (uint32_t)(std::lower_bound(offd.begin(), offd.end(), cq) - offd.begin());
              w |= code << (2 + SLOT_BITS * q);
            }
            pairs[t] = w;
          }
        }
      }
    };
    parallel_chunks(nblk, 64, blocks);
    for (int b = 0; b < nblk; b++) {
      beptr[b + 1] += beptr[b];
      bvptr[b + 1] += bvptr[b];
    }
    for (int c = 0; c < 16; c++) {
      s->asm_max_e = std::max(s->asm_max_e, part_max[c]);
      s->asm_max_v = std::max(s->asm_max_v, part_maxv[c]);
    }
    REQUIRE(s->sell_wmax < (int)PAIR_DIAG, ADMFEM_ERR_UNSUPPORTED, "row wider than 61");
    // chunk c covers a contiguous range of blocks: its lists start at the
    // offsets of its first block
    std::vector<int2> brec(std::max<int>(1, beptr[nblk]), make_int2(0, 0));
    std::vector<int> bvert(std::max<int>(1, bvptr[nblk]), 0);
    parallel_chunks(nblk, 64, [&](int64_t bb, int64_t, int c) {
      for (size_t k = 0; k < part[c].size(); k++)
        brec[beptr[bb] + k] = make_int2(part[c][k], part_loc[c][k]);
      std::copy(part_v[c].begin(), part_v[c].end(), bvert.begin() + bvptr[bb]);
    });
    s->d_row_bd.upload(s->row_bd);
    s->d_blk_eptr.upload(beptr);
    s->d_blk_rec.upload(brec);
    s->d_blk_vptr.upload(bvptr);
    s->d_blk_vert.upload(bvert);
    s->d_pairs.upload(pairs.empty() ? std::vector<uint32_t>(1, 0) : pairs);
  } catch (...) {
    system_free(s);
    throw;
  }
  return s;
}

// A system with no mesh: the CSR pattern of a sparse matrix given by the
// caller (the general K\F op), identity node order, no boundary rows.
System *system_from_csr(int n, std::vector<int> rp, std::vector<int> col) {
  REQUIRE(n > 0 && (int)rp.size() == n + 1, ADMFEM_ERR_ARG, "bad CSR pattern");
  System *s = new System();
  try {
    s->mesh = nullptr;
    s->n = n;
    s->npad = (n + 63) / 64 * 64;
    s->row_bd.assign(n, -1);
    s->rp = std::move(rp);
    s->col = std::move(col);
    build_solver_layout(s);
    s->d_row_bd.upload(s->row_bd);
  } catch (...) {
    system_free(s);
    throw;
  }
  return s;
}

void system_free(System *s) {
  if (!s) return;
  if (s->work) cg_work_free(s->work);
  delete s;
}

}  // namespace admfem
