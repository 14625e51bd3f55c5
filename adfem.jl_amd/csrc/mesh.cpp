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
         conn_dof.bytes() + d_d2e_ptr.bytes() + d_d2e.bytes() + d_edges.bytes();
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
    m->elems.assign(elems, elems + 3 * (size_t)nelem);
    // orientation fix (FinalizeTriMesh(1,0,true), Common.cpp:40): swap the
    // first two vertices of a triangle with negative Jacobian determinant
    parallel_chunks(nelem, PAR_GRAIN, [&](int64_t b, int64_t e1, int) {
      for (int e = (int)b; e < (int)e1; e++) {
        int *v = &m->elems[3 * (size_t)e];
        for (int k = 0; k < 3; k++)
          REQUIRE(v[k] >= 0 && v[k] < nnode, ADMFEM_ERR_ARG, "element vertex id out of range");
        double x1 = coords[2 * v[0]], y1 = coords[2 * v[0] + 1];
        double x2 = coords[2 * v[1]], y2 = coords[2 * v[1] + 1];
        double x3 = coords[2 * v[2]], y3 = coords[2 * v[2] + 1];
        double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        if (det < 0.0) std::swap(v[0], v[1]);
      }
    });
    // node order, connectivity, incidences, pattern: on the device
    mesh_topology(*m, coords, nullptr);
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

// The pre-BC pattern in user numbering (rows sorted), from the device
// pattern in the internal order; made on first request.
void mesh_user_pattern(Mesh &m) {
  if (!m.urp.empty()) return;
  const int n = m.nnode;
  std::vector<int> rp(n + 1), col(m.nnz_full);
  HIPCHECK(hipMemcpy(rp.data(), m.d_rp.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(col.data(), m.d_col.p, sizeof(int) * col.size(), hipMemcpyDeviceToHost));
  std::vector<int> urp(n + 1, 0);
  for (int u = 0; u < n; u++) urp[u + 1] = urp[u] + rp[m.perm[u] + 1] - rp[m.perm[u]];
  m.ucol.resize(m.nnz_full);
  parallel_chunks(n, PAR_GRAIN, [&](int64_t b, int64_t e, int) {
    for (int u = (int)b; u < (int)e; u++) {
      const int r = m.perm[u];
      int *c = &m.ucol[urp[u]];
      for (int k = rp[r]; k < rp[r + 1]; k++) c[k - rp[r]] = m.iperm[col[k]];
      std::sort(c, c + (rp[r + 1] - rp[r]));
    }
  });
  m.urp = std::move(urp);
}

size_t System::device_bytes() const {
  return d_rp.bytes() + d_col.bytes() + d_row_bd.bytes() + d_blk_eptr.bytes() +
         d_blk_rec.bytes() + d_blk_vptr.bytes() + d_blk_vert.bytes() + d_pairs.bytes() + val.bytes() + rhs.bytes() +
         d_slice_ptr.bytes() + d_sell_col.bytes() + d_sell_d16.bytes() +
         d_sell_val.bytes() + d_diag.bytes() + d_perm2.bytes() + d_iperm2.bytes() +
         d_conn2.bytes() + d_inc_ptr.bytes() + d_blk_elem.bytes() + d_pairs2.bytes() +
         d_vrow.bytes() + d_dof_ends.bytes();
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
      std::vector<int> vrow(nn);
      std::vector<int2> ends(n);
      for (int r = 0; r < nn; r++) vrow[r] = s->perm2[m->iperm[r]];
      for (int k = 0; k < n; k++) {
        const int d = s->iperm2[k];
        ends[k] = d < nn ? make_int2(m->perm[d], m->perm[d]) : make_int2(lo[d - nn], hi[d - nn]);
      }
      s->d_vrow.upload(vrow);
      s->d_dof_ends.upload(ends);
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
    s->nnz = (int64_t)s->col.size();
    s->d_rp.upload(s->rp);
    s->d_col.upload(s->col);
    solver_layout(s);
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
    // post-BC pattern (ImposeDirichlet.h:27-51 keeps triplets with i,j not on
    // the boundary, then appends (b,b,1)), solver layout and assembly blocks:
    // on the device
    system_topology(s);
  } catch (...) {
    system_free(s);
    throw;
  }
  return s;
}

// A system with no mesh: the CSR pattern of a sparse matrix given by the
// caller (the general K\F op), on the device (rp, col are taken over),
// identity node order, no boundary rows.
System *system_from_device_csr(int n, DevBuf<int> &rp, DevBuf<int> &col, int64_t nnz) {
  REQUIRE(n > 0 && (int64_t)rp.n == n + 1, ADMFEM_ERR_ARG, "bad CSR pattern");
  System *s = new System();
  try {
    s->mesh = nullptr;
    s->n = n;
    s->npad = (n + 63) / 64 * 64;
    s->row_bd.assign(n, -1);
    s->nnz = nnz;
    s->d_rp.swap(rp);
    s->d_col.swap(col);
    if (!s->d_col.p) s->d_col.alloc(1);
    solver_layout(s);
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
