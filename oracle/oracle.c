/*
 * CPU oracle for the AdFem P1 forward/adjoint path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, always as the checker or the timed CPU baseline, never as product code.
 *
 * Every loop below restates the reference C++ loop named in its comment, kept
 * serial and in the reference's iteration order so that the triplet order,
 * the duplicate handling and the floating-point operation order match it.
 * The reference's Eigen temporaries become plain scalars; the arithmetic per
 * term is the same.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Triangle rules on the reference triangle (0,0),(1,0),(0,1), weights summing
 * to 1/2.  Order 0/1: centroid; order 2: the symmetric 3-point rule with
 * points (a,a),(a,b),(b,a), a=1/6, b=2/3, weight 1/6 each, in the order MFEM's
 * IntegrationRules::Get(TRIANGLE, 2) lists them (the rule the reference takes
 * at deps/MFEM/Common.cpp:42-44 with its default order 2, src/MFEM/MFEM.jl:73). */
int orc_tri_rule(int order, double *xi, double *eta, double *weight) {
  if (order == 0 || order == 1) {
    xi[0] = 1.0 / 3.0; eta[0] = 1.0 / 3.0; weight[0] = 0.5;
    return 1;
  }
  if (order == 2) {
    const double a = 1.0 / 6.0, b = 1.0 - 2.0 * a, wt = 1.0 / 6.0;
    xi[0] = a; eta[0] = a;
    xi[1] = a; eta[1] = b;
    xi[2] = b; eta[2] = a;
    weight[0] = weight[1] = weight[2] = wt;
    return 3;
  }
  if (order == 3 || order == 4) {
    /* the 6-point degree-4 rule (Strang & Fix; Dunavant 1985, rule 4): two
     * orbits (a,a),(a,b),(b,a), b = 1-2a, listed orbit by orbit in the same
     * point order as the 3-point rule (P2's default order 4,
     * src/MFEM/MFEM.jl:74-75) */
    const double a[2] = {0.091576213509770743459571463402202508,
                         0.44594849091596488631832925388305199};
    const double wt[2] = {0.054975871827660933819163162450105264,
                          0.11169079483900573284750350421656140};
    for (int o = 0; o < 2; o++) {
      const double b = 1.0 - 2.0 * a[o];
      xi[3 * o + 0] = a[o]; eta[3 * o + 0] = a[o];
      xi[3 * o + 1] = a[o]; eta[3 * o + 1] = b;
      xi[3 * o + 2] = b;    eta[3 * o + 2] = a[o];
      weight[3 * o] = weight[3 * o + 1] = weight[3 * o + 2] = wt[o];
    }
    return 6;
  }
  return -1;
}

/* deps/MFEM/Common.cpp:9-15; Eigen's norm() is the square root of the
 * squared norm. */
static double nrm(double dx, double dy) { return sqrt(dx * dx + dy * dy); }
static double heron(const double c[3][2]) {
  double a = nrm(c[0][0] - c[1][0], c[0][1] - c[1][1]);
  double b = nrm(c[2][0] - c[1][0], c[2][1] - c[1][1]);
  double cc = nrm(c[0][0] - c[2][0], c[0][1] - c[2][1]);
  double s = (a + b + cc) / 2.0;
  return sqrt(s * (s - a) * (s - b) * (s - cc));
}

/* deps/MFEM/Common.cpp:20-142, P1 branch. */
int orc_mesh_init(const double *coords, int nnode, int *elems, int nelem,
                  int order, double *gx, double *gy, double *w, double *area,
                  double *h, double *hx, double *hy) {
  (void)nnode;
  double qx[8], qy[8], qw[8];
  int nq = orc_tri_rule(order, qx, qy, qw);
  if (nq < 0) return -1;
  for (int e = 0; e < nelem; e++) {
    int *v = elems + 3 * e;
    /* FinalizeTriMesh(1, 0, true) (Common.cpp:40): CheckElementOrientation
     * swaps vertices 0 and 1 of a triangle whose Jacobian determinant is
     * negative. */
    {
      double x1 = coords[2 * v[0]], y1 = coords[2 * v[0] + 1];
      double x2 = coords[2 * v[1]], y2 = coords[2 * v[1] + 1];
      double x3 = coords[2 * v[2]], y3 = coords[2 * v[2] + 1];
      double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
      if (det < 0.0) { int t = v[0]; v[0] = v[1]; v[1] = t; }
    }
    double c[3][2];
    for (int k = 0; k < 3; k++) {
      c[k][0] = coords[2 * v[k]];
      c[k][1] = coords[2 * v[k] + 1];
    }
    double A = heron(c); /* Common.cpp:83 */
    area[e] = A;
    /* Jacobian of the affine map, columns d/dxi, d/deta; inverse for
     * CalcPhysDShape: dshape_phys = dshape_ref * J^{-1}. */
    double J00 = c[1][0] - c[0][0], J01 = c[2][0] - c[0][0];
    double J10 = c[1][1] - c[0][1], J11 = c[2][1] - c[0][1];
    double det = J00 * J11 - J01 * J10;
    double I00 = J11 / det, I01 = -J01 / det, I10 = -J10 / det, I11 = J00 / det;
    /* reference-element gradients of (1-xi-eta, xi, eta) */
    const double dref[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    for (int i = 0; i < nq; i++) {
      int g = nq * e + i;
      double hs0 = 1 - qx[i] - qy[i], hs1 = qx[i], hs2 = qy[i]; /* 106-108 */
      gx[g] = c[0][0] * hs0 + c[1][0] * hs1 + c[2][0] * hs2;     /* 110 */
      gy[g] = c[0][1] * hs0 + c[1][1] * hs1 + c[2][1] * hs2;     /* 111 */
      w[g] = qw[i] * A / 0.5;                                    /* 116 */
      h[3 * g + 0] = hs0; h[3 * g + 1] = hs1; h[3 * g + 2] = hs2;
      for (int r = 0; r < 3; r++) {
        hx[3 * g + r] = dref[r][0] * I00 + dref[r][1] * I10;
        hy[3 * g + r] = dref[r][0] * I01 + dref[r][1] * I11;
      }
    }
  }
  return nq;
}

/* P2 shape functions on the reference triangle in barycentric form
 * (l0 = 1-xi-eta, l1 = xi, l2 = eta): vertices l_k(2 l_k - 1), then the edge
 * midpoint functions 4 l0 l1, 4 l1 l2, 4 l2 l0 -- the nodal basis of MFEM's
 * H1_TriangleElement(2) (vertices, then edges (0,1),(1,2),(2,0)), with their
 * derivatives d/dxi, d/deta. */
static void p2_ref(double x, double y, double h[6], double d[6][2]) {
  const double l0 = 1.0 - x - y;
  h[0] = l0 * (2.0 * l0 - 1.0);
  h[1] = x * (2.0 * x - 1.0);
  h[2] = y * (2.0 * y - 1.0);
  h[3] = 4.0 * l0 * x;
  h[4] = 4.0 * x * y;
  h[5] = 4.0 * y * l0;
  d[0][0] = -(4.0 * l0 - 1.0); d[0][1] = -(4.0 * l0 - 1.0);
  d[1][0] = 4.0 * x - 1.0;     d[1][1] = 0.0;
  d[2][0] = 0.0;               d[2][1] = 4.0 * y - 1.0;
  d[3][0] = 4.0 * (l0 - x);    d[3][1] = -4.0 * x;
  d[4][0] = 4.0 * y;           d[4][1] = 4.0 * x;
  d[5][0] = -4.0 * y;          d[5][1] = 4.0 * (l0 - y);
}

/* deps/MFEM/Common.cpp:20-142, P2 branch (degree 2): elems must already be
 * orientation-fixed (orc_mesh_init); elem_edge[3e+j] is the MFEM id of the
 * element's edge j ((0,1),(1,2),(2,0), orc_mesh_edges).  dofs (nelem x 6):
 * the three vertices, then nnode + edge id (Common.cpp:63-72).  h/hx/hy
 * [ngauss*6] from CalcPhysShape / CalcPhysDShape (dshape_ref J^{-1}). */
int orc_mesh_init_p2(const double *coords, int nnode, const int *elems, int nelem,
                     int order, const int64_t *elem_edge, int *dofs, double *h,
                     double *hx, double *hy) {
  double qx[8], qy[8], qw[8];
  int nq = orc_tri_rule(order, qx, qy, qw);
  if (nq < 0) return -1;
  for (int e = 0; e < nelem; e++) {
    const int *v = elems + 3 * e;
    for (int k = 0; k < 3; k++) {
      dofs[6 * e + k] = v[k];
      dofs[6 * e + 3 + k] = nnode + (int)elem_edge[3 * e + k];
    }
    double c[3][2];
    for (int k = 0; k < 3; k++) {
      c[k][0] = coords[2 * v[k]];
      c[k][1] = coords[2 * v[k] + 1];
    }
    double J00 = c[1][0] - c[0][0], J01 = c[2][0] - c[0][0];
    double J10 = c[1][1] - c[0][1], J11 = c[2][1] - c[0][1];
    double det = J00 * J11 - J01 * J10;
    double I00 = J11 / det, I01 = -J01 / det, I10 = -J10 / det, I11 = J00 / det;
    for (int i = 0; i < nq; i++) {
      int g = nq * e + i;
      double hv[6], d[6][2];
      p2_ref(qx[i], qy[i], hv, d);
      for (int r = 0; r < 6; r++) {
        h[6 * g + r] = hv[r];
        hx[6 * g + r] = d[r][0] * I00 + d[r][1] * I10;
        hy[6 * g + r] = d[r][0] * I01 + d[r][1] * I11;
      }
    }
  }
  return nq;
}

typedef struct { int64_t key; int64_t ord; } keyord;
static int cmp_keyord(const void *a, const void *b) {
  const keyord *x = (const keyord *)a, *y = (const keyord *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->ord < y->ord ? -1 : (x->ord > y->ord);
}
static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : (x > y);
}

/* Common.cpp:133-140 on MFEM's edge numbering; elem_edge (optional,
 * 3*nelem) receives each element edge's id. */
int64_t orc_mesh_edges(const int *elems, int nelem, int nnode, int64_t *out,
                       int64_t *elem_edge) {
  static const int ev[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  int64_t M = 3 * (int64_t)nelem;
  keyord *ko = (keyord *)malloc(sizeof(keyord) * (M ? M : 1));
  for (int64_t e = 0; e < nelem; e++)
    for (int j = 0; j < 3; j++) {
      int64_t a = elems[3 * e + ev[j][0]], b = elems[3 * e + ev[j][1]];
      int64_t lo = a < b ? a : b, hi = a < b ? b : a;
      ko[3 * e + j].key = lo * (int64_t)nnode + hi;
      ko[3 * e + j].ord = 3 * e + j;
    }
  qsort(ko, M, sizeof(keyord), cmp_keyord);
  /* unique keys (groups) with their first occurrence, then order the groups
   * by first occurrence: a group's rank is its edge id */
  keyord *u = (keyord *)malloc(sizeof(keyord) * (M ? M : 1));
  int64_t *grp = (int64_t *)malloc(sizeof(int64_t) * (M ? M : 1));
  int64_t ne = 0;
  for (int64_t i = 0; i < M; i++) {
    if (i == 0 || ko[i].key != ko[i - 1].key) {
      u[ne].key = ko[i].ord; /* sort key: first occurrence */
      u[ne].ord = ne;        /* group number */
      ne++;
    }
    grp[i] = ne - 1;
  }
  int64_t *gkey = (int64_t *)malloc(sizeof(int64_t) * (ne ? ne : 1));
  for (int64_t i = 0; i < M; i++) gkey[grp[i]] = ko[i].key;
  qsort(u, ne, sizeof(keyord), cmp_keyord);
  int64_t *gid = (int64_t *)malloc(sizeof(int64_t) * (ne ? ne : 1));
  for (int64_t r = 0; r < ne; r++) gid[u[r].ord] = r;
  if (out)
    for (int64_t r = 0; r < ne; r++) {
      int64_t k = gkey[u[r].ord];
      out[r] = k / nnode + 1;
      out[ne + r] = k % nnode + 1;
    }
  if (elem_edge)
    for (int64_t i = 0; i < M; i++) elem_edge[ko[i].ord] = gid[grp[i]];
  free(ko); free(u); free(grp); free(gkey); free(gid);
  return ne;
}

/* FemLaplaceScalar.h:3-26: nd*nd triplets per Gauss point, p-major then q.
 * dofs nelem x nd (user dof ids); hx/hy [ngauss*nd]. */
void orc_laplace_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                         const double *hx, const double *hy,
                         const double *kappa, int64_t *indices, double *vv) {
  int64_t s = 0, nz = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int k = 0; k < nq; k++) {
      const double *dx = hx + nd * s, *dy = hy + nd * s;
      double kw = kappa[s], ww = w[s];
      s++;
      for (int p = 0; p < nd; p++)
        for (int q = 0; q < nd; q++) {
          /* N = D * D^T * kappa * w, evaluated as Eigen does: the product
           * first, then the two scalar factors. */
          double d = dx[p] * dx[q] + dy[p] * dy[q];
          indices[2 * nz] = dofs[nd * i + p];
          indices[2 * nz + 1] = dofs[nd * i + q];
          vv[nz] = d * kw * ww;
          nz++;
        }
    }
}

/* FemLaplaceScalar.h:28-54 */
void orc_laplace_backward(int nd, int nelem, int nq, const double *w, const double *hx,
                          const double *hy, const double *grad_vv,
                          double *grad_kappa) {
  int64_t s = 0, nz = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int k = 0; k < nq; k++) {
      const double *dx = hx + nd * s, *dy = hy + nd * s;
      double v = 0.0;
      for (int p = 0; p < nd; p++)
        for (int q = 0; q < nd; q++) {
          double d = (dx[p] * dx[q] + dy[p] * dy[q]) * w[s];
          v += grad_vv[nz] * d;
          nz++;
        }
      grad_kappa[s] = v;
      s++;
    }
}

/* ComputeFemMassMatrixMfem.h:4-29 */
void orc_mass_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                      const double *h, const double *rho, int64_t *indices,
                      double *vv) {
  int64_t k = 0, k0 = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      const double *N = h + nd * k0;
      double r = rho[k0], ww = w[k0];
      k0++;
      for (int l = 0; l < nd; l++)
        for (int s = 0; s < nd; s++) {
          indices[2 * k] = dofs[nd * i + l];
          indices[2 * k + 1] = dofs[nd * i + s];
          vv[k] = N[l] * N[s] * r * ww;
          k++;
        }
    }
}

/* ComputeFemMassMatrixMfem.h:31-56 */
void orc_mass_backward(int nd, int nelem, int nq, const double *w, const double *h,
                       const double *grad_vv, double *grad_rho) {
  int64_t k = 0, k0 = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      const double *N = h + nd * k0;
      for (int l = 0; l < nd; l++)
        for (int s = 0; s < nd; s++) {
          grad_rho[k0] += N[l] * N[s] * w[k0] * grad_vv[k];
          k++;
        }
      k0++;
    }
}

/* FemSourceScalar.h:4-15 */
void orc_source_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                        const double *h, const double *f, double *rhs) {
  int64_t k = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      for (int r = 0; r < nd; r++)
        rhs[dofs[nd * i + r]] += f[k] * h[nd * k + r] * w[k];
      k++;
    }
}

/* FemSourceScalar.h:17-32 */
void orc_source_backward(const int *dofs, int nd, int nelem, int nq, const double *w,
                         const double *h, const double *grad_rhs,
                         double *grad_f) {
  int64_t k = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      grad_f[k] = 0.0;
      for (int r = 0; r < nd; r++)
        grad_f[k] += h[nd * k + r] * w[k] * grad_rhs[dofs[nd * i + r]];
      k++;
    }
}

/* std::map<int64,double> bdMap of the reference as a sorted unique key array
 * (last assignment wins), with binary-search lookups. */
typedef struct { int64_t *key; double *val; int64_t *idx; int n; } bdmap;
static bdmap bdmap_build(const int64_t *bd, const double *bdval, int bdN) {
  bdmap m;
  keyord *ko = (keyord *)malloc(sizeof(keyord) * (bdN ? bdN : 1));
  for (int i = 0; i < bdN; i++) { ko[i].key = bd[i] - 1; ko[i].ord = i; }
  qsort(ko, bdN, sizeof(keyord), cmp_keyord);
  m.key = (int64_t *)malloc(sizeof(int64_t) * (bdN ? bdN : 1));
  m.val = (double *)malloc(sizeof(double) * (bdN ? bdN : 1));
  m.idx = (int64_t *)malloc(sizeof(int64_t) * (bdN ? bdN : 1));
  m.n = 0;
  for (int i = 0; i < bdN; i++) {
    if (m.n > 0 && m.key[m.n - 1] == ko[i].key) m.n--; /* later one wins */
    m.key[m.n] = ko[i].key;
    m.val[m.n] = bdval ? bdval[ko[i].ord] : 0.0;
    m.idx[m.n] = ko[i].ord;
    m.n++;
  }
  free(ko);
  return m;
}
static int bdmap_find(const bdmap *m, int64_t k) {
  int lo = 0, hi = m->n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (m->key[mid] == k) return mid;
    if (m->key[mid] < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}
static void bdmap_free(bdmap *m) { free(m->key); free(m->val); free(m->idx); }

/* ImposeDirichlet.h:27-60 */
int64_t orc_dirichlet_forward(const int64_t *indices, const double *vv,
                              int64_t sN, const int64_t *bd, const double *bdval,
                              int bdN, const double *rhs, int N,
                              int64_t *oindices, double *ov, double *orhs) {
  bdmap m = bdmap_build(bd, bdval, bdN);
  for (int i = 0; i < N; i++) orhs[i] = rhs[i];
  int64_t z = 0;
  for (int64_t k = 0; k < sN; k++) {
    int64_t i = indices[2 * k], j = indices[2 * k + 1];
    int bi = bdmap_find(&m, i), bj = bdmap_find(&m, j);
    if (bi < 0 && bj < 0) {
      oindices[2 * z] = i; oindices[2 * z + 1] = j; ov[z] = vv[k]; z++;
    }
    if (bi < 0 && bj >= 0) orhs[i] = orhs[i] - vv[k] * m.val[bj];
  }
  for (int t = 0; t < m.n; t++) {
    oindices[2 * z] = m.key[t]; oindices[2 * z + 1] = m.key[t]; ov[z] = 1.0; z++;
    orhs[m.key[t]] = m.val[t];
  }
  bdmap_free(&m);
  return z;
}

/* ImposeDirichlet.h:63-93 */
void orc_dirichlet_backward(double *grad_vv_ipt, double *grad_rhs_ipt,
                            double *grad_bdval, const double *grad_ov,
                            const double *grad_orhs, const int64_t *indices,
                            const double *vv, int64_t sN, const int64_t *bd,
                            const double *bdval, int bdN, int N) {
  memset(grad_vv_ipt, 0, sizeof(double) * sN);
  memset(grad_rhs_ipt, 0, sizeof(double) * N);
  memset(grad_bdval, 0, sizeof(double) * bdN);
  bdmap m = bdmap_build(bd, bdval, bdN);
  int64_t z = 0;
  for (int64_t k = 0; k < sN; k++) {
    int64_t i = indices[2 * k], j = indices[2 * k + 1];
    int bi = bdmap_find(&m, i), bj = bdmap_find(&m, j);
    if (bi < 0 && bj < 0) grad_vv_ipt[k] = grad_ov[z++];
    if (bi < 0 && bj >= 0) {
      grad_vv_ipt[k] -= m.val[bj] * grad_orhs[i];
      grad_bdval[m.idx[bj]] -= vv[k] * grad_orhs[i];
    }
  }
  for (int i = 0; i < N; i++)
    if (bdmap_find(&m, i) < 0) grad_rhs_ipt[i] = grad_orhs[i];
  for (int t = 0; t < m.n; t++) grad_bdval[m.idx[t]] += grad_orhs[m.key[t]];
  bdmap_free(&m);
}

int64_t orc_coo_to_csr(const int64_t *indices, const double *vv, int64_t sN,
                       int N, int64_t *row_ptr, int32_t *col, double *val) {
  int64_t *cnt = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
  for (int64_t k = 0; k < sN; k++) cnt[indices[2 * k] + 1]++;
  for (int i = 0; i < N; i++) cnt[i + 1] += cnt[i];
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * ((size_t)N + 1));
  memcpy(pos, cnt, sizeof(int64_t) * ((size_t)N + 1));
  int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (sN ? sN : 1));
  for (int64_t k = 0; k < sN; k++) perm[pos[indices[2 * k]]++] = k; /* stable */
  int64_t nnz = 0;
  row_ptr[0] = 0;
  keyord *buf = NULL;
  int64_t cap = 0;
  for (int i = 0; i < N; i++) {
    int64_t a = cnt[i], b = cnt[i + 1], n = b - a;
    if (n > cap) { cap = n; buf = (keyord *)realloc(buf, sizeof(keyord) * cap); }
    for (int64_t t = 0; t < n; t++) {
      buf[t].key = indices[2 * perm[a + t] + 1];
      buf[t].ord = perm[a + t];
    }
    qsort(buf, n, sizeof(keyord), cmp_keyord); /* by column, then triplet order */
    for (int64_t t = 0; t < n; t++) {
      if (t > 0 && buf[t].key == buf[t - 1].key) {
        val[nnz - 1] += vv[buf[t].ord];
      } else {
        col[nnz] = (int32_t)buf[t].key;
        val[nnz] = vv[buf[t].ord];
        nnz++;
      }
    }
    row_ptr[i + 1] = nnz;
  }
  free(buf); free(perm); free(pos); free(cnt);
  return nnz;
}

/* src/MFEM/MCore.jl:385-406 (bcedge: an edge toggled an odd number of times
 * is a boundary edge) and 439-450 (bcnode: the set of their vertices). */
int orc_bcnode(const int *elems, int nelem, int nnode, int64_t *out) {
  static const int ev[3][2] = {{0, 1}, {2, 1}, {0, 2}};
  int64_t M = 3 * (int64_t)nelem;
  int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (M ? M : 1));
  for (int64_t e = 0; e < nelem; e++)
    for (int j = 0; j < 3; j++) {
      int64_t a = elems[3 * e + ev[j][0]], b = elems[3 * e + ev[j][1]];
      keys[3 * e + j] = (a < b ? a : b) * (int64_t)nnode + (a < b ? b : a);
    }
  qsort(keys, M, sizeof(int64_t), cmp_i64);
  char *isbd = (char *)calloc((size_t)nnode, 1);
  for (int64_t i = 0; i < M;) {
    int64_t j = i;
    while (j < M && keys[j] == keys[i]) j++;
    if ((j - i) % 2 == 1) { isbd[keys[i] / nnode] = 1; isbd[keys[i] % nnode] = 1; }
    i = j;
  }
  int n = 0;
  for (int v = 0; v < nnode; v++)
    if (isbd[v]) out[n++] = v + 1;
  free(isbd); free(keys);
  return n;
}

/* FemToGaussPointsMfem.h:6-18 (forward) and 20-33 (backward) with the
 * vertex functions hs (P1 h, nd = 3); DofToGaussPointsMfem.h:6-32 is the
 * same loop over the element's dofs with h (nd = 3 or 6). */
void orc_fem_to_gauss_forward(const int *dofs, int nd, int nelem, int nq, const double *h,
                              const double *u, double *out) {
  int64_t k = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      double v = 0.0;
      for (int r = 0; r < nd; r++) v += u[dofs[nd * i + r]] * h[nd * k + r];
      out[k] = v;
      k++;
    }
}

void orc_fem_to_gauss_backward(const int *dofs, int nd, int nelem, int nq, const double *h,
                               const double *grad_out, double *grad_u) {
  int64_t k = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      for (int r = 0; r < nd; r++) grad_u[dofs[nd * i + r]] += grad_out[k] * h[nd * k + r];
      k++;
    }
}

/* FemGradMfem.h:8-22 (forward) and 24-41 (backward). */
void orc_fem_grad_forward(const int *dofs, int nd, int nelem, int nq, const double *hx,
                          const double *hy, const double *u, double *out) {
  int64_t k = 0, g = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++, g++) {
      out[k] = 0.0;
      for (int r = 0; r < nd; r++) out[k] += u[dofs[nd * i + r]] * hx[nd * g + r];
      k++;
      out[k] = 0.0;
      for (int r = 0; r < nd; r++) out[k] += u[dofs[nd * i + r]] * hy[nd * g + r];
      k++;
    }
}

void orc_fem_grad_backward(const int *dofs, int nd, int nelem, int nq, const double *hx,
                           const double *hy, const double *grad_out, double *grad_u) {
  int64_t k = 0, g = 0;
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++, g++) {
      for (int r = 0; r < nd; r++) grad_u[dofs[nd * i + r]] += grad_out[k] * hx[nd * g + r];
      k++;
      for (int r = 0; r < nd; r++) grad_u[dofs[nd * i + r]] += grad_out[k] * hy[nd * g + r];
      k++;
    }
}

/* ComputeLaplaceTermMfem.h:4-17 (forward) and 19-39 (backward). */
void orc_laplace_term_forward(const int *dofs, int nd, int nelem, int nq, const double *w,
                              const double *hx, const double *hy, const double *nu,
                              const double *u, double *out) {
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      const int64_t g = (int64_t)nq * i + j;
      const double nv = nu[g];
      for (int p = 0; p < nd; p++)
        for (int q = 0; q < nd; q++)
          out[dofs[nd * i + p]] +=
              nv * (hx[nd * g + p] * hx[nd * g + q] * u[dofs[nd * i + q]] * w[g] +
                    hy[nd * g + p] * hy[nd * g + q] * u[dofs[nd * i + q]] * w[g]);
    }
}

void orc_laplace_term_backward(const int *dofs, int nd, int nelem, int nq, const double *w,
                               const double *hx, const double *hy, const double *nu,
                               const double *u, const double *grad_out, double *grad_nu,
                               double *grad_u) {
  for (int64_t i = 0; i < nelem; i++)
    for (int j = 0; j < nq; j++) {
      const int64_t g = (int64_t)nq * i + j;
      const double nv = nu[g];
      for (int p = 0; p < nd; p++)
        for (int q = 0; q < nd; q++) {
          const double go = grad_out[dofs[nd * i + p]];
          grad_nu[g] += go * (hx[nd * g + p] * hx[nd * g + q] * u[dofs[nd * i + q]] * w[g] +
                              hy[nd * g + p] * hy[nd * g + q] * u[dofs[nd * i + q]] * w[g]);
          grad_u[dofs[nd * i + q]] += go * nv * hx[nd * g + p] * hx[nd * g + q] * w[g];
          grad_u[dofs[nd * i + q]] += go * nv * hy[nd * g + p] * hy[nd * g + q] * w[g];
        }
    }
}
