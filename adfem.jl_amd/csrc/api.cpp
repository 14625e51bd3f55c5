// C-ABI of libadmfem: argument checks, handle management, error codes, and
// the reference-named entry points on the default mesh.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "internal.h"

struct admfem_mesh_s {
  admfem::Mesh *m;
};
struct admfem_system_s {
  admfem::System *s;
  admfem::DevBuf<double> u_int, lam_int, g_int, tmp_int;
  admfem::DevBuf<double2> ul_int;  // (u, lam) pairs of the gradient call
  admfem::DevBuf<double> loss2;    // loss and index-error flag of forward_adjoint
};
struct admfem_server_s {
  admfem::Server *sv;
};

namespace admfem {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

template <class F>
static int guarded(F &&f) {
  try {
    f();
    return ADMFEM_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host out of memory");
    return ADMFEM_ERR_ARG;
  } catch (const std::exception &e) {
    set_error(e.what());
    return ADMFEM_ERR_ARG;
  }
}

static hipStream_t S(void *stream) { return (hipStream_t)stream; }

// The process-wide default mesh behind the reference-named symbols (the
// reference's `mmesh` singleton, deps/MFEM/Common.cpp:7).
static std::mutex g_default_mu;
static admfem_mesh g_default = nullptr;

static Mesh &M(admfem_mesh h) {
  REQUIRE(h && h->m, ADMFEM_ERR_ARG, "null mesh handle");
  return *h->m;
}
static System &SY(admfem_system h) {
  REQUIRE(h && h->s, ADMFEM_ERR_ARG, "null system handle");
  return *h->s;
}

template <class T>
static void d2h(T *dst, const T *src, size_t n) {
  if (n) HIPCHECK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace admfem

using namespace admfem;

extern "C" {

const char *admfem_last_error(void) { return g_err.c_str(); }
int admfem_version(void) { return 1; }

// ------------------------------------------------------------ B. meshes
int admfem_mesh_create(const double *coords, int nnode, const int *elems, int nelem,
                       int order, int degree, admfem_mesh *out) {
  return guarded([&] {
    REQUIRE(out, ADMFEM_ERR_ARG, "null out");
    *out = nullptr;
    REQUIRE(degree == 1 || degree == 2, ADMFEM_ERR_UNSUPPORTED,
            "only degree 1 (P1) and 2 (P2) elements are built");
    Mesh *m = mesh_build(coords, nnode, elems, nelem, order, degree);
    *out = new admfem_mesh_s{m};
  });
}

int admfem_mesh_destroy(admfem_mesh h) {
  return guarded([&] {
    if (!h) return;
    {
      std::lock_guard<std::mutex> lk(g_default_mu);
      if (h == g_default) g_default = nullptr;
    }
    delete h->m;
    delete h;
  });
}

int admfem_mesh_info(admfem_mesh h, int64_t info[8]) {
  return guarded([&] {
    Mesh &m = M(h);
    mesh_edges(m);
    info[0] = m.nnode;
    info[1] = m.nelem;
    info[2] = m.ngauss;
    info[3] = m.ndof;
    info[4] = m.nd;
    info[5] = m.nedges;
    info[6] = m.nnz_full;
    info[7] = m.rule.nq;
  });
}

int admfem_mesh_gauss(admfem_mesh h, double *x, double *y) {
  return guarded([&] {
    Mesh &m = M(h);
    d2h(x, m.gx.p, m.ngauss);
    d2h(y, m.gy.p, m.ngauss);
  });
}

int admfem_mesh_weights(admfem_mesh h, double *w) {
  return guarded([&] { d2h(w, M(h).gw.p, M(h).ngauss); });
}

int admfem_mesh_area(admfem_mesh h, double *a) {
  return guarded([&] { d2h(a, M(h).area.p, (size_t)M(h).nelem); });
}

int admfem_mesh_elems(admfem_mesh h, int *elems) {
  return guarded([&] {
    Mesh &m = M(h);
    std::copy(m.elems.begin(), m.elems.end(), elems);
  });
}

int admfem_mesh_dofs(admfem_mesh h, int *dofs) {
  return guarded([&] {
    Mesh &m = M(h);
    REQUIRE(dofs, ADMFEM_ERR_ARG, "null pointer");
    if (m.degree == 2) std::copy(m.dofs.begin(), m.dofs.end(), dofs);
    else std::copy(m.elems.begin(), m.elems.end(), dofs);
  });
}

int admfem_mesh_edges(admfem_mesh h, long long *edges) {
  return guarded([&] {
    Mesh &m = M(h);
    mesh_edges_host(m);
    std::copy(m.edges.begin(), m.edges.end(), edges);
  });
}

int admfem_mesh_bcnode(admfem_mesh h, long long *out, int *count) {
  return guarded([&] {
    Mesh &m = M(h);
    // nodes of the edges used by exactly one element (src/MFEM/MCore.jl:385-406,
    // 439-450), flagged by the edge pass
    mesh_edges(m);
    int c = 0;
    for (int u = 0; u < m.nnode; u++)
      if (m.bd_node[u]) out[c++] = u + 1;
    // P2: then the boundary edges' dofs, nnode + edge id (MCore.jl:441-443)
    for (size_t i = 0; i < m.bd_edge.size(); i++)
      if (m.bd_edge[i]) out[c++] = m.nnode + (long long)i + 1;
    *count = c;
  });
}

int admfem_mesh_pattern(admfem_mesh h, int *row_ptr, int *col) {
  return guarded([&] {
    Mesh &m = M(h);
    mesh_user_pattern(m);
    std::copy(m.urp.begin(), m.urp.end(), row_ptr);
    std::copy(m.ucol.begin(), m.ucol.end(), col);
  });
}

// ------------------------------------------------------------ C. operators
int admfem_fem_laplace_scalar(admfem_mesh h, const double *kappa, long long *indices,
                              double *vv, void *stream) {
  return guarded([&] {
    REQUIRE(kappa && indices && vv, ADMFEM_ERR_ARG, "null pointer");
    launch_laplace_coo(M(h), kappa, indices, vv, S(stream));
  });
}

int admfem_fem_laplace_scalar_grad(admfem_mesh h, const double *grad_vv,
                                   double *grad_kappa, void *stream) {
  return guarded([&] {
    REQUIRE(grad_vv && grad_kappa, ADMFEM_ERR_ARG, "null pointer");
    launch_laplace_coo_grad(M(h), grad_vv, grad_kappa, S(stream));
  });
}

int admfem_fem_mass_matrix(admfem_mesh h, const double *rho, long long *indices,
                           double *vv, void *stream) {
  return guarded([&] {
    REQUIRE(rho && indices && vv, ADMFEM_ERR_ARG, "null pointer");
    launch_mass_coo(M(h), rho, indices, vv, S(stream));
  });
}

int admfem_fem_mass_matrix_grad(admfem_mesh h, const double *grad_vv, double *grad_rho,
                                void *stream) {
  return guarded([&] {
    REQUIRE(grad_vv && grad_rho, ADMFEM_ERR_ARG, "null pointer");
    launch_mass_coo_grad(M(h), grad_vv, grad_rho, S(stream));
  });
}

int admfem_fem_source_scalar(admfem_mesh h, const double *f, double *rhs, void *stream) {
  return guarded([&] {
    REQUIRE(f && rhs, ADMFEM_ERR_ARG, "null pointer");
    launch_source(M(h), f, rhs, false, S(stream));
  });
}

int admfem_fem_source_scalar_grad(admfem_mesh h, const double *grad_rhs, double *grad_f,
                                  void *stream) {
  return guarded([&] {
    REQUIRE(grad_rhs && grad_f, ADMFEM_ERR_ARG, "null pointer");
    launch_source_grad(M(h), grad_rhs, grad_f, S(stream));
  });
}

int admfem_fem_to_gauss_points(admfem_mesh h, const double *u, double *out, void *stream) {
  return guarded([&] {
    REQUIRE(u && out, ADMFEM_ERR_ARG, "null pointer");
    launch_to_gauss(M(h), u, out, S(stream));
  });
}

int admfem_fem_to_gauss_points_grad(admfem_mesh h, const double *grad_out, double *grad_u,
                                    void *stream) {
  return guarded([&] {
    REQUIRE(grad_out && grad_u, ADMFEM_ERR_ARG, "null pointer");
    launch_to_gauss_grad(M(h), grad_out, grad_u, S(stream));
  });
}

int admfem_dof_to_gauss_points(admfem_mesh h, const double *u, double *out, void *stream) {
  return guarded([&] {
    REQUIRE(u && out, ADMFEM_ERR_ARG, "null pointer");
    launch_dof_to_gauss(M(h), u, out, S(stream));
  });
}

int admfem_dof_to_gauss_points_grad(admfem_mesh h, const double *grad_out, double *grad_u,
                                    void *stream) {
  return guarded([&] {
    REQUIRE(grad_out && grad_u, ADMFEM_ERR_ARG, "null pointer");
    launch_dof_to_gauss_grad(M(h), grad_out, grad_u, S(stream));
  });
}

int admfem_fem_grad(admfem_mesh h, const double *u, double *out, void *stream) {
  return guarded([&] {
    REQUIRE(u && out, ADMFEM_ERR_ARG, "null pointer");
    launch_fem_grad(M(h), u, out, S(stream));
  });
}

int admfem_fem_grad_grad(admfem_mesh h, const double *grad_out, double *grad_u, void *stream) {
  return guarded([&] {
    REQUIRE(grad_out && grad_u, ADMFEM_ERR_ARG, "null pointer");
    launch_fem_grad_grad(M(h), grad_out, grad_u, S(stream));
  });
}

int admfem_fem_laplace_term(admfem_mesh h, const double *u, const double *nu, double *out,
                            void *stream) {
  return guarded([&] {
    REQUIRE(u && nu && out, ADMFEM_ERR_ARG, "null pointer");
    launch_laplace_term(M(h), u, nu, out, S(stream));
  });
}

int admfem_fem_laplace_term_grad(admfem_mesh h, const double *grad_out, const double *u,
                                 const double *nu, double *grad_u, double *grad_nu,
                                 void *stream) {
  return guarded([&] {
    REQUIRE(grad_out && u && nu, ADMFEM_ERR_ARG, "null pointer");
    if (grad_u) launch_laplace_term(M(h), grad_out, nu, grad_u, S(stream));
    if (grad_nu) launch_laplace_term_nu_grad(M(h), grad_out, u, grad_nu, S(stream));
  });
}

int admfem_impose_dirichlet(const long long *indices, const double *vv, int64_t sN,
                            const long long *bd, const double *bdval, int bdN,
                            const double *rhs, int N, long long *oindices, double *ov,
                            double *orhs, int64_t *out_count, void *stream) {
  return guarded([&] {
    REQUIRE(out_count && orhs && rhs && (sN == 0 || (indices && vv)) &&
                (sN + bdN == 0 || (oindices && ov)) && (bdN == 0 || (bd && bdval)),
            ADMFEM_ERR_ARG, "null pointer");
    *out_count = run_impose_dirichlet(indices, vv, sN, bd, bdval, bdN, rhs, N, oindices, ov,
                                      orhs, S(stream));
  });
}

int admfem_impose_dirichlet_grad(const double *grad_ov, const double *grad_orhs,
                                 const long long *indices, const double *vv, int64_t sN,
                                 const long long *bd, const double *bdval, int bdN, int N,
                                 double *grad_vv, double *grad_rhs, double *grad_bdval,
                                 void *stream) {
  return guarded([&] {
    REQUIRE(grad_orhs && grad_rhs && (sN == 0 || (indices && vv && grad_vv)) &&
                (bdN == 0 || (bd && bdval && grad_bdval)),
            ADMFEM_ERR_ARG, "null pointer");
    run_impose_dirichlet_grad(grad_ov, grad_orhs, indices, vv, sN, bd, bdval, bdN, N,
                              grad_vv, grad_rhs, grad_bdval, S(stream));
  });
}

// ------------------------------------------------------------ D. systems
int admfem_system_create(admfem_mesh h, const long long *bdnode, int nbd,
                         admfem_system *out) {
  return guarded([&] {
    REQUIRE(out, ADMFEM_ERR_ARG, "null out");
    *out = nullptr;
    System *s = system_build(&M(h), bdnode, nbd);
    admfem_system hs = new admfem_system_s();
    hs->s = s;
    try {
      hs->u_int.alloc(s->npad);
      hs->lam_int.alloc(s->npad);
      hs->g_int.alloc(s->npad);
      hs->tmp_int.alloc(s->npad);
      hs->ul_int.alloc(s->npad);
      hs->loss2.alloc(2);
    } catch (...) {
      system_free(s);
      delete hs;
      throw;
    }
    *out = hs;
  });
}

int admfem_system_destroy(admfem_system h) {
  return guarded([&] {
    if (!h) return;
    system_free(h->s);
    delete h;
  });
}

int admfem_system_info(admfem_system h, int64_t info[4]) {
  return guarded([&] {
    System &s = SY(h);
    info[0] = s.n;
    info[1] = s.nnz;
    info[2] = s.nbd_unique;
    info[3] = (int64_t)(s.device_bytes() + s.mesh->device_bytes() +
                        6 * sizeof(double) * (size_t)s.npad);  // + the handle's vectors
  });
}

int admfem_system_pattern(admfem_system h, int *row_ptr, int *col) {
  return guarded([&] {
    System &s = SY(h);
    system_host_pattern(s);
    const std::vector<int> &P = s.uperm(), &IP = s.uiperm();
    int pos = 0;
    row_ptr[0] = 0;
    for (int u = 0; u < s.n; u++) {
      int r = P[u];
      int start = pos;
      for (int k = s.rp[r]; k < s.rp[r + 1]; k++) col[pos++] = IP[s.col[k]];
      std::sort(col + start, col + pos);
      row_ptr[u + 1] = pos;
    }
  });
}

int admfem_system_csr(admfem_system h, int *row_ptr, int *col, double *val, double *rhs) {
  return guarded([&] {
    System &s = SY(h);
    system_host_pattern(s);
    const std::vector<int> &P = s.uperm(), &IP = s.uiperm();
    // the assembled values live in the solver format: diagonal + SELL slices
    std::vector<double> sv(std::max(1, s.slice_ptr.back())), dg(s.npad), b(s.n);
    HIPCHECK(hipDeviceSynchronize());
    d2h(sv.data(), s.d_sell_val.p, sv.size());
    d2h(dg.data(), s.d_diag.p, dg.size());
    d2h(b.data(), s.rhs.p, b.size());
    int pos = 0;
    row_ptr[0] = 0;
    std::vector<std::pair<int, double>> row;
    for (int u = 0; u < s.n; u++) {
      int r = P[u];
      row.clear();
      const int off = s.slice_ptr[r >> 6] + (r & 63);
      int kk = 0;
      for (int k = s.rp[r]; k < s.rp[r + 1]; k++)
        row.push_back({IP[s.col[k]], s.col[k] == r ? dg[r] : sv[off + 64 * kk++]});
      std::sort(row.begin(), row.end(),
                [](const std::pair<int, double> &a, const std::pair<int, double> &c) {
                  return a.first < c.first;
                });
      for (auto &e : row) {
        col[pos] = e.first;
        val[pos] = e.second;
        pos++;
      }
      row_ptr[u + 1] = pos;
      if (rhs) rhs[u] = b[r];
    }
  });
}

static void assemble_impl(admfem_system h, const admfem_problem *p, hipStream_t st) {
  System &s = SY(h);
  REQUIRE(p, ADMFEM_ERR_ARG, "null problem");
  admfem_problem q = *p;
  if (p->u_prev) {  // node vector: user -> internal order
    launch_scatter(s.d_uperm(), p->u_prev, h->tmp_int.p, s.n, st);
    q.u_prev = h->tmp_int.p;
  }
  launch_assemble(s, q, st);
}

int admfem_system_assemble(admfem_system h, const admfem_problem *p, void *stream) {
  return guarded([&] { assemble_impl(h, p, S(stream)); });
}

int admfem_system_solve(admfem_system h, const double *b, double *x,
                        const admfem_solve_opts *o, admfem_solve_stats *st, void *stream) {
  return guarded([&] {
    System &s = SY(h);
    REQUIRE(x && o, ADMFEM_ERR_ARG, "null pointer");
    hipStream_t str = S(stream);
    const double *b_int = s.rhs.p;
    if (b) {
      launch_scatter(s.d_uperm(), b, h->g_int.p, s.n, str);
      b_int = h->g_int.p;
    }
    admfem_solve_stats tmp{};
    cg_solve(s, b_int, h->u_int.p, *o, tmp, str);
    launch_scatter(s.d_uiperm(), h->u_int.p, x, s.n, str);
    HIPCHECK(hipStreamSynchronize(str));
    if (st) *st = tmp;
    if (tmp.status != ADMFEM_OK) {
      set_error("CG did not converge (status " + std::to_string(tmp.status) + ")");
      throw Error{tmp.status, admfem_last_error()};
    }
  });
}

int admfem_system_kappa_gradient(admfem_system h, const double *u, const double *lam,
                                 double scale, double *grad_kappa, void *stream) {
  return guarded([&] {
    System &s = SY(h);
    REQUIRE(u && lam && grad_kappa, ADMFEM_ERR_ARG, "null pointer");
    hipStream_t str = S(stream);
    launch_kappa_gradient_user(s, u, lam, h->ul_int.p, scale, grad_kappa, str);
  });
}

int admfem_system_kappa_gradient_noalloc(admfem_system h, const double *u, const double *lam,
                                         double scale, double *grad_kappa, void *stream) {
  return guarded([&] {
    System &s = SY(h);
    REQUIRE(u && lam && grad_kappa, ADMFEM_ERR_ARG, "null pointer");
    kappa_gradient_noalloc(s, u, lam, scale, grad_kappa, S(stream));
  });
}

int admfem_system_forward_adjoint(admfem_system h, const admfem_problem *p,
                                  const int *obs_idx, const double *obs_val, int nobs,
                                  const admfem_solve_opts *o, double *u, double *lam,
                                  double *grad_kappa, double *loss,
                                  admfem_solve_stats st[2], void *stream) {
  return guarded([&] {
    System &s = SY(h);
    REQUIRE(o && (nobs == 0 || (obs_idx && obs_val)) && nobs >= 0, ADMFEM_ERR_ARG,
            "bad observation arguments");
    hipStream_t str = S(stream);
    assemble_impl(h, p, str);
    admfem_solve_stats s0{}, s1{};
    cg_solve(s, s.rhs.p, h->u_int.p, *o, s0, str);
    double L = loss_and_grad(s, h->u_int.p, obs_idx, obs_val, nobs, h->g_int.p, h->loss2.p,
                             str);
    cg_solve(s, h->g_int.p, h->lam_int.p, *o, s1, str);
    if (grad_kappa)
      launch_kappa_gradient(s, h->u_int.p, h->lam_int.p, p->lap_coef, grad_kappa, str);
    if (u) launch_scatter(s.d_uiperm(), h->u_int.p, u, s.n, str);
    if (lam) launch_scatter(s.d_uiperm(), h->lam_int.p, lam, s.n, str);
    HIPCHECK(hipStreamSynchronize(str));
    if (loss) *loss = L;
    if (st) {
      st[0] = s0;
      st[1] = s1;
    }
    int bad = s0.status != ADMFEM_OK ? s0.status : s1.status;
    if (bad != ADMFEM_OK) throw Error{bad, "CG did not converge in the forward or adjoint solve"};
  });
}

int admfem_sparse_solve(const long long *indices, const double *vv, int64_t nnz, int n,
                        const double *rhs, double *u, const admfem_solve_opts *o,
                        admfem_solve_stats *st, void *stream) {
  return guarded([&] {
    REQUIRE(rhs && u && o && (nnz == 0 || (indices && vv)), ADMFEM_ERR_ARG, "null pointer");
    admfem_solve_stats tmp{};
    sparse_solve(indices, vv, nnz, n, rhs, u, *o, tmp, S(stream));
    if (st) *st = tmp;
    if (tmp.status != ADMFEM_OK) throw Error{tmp.status, "sparse solve did not converge"};
  });
}

int admfem_sparse_solve_grad(const long long *indices, const double *vv, int64_t nnz, int n,
                             const double *u, const double *grad_u, double *grad_vv,
                             double *grad_rhs, const admfem_solve_opts *o,
                             admfem_solve_stats *st, void *stream) {
  return guarded([&] {
    REQUIRE(u && grad_u && grad_rhs && o && (nnz == 0 || (indices && vv && grad_vv)),
            ADMFEM_ERR_ARG, "null pointer");
    admfem_solve_stats tmp{};
    sparse_solve_grad(indices, vv, nnz, n, u, grad_u, grad_vv, grad_rhs, *o, tmp, S(stream));
    if (st) *st = tmp;
    if (tmp.status != ADMFEM_OK) throw Error{tmp.status, "adjoint solve did not converge"};
  });
}

int admfem_sparse_solve_info(int64_t info[4]) {
  return guarded([&] {
    REQUIRE(info, ADMFEM_ERR_ARG, "null pointer");
    sparse_solve_info(info);
  });
}

int admfem_system_spmv_repeat(admfem_system h, int precision, int repeat, double *bytes,
                              void *stream) {
  return guarded([&] {
    System &s = SY(h);
    REQUIRE(repeat >= 0, ADMFEM_ERR_ARG, "bad repeat");
    spmv_repeat(s, precision, repeat, S(stream));
    if (bytes) {
      const double vb = precision == ADMFEM_PREC_MIXED ? 4.0 : 8.0;
      const double nnz = (double)s.nnz, n = s.n;
      // one CG iteration of the canonical CSR form: the SpMV plus the
      // iteration's vector updates (6 n s)
      *bytes = nnz * (vb + 4.0) + 4.0 * (n + 1) + 2.0 * n * vb + 6.0 * n * vb;
    }
  });
}

// ------------------------------------------------------------ E. serving
int admfem_server_create(const admfem_server_opts *o, admfem_server *out) {
  return guarded([&] {
    REQUIRE(o && out, ADMFEM_ERR_ARG, "null pointer");
    *out = nullptr;
    Server *sv = server_create(*o);
    *out = new admfem_server_s{sv};
  });
}

int admfem_server_destroy(admfem_server h) {
  return guarded([&] {
    if (!h) return;
    server_free(h->sv);
    delete h;
  });
}

static Server &SV(admfem_server h) {
  REQUIRE(h && h->sv, ADMFEM_ERR_ARG, "null server handle");
  return *h->sv;
}

int admfem_server_add_topology(admfem_server h, admfem_system s, double jitter, int *topo_id) {
  return guarded([&] {
    REQUIRE(topo_id && jitter >= 0, ADMFEM_ERR_ARG, "bad argument");
    *topo_id = server_add_topology(&SV(h), &SY(s), jitter);
  });
}

int admfem_server_submit(admfem_server h, const admfem_request *r, int64_t *req_id) {
  return guarded([&] {
    REQUIRE(r && req_id, ADMFEM_ERR_ARG, "null pointer");
    *req_id = server_submit(&SV(h), *r);
  });
}

int admfem_server_run(admfem_server h, double timeout_s) {
  return guarded([&] { server_run(&SV(h), timeout_s); });
}

int admfem_server_result(admfem_server h, int64_t id, admfem_result *out) {
  return guarded([&] {
    REQUIRE(out, ADMFEM_ERR_ARG, "null pointer");
    server_result(&SV(h), id, *out);
  });
}

int admfem_server_stats(admfem_server h, double stats[8]) {
  return guarded([&] {
    REQUIRE(stats, ADMFEM_ERR_ARG, "null pointer");
    server_stats(&SV(h), stats);
  });
}

int admfem_server_spmv_repeat(admfem_server h, const int *topos, int nsys, int repeat,
                              double *bytes, double *seconds, void *stream) {
  return guarded([&] {
    REQUIRE(topos && nsys > 0 && repeat >= 0, ADMFEM_ERR_ARG, "bad arguments");
    double sec = 0.0;
    const double b = server_spmv_repeat(&SV(h), topos, nsys, repeat, S(stream), &sec);
    if (bytes) *bytes = b;
    if (seconds) *seconds = sec;
  });
}

// ------------------------------------------------- A. reference-named symbols
admfem_mesh admfem_default_mesh(void) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  return g_default;
}

long long *init_nnfem_mesh(double *vertices, int num_vertices, int *element_indices,
                           int num_elements, int order, int lorder, int degree,
                           long long *nedges) {
  (void)lorder;  // line-integral order: boundary-edge integrals are not built
  long long *ret = nullptr;
  int rc = guarded([&] {
    REQUIRE(vertices && element_indices && nedges, ADMFEM_ERR_ARG, "null pointer");
    REQUIRE(degree == 1 || degree == 2, ADMFEM_ERR_UNSUPPORTED,
            "only degree 1 (P1) and 2 (P2) elements are built");
    std::vector<double> c(2 * (size_t)num_vertices);
    for (int i = 0; i < num_vertices; i++) {
      c[2 * i] = vertices[3 * i];
      c[2 * i + 1] = vertices[3 * i + 1];
    }
    Mesh *m = mesh_build(c.data(), num_vertices, element_indices, num_elements, order, degree);
    admfem_mesh h = new admfem_mesh_s{m};
    {
      std::lock_guard<std::mutex> lk(g_default_mu);
      if (g_default) {
        printf("WARNING: Internal mesh is being overwritten!\n");  // API.cpp:7
        delete g_default->m;
        delete g_default;
      }
      g_default = h;
    }
    mesh_edges_host(*m);
    *nedges = m->nedges;
    ret = (long long *)malloc(sizeof(long long) * std::max<size_t>(1, m->edges.size()));
    REQUIRE(ret, ADMFEM_ERR_ARG, "host out of memory");
    std::copy(m->edges.begin(), m->edges.end(), ret);
  });
  if (rc != ADMFEM_OK) {
    fprintf(stderr, "init_nnfem_mesh: %s\n", admfem_last_error());
    return nullptr;
  }
  return ret;
}

static Mesh *default_or_warn(const char *fn) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    fprintf(stderr, "%s: no mesh (call init_nnfem_mesh first)\n", fn);
    return nullptr;
  }
  return g_default->m;
}

int mfem_get_ngauss(void) {
  Mesh *m = default_or_warn("mfem_get_ngauss");
  return m ? (int)m->ngauss : 0;
}

void mfem_get_gauss(double *x, double *y) {
  Mesh *m = default_or_warn("mfem_get_gauss");
  if (m) guarded([&] {
      d2h(x, m->gx.p, m->ngauss);
      d2h(y, m->gy.p, m->ngauss);
    });
}

void mfem_get_gauss_weights(double *w) {
  Mesh *m = default_or_warn("mfem_get_gauss_weights");
  if (m) guarded([&] { d2h(w, m->gw.p, m->ngauss); });
}

void mfem_get_area(double *a) {
  Mesh *m = default_or_warn("mfem_get_area");
  if (m) guarded([&] { d2h(a, m->area.p, (size_t)m->nelem); });
}

int mfem_get_elem_ndof(void) {
  Mesh *m = default_or_warn("mfem_get_elem_ndof");
  return m ? m->nd : 0;
}

int mfem_get_ndof(void) {
  Mesh *m = default_or_warn("mfem_get_ndof");
  return m ? (int)m->ndof : 0;
}

void mfem_get_connectivity(long long *conn) {
  Mesh *m = default_or_warn("mfem_get_connectivity");
  if (!m) return;
  const std::vector<int> &d = m->degree == 2 ? m->dofs : m->elems;
  for (size_t k = 0; k < d.size(); k++) conn[k] = d[k] + 1;
}

void mfem_get_element_to_vertices(long long *elems) {
  Mesh *m = default_or_warn("mfem_get_element_to_vertices");
  if (!m) return;
  for (int i = 0; i < m->nelem; i++)
    for (int k = 0; k < 3; k++) elems[(size_t)k * m->nelem + i] = m->elems[3 * (size_t)i + k] + 1;
}

void FemLaplaceScalar_forward_Julia(long long *indices, double *vv, const double *kappa) {
  Mesh *m = default_or_warn("FemLaplaceScalar_forward_Julia");
  if (!m) return;
  int rc = guarded([&] {
    size_t nt = (size_t)m->nd * m->nd * (size_t)m->ngauss;
    DevBuf<double> dk, dv;
    DevBuf<long long> di;
    dk.upload(kappa, m->ngauss);
    dv.alloc(nt);
    di.alloc(2 * nt);
    launch_laplace_coo(*m, dk.p, di.p, dv.p, nullptr);
    d2h(indices, di.p, 2 * nt);
    d2h(vv, dv.p, nt);
  });
  if (rc) fprintf(stderr, "FemLaplaceScalar_forward_Julia: %s\n", admfem_last_error());
}

void pcl_FemLaplaceScalar_Jacobian(double *H) {
  Mesh *m = default_or_warn("pcl_FemLaplaceScalar_Jacobian");
  if (!m) return;
  int rc = guarded([&] {
    // the triplet values are linear in kappa, each in one Gauss point's:
    // d vv[t] / d kappa[t / nd^2] = vv[t] at kappa = 1
    const size_t nd2 = (size_t)m->nd * m->nd, ng = (size_t)m->ngauss, nt = nd2 * ng;
    std::vector<double> one(ng, 1.0), vv(nt);
    DevBuf<double> dk, dv(nt);
    DevBuf<long long> di(2 * nt);
    dk.upload(one);
    launch_laplace_coo(*m, dk.p, di.p, dv.p, nullptr);
    d2h(vv.data(), dv.p, nt);
    for (size_t t = 0; t < nt; t++) H[t / nd2 + t * ng] = vv[t];
  });
  if (rc) fprintf(stderr, "pcl_FemLaplaceScalar_Jacobian: %s\n", admfem_last_error());
}

void pcl_ImposeDirichlet(double *J, const long long *indices, const long long *bd, int bdN,
                         int sN) {
  int rc = guarded([&] {
    REQUIRE(sN >= 0 && bdN >= 0 && (sN == 0 || (J && indices)) && (bdN == 0 || bd),
            ADMFEM_ERR_ARG, "bad arguments");
    long long nmax = 0;
    for (int k = 0; k < 2 * sN; k++) nmax = std::max(nmax, indices[k]);
    for (int k = 0; k < bdN; k++) nmax = std::max(nmax, bd[k]);
    std::vector<char> isbd((size_t)nmax + 1, 0);
    for (int k = 0; k < bdN; k++)
      if (bd[k] >= 1) isbd[bd[k]] = 1;  // 1-based
    size_t s = 0;
    for (int k = 0; k < sN; k++) {
      const long long i = indices[k], j = indices[k + sN];
      const bool ib = i >= 1 && isbd[i], jb = j >= 1 && isbd[j];
      if (!ib && !jb) J[(size_t)k + s++ * (size_t)sN] = 1.0;
    }
  });
  if (rc) fprintf(stderr, "pcl_ImposeDirichlet: %s\n", admfem_last_error());
}

void FemSourceScalar_forward_Julia(double *rhs, const double *f) {
  Mesh *m = default_or_warn("FemSourceScalar_forward_Julia");
  if (!m) return;
  int rc = guarded([&] {
    DevBuf<double> df, dr;
    df.upload(f, m->ngauss);
    dr.upload(rhs, m->ndof);
    launch_source(*m, df.p, dr.p, true, nullptr);
    d2h(rhs, dr.p, (size_t)m->ndof);
  });
  if (rc) fprintf(stderr, "FemSourceScalar_forward_Julia: %s\n", admfem_last_error());
}

void FemToGaussPointsMfem_Julia(double *out, const double *u) {
  Mesh *m = default_or_warn("FemToGaussPointsMfem_Julia");
  if (!m) return;
  int rc = guarded([&] {
    DevBuf<double> du, dout(m->ngauss);
    du.upload(u, m->nnode);
    launch_to_gauss(*m, du.p, dout.p, nullptr);
    d2h(out, dout.p, (size_t)m->ngauss);
  });
  if (rc) fprintf(stderr, "FemToGaussPointsMfem_Julia: %s\n", admfem_last_error());
}

void DofToGaussPointsMfem_forward_Julia(double *out, const double *u) {
  Mesh *m = default_or_warn("DofToGaussPointsMfem_forward_Julia");
  if (!m) return;
  int rc = guarded([&] {
    DevBuf<double> du, dout(m->ngauss);
    du.upload(u, m->ndof);
    launch_dof_to_gauss(*m, du.p, dout.p, nullptr);
    d2h(out, dout.p, (size_t)m->ngauss);
  });
  if (rc) fprintf(stderr, "DofToGaussPointsMfem_forward_Julia: %s\n", admfem_last_error());
}

void ComputeLaplaceTermMfem_forward_Julia(double *out, const double *nu, const double *u) {
  Mesh *m = default_or_warn("ComputeLaplaceTermMfem_forward_Julia");
  if (!m) return;
  int rc = guarded([&] {
    DevBuf<double> du, dnu, dk(m->ndof);
    du.upload(u, m->ndof);
    dnu.upload(nu, m->ngauss);
    launch_laplace_term(*m, du.p, dnu.p, dk.p, nullptr);
    std::vector<double> k(m->ndof);
    d2h(k.data(), dk.p, k.size());
    for (size_t i = 0; i < k.size(); i++) out[i] += k[i];  // accumulates, as the reference
  });
  if (rc) fprintf(stderr, "ComputeLaplaceTermMfem_forward_Julia: %s\n", admfem_last_error());
}

}  // extern "C"
