// Internal structures of libadmfem (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "../../include/admfem.h"

namespace admfem {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
struct Error {
  int code;
  std::string msg;
};
#define HIPCHECK(expr)                                                        \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      throw ::admfem::Error{ADMFEM_ERR_HIP, std::string(#expr) + ": " +       \
                                                hipGetErrorString(_e)};       \
  } while (0)
#define REQUIRE(cond, code, msg)                                              \
  do {                                                                        \
    if (!(cond)) throw ::admfem::Error{code, msg};                            \
  } while (0)

// Device buffer (RAII over hipMalloc/hipFree).
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIPCHECK(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) HIPCHECK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Packed per-incidence slot word: bits 0-1 local vertex p of the row node in
// the element, then three 6-bit slots (offset of column conn[e][q] inside the
// row, q = 0,1,2).  SLOT_LIFT marks a boundary column whose contribution is
// moved to the right-hand side (Dirichlet lifting).
constexpr uint32_t SLOT_BITS = 6;
constexpr uint32_t SLOT_LIFT = 63;
constexpr int MAX_ROW = 62;
__host__ __device__ inline uint32_t slot_of(uint32_t w, int q) {
  return (w >> (2 + SLOT_BITS * q)) & 63u;
}

// Assembly pair word of a (row, incident element) pair of a post-BC system:
// bits 0-1 the row node's local vertex p, then three 6-bit codes for the
// element's vertices q = 0,1,2 -- the index of column conn[e][q] among the
// row's off-diagonal entries (its SELL position), PAIR_DIAG for the row
// itself or SLOT_LIFT for a boundary column -- and bits 20-31 the element's
// index in its assembly block's element list.
constexpr uint32_t PAIR_DIAG = 62;
constexpr int PAIR_MAX_ELEMS = 4096;
// P2 pair word (64 bits): bits 0-2 the row dof's local index p in the
// element, six 6-bit codes (as above) for the element's dofs q = 0..5 from
// bit 3, and the element's index in its block's element list from bit 40.
constexpr int P2_PAIR_CODE0 = 3;
constexpr int P2_PAIR_ELEM = 40;
// An assembly block's element record is {element id, its 3 vertices as
// 10-bit indices into the block's sorted list of distinct vertices}.
constexpr int BLK_VERT_BITS = 10;
constexpr int BLK_MAX_VERTS = 1 << BLK_VERT_BITS;
// Mixed-precision CG: default residual replacement factor (a replacement
// when the FP32 residual fell by this factor since the last one) and the
// stall guard (a replacement when the FP32 residual has not fallen by 1%
// for this many iterations).
constexpr double REPLACE_DELTA = 0.3;
constexpr int STALL_ITERS = 64;
constexpr int AS_ROWS = 64;  // rows per assembly block (one SELL slice)
constexpr int MAXQ = 6;  // quadrature points per element carried here
struct QuadRule {
  int nq;
  double xi[MAXQ], eta[MAXQ], w[MAXQ];  // reference triangle, weights sum 1/2
};
int quad_rule(int order, QuadRule &r);

// ------------------------------------------------------- host parallelism
// Runs f(begin, end, chunk) over [0, n) in contiguous chunks, one per host
// thread (at most 16, at least `grain` items each; chunk c covers
// [c*n/nc, (c+1)*n/nc)).  Returns the number of chunks.  An exception thrown
// by any chunk is rethrown here after every thread has joined.
template <class F>
int parallel_chunks(int64_t n, int64_t grain, F &&f) {
  const int hw = std::max(1u, std::thread::hardware_concurrency());
  const int nc = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(hw, 16), n / std::max<int64_t>(grain, 1)));
  if (nc == 1) {
    f((int64_t)0, n, 0);
    return 1;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nc);
  for (int c = 0; c < nc; c++)
    th.emplace_back([&, c] {
      try {
        f(n * c / nc, n * (c + 1) / nc, c);
      } catch (...) {
        err[c] = std::current_exception();
      }
    });
  for (auto &t : th) t.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
  return nc;
}
constexpr int64_t PAR_GRAIN = 16384;  // items per host thread, at least

// --------------------------------------------------------------------- mesh
struct Mesh {
  int nnode = 0, nelem = 0, order = 2;
  int degree = 1, nd = 3;  // element degree and dofs per element (3 or 6)
  int64_t ndof = 0;        // nnode (P1) or nnode + nedges (P2)
  QuadRule rule{};
  int64_t ngauss = 0;
  // host copies
  std::vector<int> elems;       // oriented, user node ids, nelem x 3
  std::vector<int> perm;        // user node id -> internal id
  std::vector<int> iperm;       // internal id -> user node id
  int64_t nnz_full = 0;         // entries of the pre-BC pattern
  std::vector<int> urp, ucol;   // pre-BC pattern, user numbering (rows sorted),
                                // made on request (mesh_user_pattern)
  int64_t nedges = -1;
  std::vector<long long> edges;  // host copy of d_edges, made on request (mesh_edges_host)
  DevBuf<long long> d_edges;     // 2*nedges, 1-based user ids: [min ...; max ...]
  std::vector<char> bd_node;     // built with edges: node on a boundary edge
  std::vector<char> bd_edge;     // built with edges (P2): edge used by one element
  std::vector<int> elem_edge;    // P2: MFEM id of element edge j at 3e+j
  std::vector<int> dofs;         // P2: nelem x 6 user dof ids
  // device
  DevBuf<double> x, y;          // internal order
  DevBuf<double2> xy;           // internal order, interleaved: one 16-byte gather per vertex
  DevBuf<int> conn;             // [3][nelem] internal ids (SoA)
  DevBuf<int> conn_user;        // [3][nelem] user ids (SoA)
  DevBuf<int> d_perm, d_iperm;
  DevBuf<int> d_n2e_ptr, d_n2e_elem;  // internal row -> incident elements (ascending)
  DevBuf<uint32_t> d_slots_full;  // per incidence, slots in the pre-BC row
  DevBuf<int> d_rp, d_col;        // pre-BC pattern, internal numbering
  DevBuf<double> gx, gy, gw, area;  // geometry (user element order)
  DevBuf<int> conn_dof;             // P2: [6][nelem] user dof ids (SoA)
  DevBuf<int> d_d2e_ptr, d_d2e;     // P2: dof -> (element << 3 | local dof), ascending
  size_t device_bytes() const;
};

Mesh *mesh_build(const double *coords, int nnode, const int *elems, int nelem,
                 int order, int degree = 1);
// topology.hip: the MFEM edge table on the device (m.nedges, m.d_edges,
// m.bd_node; P2: m.elem_edge, m.bd_edge), and its host copy m.edges
void mesh_edges(Mesh &m);
void mesh_edges_host(Mesh &m);
// topology.hip: node order (host perm / iperm) and the device topology
// (connectivity, incidences, pre-BC pattern, slot words, coordinates) of a
// mesh whose elems are set and oriented; coords in user order (host)
void mesh_topology(Mesh &m, const double *coords, hipStream_t st);
void mesh_user_pattern(Mesh &m);

// ------------------------------------------------------------------ system
struct CGWork;
struct System {
  Mesh *mesh = nullptr;
  int degree = 1;  // the mesh's element degree
  int n = 0;       // rows (P1: nnode, P2: ndof)
  int npad = 0;    // rows padded to a multiple of 64
  int nbd = 0;     // entries of bdnode as given
  int nbd_unique = 0;
  std::vector<int> row_bd;   // internal row -> index into bdval (-1 interior)
  int64_t nnz = 0;           // entries of the post-BC pattern
  // post-BC pattern, internal numbering (host copy of d_rp / d_col; for a
  // P1 mesh system made on request, system_host_pattern)
  std::vector<int> rp, col;
  // device: pattern + assembly maps
  DevBuf<int> d_rp, d_col, d_row_bd;
  // assembly: per block of AS_ROWS rows the distinct elements incident to
  // its interior rows (ascending) as {element, 3 local vertex indices}, the
  // distinct vertices of those elements (ascending internal ids), and a pair
  // word per incidence (row-major, parallel to the mesh's n2e lists)
  DevBuf<int> d_blk_eptr;
  DevBuf<int2> d_blk_rec;
  DevBuf<int> d_blk_vptr, d_blk_vert;
  DevBuf<uint32_t> d_pairs;
  int asm_max_e = 0;   // most elements of one block
  int asm_max_v = 0;   // most vertices of one block
  int sell_wmax = 0;   // widest SELL slice
  // CSR values (general sparse systems only) and the right-hand side
  DevBuf<double> val, rhs;
  // solver format: SELL-64 off-diagonals + separate diagonal
  int nslices = 0;
  std::vector<int> slice_ptr;  // offsets (elements) of each slice
  DevBuf<int> d_slice_ptr;
  DevBuf<int> d_sell_col;      // absolute internal column ids
  DevBuf<int16_t> d_sell_d16;  // the same as 16-bit offsets from the row (empty
                               // when an offset does not fit)
  DevBuf<double> d_sell_val;   // FP64 values
  DevBuf<double> d_diag;
  // P2: the internal dof order (each vertex dof at the mesh's node position,
  // followed by the edge dofs whose lower-numbered end it is), the element
  // dofs as internal rows ([6][nelem]), per row its incidences, and per
  // assembly block its element ids and a 64-bit pair word per incidence
  std::vector<int> perm2, iperm2;  // user dof id <-> internal row
  DevBuf<int> d_perm2, d_iperm2, d_conn2, d_inc_ptr, d_blk_elem;
  DevBuf<uint64_t> d_pairs2;
  // P2, for per-request inputs: the row of each vertex (mesh-internal vertex
  // id -> row) and the end vertices of each row's dof (a vertex dof: twice
  // its vertex; an edge dof: the edge's ends), mesh-internal ids
  DevBuf<int> d_vrow;
  DevBuf<int2> d_dof_ends;
  CGWork *work = nullptr;
  size_t device_bytes() const;
  // user node / dof id -> internal row, and back (mesh systems)
  const std::vector<int> &uperm() const { return degree == 2 ? perm2 : mesh->perm; }
  const std::vector<int> &uiperm() const { return degree == 2 ? iperm2 : mesh->iperm; }
  const int *d_uperm() const { return degree == 2 ? d_perm2.p : mesh->d_perm.p; }
  const int *d_uiperm() const { return degree == 2 ? d_iperm2.p : mesh->d_iperm.p; }
};

System *system_build(Mesh *m, const long long *bdnode, int nbd);
void system_free(System *s);
// topology.hip: solver layout from d_rp / d_col (and s->nnz); the P1
// system's pattern, layout and assembly blocks from s->row_bd; host copy of
// the pattern
void solver_layout(System *s);
void system_topology(System *s);
void system_host_pattern(System &s);
// topology.hip: graph helpers over a device CSR pattern
void csr_row_ptr(const int *srow, int m, int n, int *rp, hipStream_t st);  // rows sorted
int64_t device_bandwidth(int n, const int *rp, const int *col, const int *inv, hipStream_t st);
int device_first_coupled(int n, const int *rp, const int *col, hipStream_t st);
int64_t device_bfs_max_level(int n, const int *rp, const int *col, int s, hipStream_t st);
void rcm(int n, const int *rp, const int *col, DevBuf<int> &order, DevBuf<int> &inv,
         hipStream_t st);

// ------------------------------------------------------------- launchers
// geometry.hip
void launch_geometry(const Mesh &m, hipStream_t st);
void launch_laplace_coo(const Mesh &m, const double *kappa, long long *ind,
                        double *vv, hipStream_t st);
void launch_laplace_coo_grad(const Mesh &m, const double *gvv, double *gk,
                             hipStream_t st);
void launch_mass_coo(const Mesh &m, const double *rho, long long *ind,
                     double *vv, hipStream_t st);
void launch_mass_coo_grad(const Mesh &m, const double *gvv, double *gr,
                          hipStream_t st);
void launch_source(const Mesh &m, const double *f, double *rhs_user,
                   bool accumulate, hipStream_t st);
void launch_source_grad(const Mesh &m, const double *grhs_user, double *gf,
                        hipStream_t st);
// fields.hip (user node order / user Gauss order)
void launch_to_gauss(const Mesh &m, const double *u, double *out, hipStream_t st);
void launch_to_gauss_grad(const Mesh &m, const double *grad_out, double *grad_u,
                          hipStream_t st);
void launch_fem_grad(const Mesh &m, const double *u, double *out, hipStream_t st);
void launch_fem_grad_grad(const Mesh &m, const double *grad_out, double *grad_u,
                          hipStream_t st);
void launch_laplace_term(const Mesh &m, const double *u, const double *nu, double *out,
                         hipStream_t st);
void launch_laplace_term_nu_grad(const Mesh &m, const double *grad_out, const double *u,
                                 double *grad_nu, hipStream_t st);
// DofToGaussPoints (DofToGaussPointsMfem.h:6-32): the degree's basis (for P1
// the vertex functions, as launch_to_gauss).
void launch_dof_to_gauss(const Mesh &m, const double *u, double *out, hipStream_t st);
void launch_dof_to_gauss_grad(const Mesh &m, const double *grad_out, double *grad_u,
                              hipStream_t st);
// P2 kernels (p2.hip); the generic launchers above dispatch to them.
void launch_p2_coo(const Mesh &m, bool mass, const double *coef, long long *ind, double *vv,
                   hipStream_t st);
void launch_p2_coo_grad(const Mesh &m, bool mass, const double *gvv, double *out,
                        hipStream_t st);
void launch_p2_source(const Mesh &m, const double *f, double *rhs, bool accumulate,
                      hipStream_t st);
void launch_p2_source_grad(const Mesh &m, const double *grhs, double *gf, hipStream_t st);
void launch_p2_dof_to_gauss(const Mesh &m, const double *u, double *out, hipStream_t st);
void launch_p2_dof_to_gauss_grad(const Mesh &m, const double *g, double *gu, hipStream_t st);
void launch_p2_fem_grad(const Mesh &m, const double *u, double *out, hipStream_t st);
void launch_p2_fem_grad_grad(const Mesh &m, const double *g, double *gu, hipStream_t st);
void launch_p2_laplace_term(const Mesh &m, const double *u, const double *nu, double *out,
                            hipStream_t st);
void launch_p2_laplace_term_nu_grad(const Mesh &m, const double *go, const double *u,
                                    double *gnu, hipStream_t st);
// assemble.hip
// Fused assembly of a mesh system straight into the solver format: FP64
// SELL values, diagonal, rhs.  (The mixed-precision solvers build their FP32
// matrix from these, scaled.)  P1 and P2 systems (the P2 kernels are in
// p2_system.hip).
struct SellOut {
  double *sv, *diag, *rhs;
};
void launch_assemble(System &s, const admfem_problem &p, hipStream_t st);
void launch_assemble_to(const System &s, const double2 *xy, const admfem_problem &p,
                        const SellOut &o, hipStream_t st);
// CSR values (s.val, general sparse systems) -> solver format
void launch_to_sell(System &s, hipStream_t st);
void launch_kappa_gradient(const System &s, const double *u_int,
                           const double *lam_int, double scale, double *gk,
                           hipStream_t st);
void launch_kappa_gradient_to(const System &s, const double2 *xy, const double *u_int,
                              const double *lam_int, double scale, double *gk, hipStream_t st);
// the same from u and lam in the user's node order: both scattered into
// ul_int [n] as (u, lam) pairs in the internal order, then the gradient
void launch_kappa_gradient_user(const System &s, const double *u, const double *lam,
                                double2 *ul_int, double scale, double *gk, hipStream_t st);
void launch_loss(const System &s, const double *u_int, const int *obs_idx,
                 const double *obs_val, int nobs, double *g_int, double *loss2,
                 hipStream_t st);
// permutation helpers (internal <-> user order)
// dst[idx[i]] = src[i] (idx a permutation): the same data movement as a
// gather with the inverse permutation, with coalesced reads and scattered
// writes, which do not stall the wave (measured faster for user <-> internal
// node order)
void launch_scatter(const int *idx, const double *src, double *dst, int n, hipStream_t st);
 // dst[i] = src[idx[i]]
// dirichlet.hip
int64_t run_impose_dirichlet(const long long *ind, const double *vv,
                             int64_t sN, const long long *bd,
                             const double *bdval, int bdN, const double *rhs,
                             int N, long long *oind, double *ov, double *orhs,
                             hipStream_t st);
void run_impose_dirichlet_grad(const double *gov, const double *gorhs,
                               const long long *ind, const double *vv,
                               int64_t sN, const long long *bd,
                               const double *bdval, int bdN, int N,
                               double *gvv, double *grhs, double *gbd,
                               hipStream_t st);
// p2_system.hip: assembly and adjoint gradient of a P2 system (the
// launchers above dispatch to these); u and lam internal order, either
// separate (u_int, lam_int) or as (u, lam) pairs (ul_int)
void launch_assemble_p2(const System &s, const double2 *xy, const admfem_problem &p,
                        const SellOut &o, hipStream_t st);
void launch_kappa_gradient_p2(const System &s, const double2 *xy, const double *u_int,
                              const double *lam_int, const double2 *ul_int, double scale,
                              double *gk, hipStream_t st);
// noalloc.hip: the gradient with its (u, lam) staging in a fixed device array
size_t kappa_gradient_noalloc_max_nodes();
void kappa_gradient_noalloc(const System &s, const double *u, const double *lam, double scale,
                            double *gk, hipStream_t st);
// cg.hip
void cg_solve(System &s, const double *b_int, double *x_int,
              const admfem_solve_opts &o, admfem_solve_stats &stats,
              hipStream_t st);
void cg_work_free(CGWork *w);
void spmv_repeat(System &s, int precision, int repeat, hipStream_t st);
// sparse_op.hip
void sparse_solve(const long long *ind, const double *vv, int64_t nnz, int n, const double *rhs,
                  double *u, const admfem_solve_opts &o, admfem_solve_stats &stats,
                  hipStream_t st);
void sparse_solve_info(int64_t info[4]);
void sparse_solve_grad(const long long *ind, const double *vv, int64_t nnz, int n,
                       const double *u, const double *grad_u, double *grad_vv,
                       double *grad_rhs, const admfem_solve_opts &o, admfem_solve_stats &stats,
                       hipStream_t st);
// server.cpp
struct Server;
Server *server_create(const admfem_server_opts &o);
void server_free(Server *sv);
int server_add_topology(Server *sv, System *s, double jitter);
int64_t server_submit(Server *sv, const admfem_request &r);
void server_run(Server *sv, double timeout);
void server_result(Server *sv, int64_t id, admfem_result &out);
void server_stats(Server *sv, double *stats);
double server_spmv_repeat(Server *sv, const int *topos, int nsys, int repeat,
                          hipStream_t st, double *seconds);
// loss gradient: g_int = 0; g[perm[idx]] = 2 (u - obs); returns loss on host
// (loss2: 2 device doubles of the caller's, loss and an index-error flag)
double loss_and_grad(const System &s, const double *u_int, const int *obs_idx,
                     const double *obs_val, int nobs, double *g_int, double *loss2,
                     hipStream_t st);

}  // namespace admfem
