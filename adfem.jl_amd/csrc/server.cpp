// Serving engine: continuous batching of forward+adjoint requests on one GPU.
//
// A request runs the reference's pipeline for its own coefficients
// (assemble -> Dirichlet -> u = K^-1 F -> L -> lambda = K^-1 dL/du ->
// dL/dkappa).  The engine admits arrived requests while their device memory
// fits the HBM budget (an arena carved once at creation, so admission and
// release never call hipMalloc/hipFree), advances every admitted system with
// one batched launch per CG step (batched.hip), and moves systems that
// converge to their next phase at the chunk boundaries, where finished
// requests release their memory to the waiting queue.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <random>
#include <thread>
#include <unordered_set>

#include "batched.h"

namespace admfem {

namespace {

constexpr size_t ALIGN = 256;
inline size_t up(size_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }

// Best-fit allocator over one device allocation, coalescing on free.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  void init(size_t bytes) {
    cap = up(bytes);
    HIPCHECK(hipMalloc(&base, cap));
    free_.clear();
    free_[0] = cap;
  }
  ~Arena() {
    if (base) (void)hipFree(base);
  }
  bool alloc(size_t bytes, size_t &off) {
    bytes = up(bytes);
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
    if (best == free_.end()) return false;
    off = best->first;
    size_t rest = best->second - bytes;
    free_.erase(best);
    if (rest) free_[off + bytes] = rest;
    used += bytes;
    peak = std::max(peak, used);
    return true;
  }
  void release(size_t off, size_t bytes) {
    bytes = up(bytes);
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Topo {
  System *sys;
  double jitter;
  int nblk;                  // batched blocks of 256 rows
  size_t bytes;              // device bytes of one request on it
  std::vector<int> interior; // user ids of interior nodes
};

enum { Q_WAIT = 0, Q_FWD = 1, Q_ADJ = 2, Q_DONE = 3 };

struct Req {
  admfem_request r;
  admfem_result res{};
  int state = Q_WAIT;
  int slot = -1;
  size_t off = 0;
  // device buffers (inside the arena)
  double2 *xy;
  double *x, *y, *kappa, *f, *uprev, *b, *sv64, *diag64, *dsq, *xa, *xb, *u64, *lam64,
      *grad, *obsv, *loss2, *part, *stage;
  float *sv32, *ysa, *ysb, *ra, *rb, *pa, *pb, *qa, *qb;
  int *obsi;
  int nobs = 0;
  std::vector<int> h_obs;  // synthetic observation indices, kept until the request is done
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// small kernels of the engine (batched.hip)
void launch_perm_coords(int n, const int *iperm, const double *coords_user, double *x,
                        double *y, double2 *xy, hipStream_t st);

struct Server {
  admfem_server_opts o{};
  hipStream_t st = nullptr;
  Arena arena;
  std::vector<Topo> topos;
  std::vector<Req> reqs;
  int max_slots = 0;
  std::vector<int> free_slots;
  DevBuf<BSys> d_desc;
  DevBuf<BScal> d_scal;
  BScal *h_scal = nullptr;       // pinned mirror
  double *h_loss = nullptr;      // pinned, per slot: loss and bad-index flag
  BSys *h_desc = nullptr;        // pinned staging of the descriptors
  int *h_list = nullptr;         // pinned staging: the batch's slots, then
                                 // their first logical blocks (list_cap each)
  size_t map_cap = 0, list_cap = 0;
  DevBuf<int2> d_map;
  DevBuf<int> d_list;            // slots [list_cap], first blocks [list_cap + 1]
  bool started = false;
  bool use_graph = false;
  std::chrono::steady_clock::time_point t0;
  // stats
  int64_t steps = 0, iters = 0, batch_sum = 0;
  int peak_active = 0;
  double wall = 0;

  ~Server() {
    drop_graphs();
    if (h_scal) (void)hipHostFree(h_scal);
    if (h_loss) (void)hipHostFree(h_loss);
    if (h_desc) (void)hipHostFree(h_desc);
    if (h_list) (void)hipHostFree(h_list);
    if (st) (void)hipStreamDestroy(st);
  }

  // ------------------------------------------------------------ layout
  static size_t layout(const System &s, Req *q, char *base, int nobs) {
    const Mesh &m = *s.mesh;
    const size_t np = s.npad, ng = m.ngauss, ns = std::max<size_t>(1, s.slice_ptr.back());
    const size_t nblk = (np + BATCH_NT - 1) / BATCH_NT;
    const size_t nv = s.degree == 2 ? (size_t)m.nnode : np;  // vertices (coordinates)
    size_t off = 0;
    auto take = [&](size_t bytes) {
      size_t o = off;
      off += up(bytes);
      return base ? (void *)(base + o) : nullptr;
    };
    auto D = [&](size_t n) { return (double *)take(n * sizeof(double)); };
    auto F = [&](size_t n) { return (float *)take(n * sizeof(float)); };
    double *x = D(nv), *y = D(nv), *kappa = D(ng), *f = D(ng), *uprev = D(np), *b = D(np),
           *sv64 = D(ns), *diag64 = D(np), *dsq = D(np), *xa = D(np), *xb = D(np),
           *u64 = D(np), *lam64 = D(np), *grad = D(ng), *obsv = D(nobs + 1), *loss2 = D(2),
           *part = D(4 * nblk), *stage = D(2 * np);
    double2 *xy = (double2 *)take(nv * sizeof(double2));
    float *sv32 = F(ns), *ysa = F(np), *ysb = F(np), *ra = F(np), *rb = F(np), *pa = F(np),
          *pb = F(np), *qa = F(np), *qb = F(np);
    int *obsi = (int *)take((nobs + 1) * sizeof(int));
    if (q) {
      q->x = x; q->y = y; q->xy = xy; q->kappa = kappa; q->f = f; q->uprev = uprev; q->b = b;
      q->sv64 = sv64; q->diag64 = diag64; q->dsq = dsq; q->xa = xa; q->xb = xb;
      q->u64 = u64; q->lam64 = lam64; q->grad = grad; q->obsv = obsv; q->loss2 = loss2;
      q->part = part; q->stage = stage; q->sv32 = sv32; q->ysa = ysa; q->ysb = ysb;
      q->ra = ra; q->rb = rb; q->pa = pa; q->pb = pb; q->qa = qa; q->qb = qb; q->obsi = obsi;
    }
    return off;
  }

  int nobs_of(const Req &q) const {
    if (q.r.source == ADMFEM_SRC_SYNTH)
      return std::max<int>(1, (int)topos[q.r.topo].interior.size() / 100);
    return q.r.nobs;
  }

  // ------------------------------------------------------------ phases
  void write_desc(Req &q, const double *b) {
    const System &s = *topos[q.r.topo].sys;
    // staged in pinned memory: the copy is asynchronous, and a slot's entry
    // is rewritten only after the stream has drained its previous request
    BSys &d = h_desc[q.slot];
    d = BSys{};
    d.sp = s.d_slice_ptr.p;
    d.sc = s.d_sell_col.p;
    d.sd = s.d_sell_d16.p;
    d.row_bd = s.d_row_bd.p;
    d.sv64 = q.sv64;
    d.diag64 = q.diag64;
    d.sv32 = q.sv32;
    d.dsq = q.dsq;
    d.b = b;
    d.x[0] = q.xa;
    d.x[1] = q.xb;
    d.ys[0] = q.ysa;
    d.ys[1] = q.ysb;
    d.r[0] = q.ra;
    d.r[1] = q.rb;
    d.p[0] = q.pa;
    d.p[1] = q.pb;
    d.q[0] = q.qa;
    d.q[1] = q.qb;
    d.part = q.part;
    d.n = s.n;
    d.npad = s.npad;
    d.nblk = topos[q.r.topo].nblk;
    HIPCHECK(hipMemcpyAsync(d_desc.p + q.slot, &d, sizeof(BSys), hipMemcpyHostToDevice, st));
  }

  void inputs(Req &q) {
    const Topo &T = topos[q.r.topo];
    const System &s = *T.sys;
    const Mesh &m = *s.mesh;
    const admfem_request &r = q.r;
    const bool heat = r.kind == 1;
    if (r.source == ADMFEM_SRC_SYNTH) {
      std::mt19937_64 rng(r.seed);
      std::normal_distribution<double> nrm(0.0, 1.0 / 0.1);
      std::uniform_real_distribution<double> ph(0.0, 2.0 * M_PI);
      SynField kf{}, uf{};
      kf.nmodes = uf.nmodes = 32;
      kf.scale = uf.scale = std::sqrt(2.0 / 32.0);
      for (int j = 0; j < 32; j++) {
        kf.wx[j] = nrm(rng);
        kf.wy[j] = nrm(rng);
        kf.phi[j] = ph(rng);
      }
      for (int j = 0; j < 32; j++) {
        uf.wx[j] = nrm(rng);
        uf.wy[j] = nrm(rng);
        uf.phi[j] = ph(rng);
      }
      synth_inputs(s, m.x.p, m.y.p, r.seed, T.jitter, kf, heat ? &uf : nullptr, q.x, q.y,
                   q.xy, q.kappa, q.f, q.uprev, st);
      // observations: 1% of the interior nodes, obs_val = 0; the indices'
      // host copy lives with the request (no wait for the copy here)
      std::vector<int> &idx = q.h_obs;
      idx.clear();
      std::unordered_set<int> seen;
      std::uniform_int_distribution<size_t> pick(0, T.interior.size() - 1);
      while ((int)idx.size() < q.nobs) {
        int u = T.interior[pick(rng)];
        if (seen.insert(u).second) idx.push_back(u);
      }
      HIPCHECK(hipMemcpyAsync(q.obsi, idx.data(), sizeof(int) * q.nobs, hipMemcpyHostToDevice, st));
      HIPCHECK(hipMemsetAsync(q.obsv, 0, sizeof(double) * q.nobs, st));
      return;
    }
    const hipMemcpyKind kind =
        r.source == ADMFEM_SRC_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    // (the inputs' presence was checked by server_submit)
    // user-order inputs are staged, then permuted to the internal order
    const int nn = m.nnode;  // vertices
    if (r.coords) {
      HIPCHECK(hipMemcpyAsync(q.stage, r.coords, sizeof(double) * 2 * nn, kind, st));
      launch_perm_coords(nn, m.d_iperm.p, q.stage, q.x, q.y, q.xy, st);
    } else {
      HIPCHECK(hipMemcpyAsync(q.x, m.x.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
      HIPCHECK(hipMemcpyAsync(q.y, m.y.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
      HIPCHECK(hipMemcpyAsync(q.xy, m.xy.p, sizeof(double2) * nn, hipMemcpyDeviceToDevice, st));
    }
    HIPCHECK(hipMemcpyAsync(q.kappa, r.kappa, sizeof(double) * m.ngauss, kind, st));
    HIPCHECK(hipMemcpyAsync(q.f, r.f, sizeof(double) * m.ngauss, kind, st));
    if (heat) {
      HIPCHECK(hipMemcpyAsync(q.stage, r.u_prev, sizeof(double) * s.n, kind, st));
      launch_scatter(s.d_uperm(), q.stage, q.uprev, s.n, st);
    }
    if (r.nobs) {
      HIPCHECK(hipMemcpyAsync(q.obsi, r.obs_idx, sizeof(int) * r.nobs, kind, st));
      HIPCHECK(hipMemcpyAsync(q.obsv, r.obs_val, sizeof(double) * r.nobs, kind, st));
    }
    if (r.source == ADMFEM_SRC_HOST) HIPCHECK(hipStreamSynchronize(st));
  }

  admfem_problem problem(const Req &q) const {
    admfem_problem p{};
    const bool heat = q.r.kind == 1;
    p.kappa = q.kappa;
    p.f = q.f;
    p.u_prev = heat ? q.uprev : nullptr;
    p.lap_coef = heat ? q.r.dt : 1.0;
    p.mass_coef = heat ? 1.0 : 0.0;
    p.src_coef = heat ? q.r.dt : 1.0;
    return p;
  }

  void assemble(Req &q) {
    const System &s = *topos[q.r.topo].sys;
    inputs(q);
    admfem_problem p = problem(q);
    // FP64 matrix and rhs; the FP32 scaled matrix of the inner iteration
    launch_assemble_to(s, q.xy, p, SellOut{q.sv64, q.diag64, q.b}, st);
    batch_scale(s, q.sv64, q.diag64, q.dsq, q.sv32, st);
  }

  void start_cg(Req &q) {
    const double d = o.replace_delta > 0 ? o.replace_delta : REPLACE_DELTA;
    batch_init(d_desc.p, d_scal.p, q.slot, topos[q.r.topo].nblk, o.rtol * o.rtol, d * d,
               o.max_iter, st);
  }

  bool admit(Req &q, double t) {
    const Topo &T = topos[q.r.topo];
    if (free_slots.empty()) return false;
    q.nobs = nobs_of(q);
    size_t bytes = layout(*T.sys, nullptr, nullptr, q.nobs);
    if (bytes > arena.cap) {
      q.state = Q_DONE;
      q.res.done = 1;
      q.res.status = ADMFEM_ERR_BUDGET;
      q.res.t_done = t;
      return true;
    }
    size_t off;
    if (!arena.alloc(bytes, off)) return false;
    q.off = off;
    q.res.device_bytes = (double)bytes;
    layout(*T.sys, &q, arena.base + off, q.nobs);
    q.slot = free_slots.back();
    free_slots.pop_back();
    q.res.t_admit = t;
    try {
      assemble(q);
      write_desc(q, q.b);
      start_cg(q);
    } catch (const Error &e) {
      // the request fails alone: its memory and slot go back to the queue,
      // the requests behind it are admitted as usual
      (void)hipStreamSynchronize(st);
      set_error(e.msg);
      fail(q, e.code == ADMFEM_ERR_HIP ? ADMFEM_ERR_HIP : ADMFEM_ERR_ARG);
      return true;
    }
    q.state = Q_FWD;
    return true;
  }

  void release(Req &q) {
    arena.release(q.off, layout(*topos[q.r.topo].sys, nullptr, nullptr, q.nobs));
    free_slots.push_back(q.slot);
    q.slot = -1;
  }

  // forward solve finished: u, loss, dL/du; start the adjoint solve
  void to_adjoint(Req &q, const BScal &sc) {
    const System &s = *topos[q.r.topo].sys;
    q.res.iters_fwd = sc.it;
    batch_final(d_desc.p, d_scal.p, q.slot, s.npad, q.u64, st);
    launch_loss(s, q.u64, q.obsi, q.obsv, q.nobs, q.b, q.loss2, st);  // g overwrites b
    start_cg(q);
    q.state = Q_ADJ;
  }

  // adjoint solve finished: the gradient and the outputs are enqueued, the
  // loss goes to the slot's pinned staging; complete() finishes the request
  // once the stream has drained (one synchronisation for every request that
  // finished at the same chunk boundary)
  void finish(Req &q, const BScal &sc) {
    const System &s = *topos[q.r.topo].sys;
    const Mesh &m = *s.mesh;
    q.res.iters_adj = sc.it;
    batch_final(d_desc.p, d_scal.p, q.slot, s.npad, q.lam64, st);
    admfem_problem p = problem(q);
    launch_kappa_gradient_to(s, q.xy, q.u64, q.lam64, p.lap_coef, q.grad, st);
    if (q.r.u_out) {
      launch_scatter(s.d_uiperm(), q.u64, q.stage, s.n, st);  // user order
      HIPCHECK(hipMemcpyAsync(q.r.u_out, q.stage, sizeof(double) * s.n, hipMemcpyDeviceToHost,
                              st));
    }
    if (q.r.grad_out)
      HIPCHECK(hipMemcpyAsync(q.r.grad_out, q.grad, sizeof(double) * m.ngauss,
                              hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h_loss + 2 * q.slot, q.loss2, 2 * sizeof(double),
                            hipMemcpyDeviceToHost, st));
  }

  // after the stream has drained: a finished request's loss and status, its
  // memory and slot back to the queue
  void complete(Req &q) {
    std::vector<int>().swap(q.h_obs);
    const double *l2 = h_loss + 2 * q.slot;
    q.res.loss = l2[0];
    if (l2[1] != 0.0 && q.res.status == ADMFEM_OK) q.res.status = ADMFEM_ERR_ARG;
    q.res.t_done = now_s(t0);
    q.res.done = 1;
    q.state = Q_DONE;
    release(q);
  }

  void fail(Req &q, int code) {
    HIPCHECK(hipStreamSynchronize(st));
    q.res.status = code;
    q.res.done = 1;
    q.res.t_done = now_s(t0);
    q.state = Q_DONE;
    release(q);
  }

  static int code_of(int status) {
    switch (status) {
      case 1: return ADMFEM_OK;
      case 3: return ADMFEM_ERR_NAN;
      case 4: return ADMFEM_ERR_BREAKDOWN;
      default: return ADMFEM_ERR_NOT_CONVERGED;
    }
  }

  // ------------------------------------------------------------ batched run
  // Device map of gmap grid blocks and host/device lists for nlist systems;
  // growing them drops the captured graphs (they point at the old buffers).
  void ensure_map(size_t gmap, size_t nlist) {
    if (gmap > map_cap) {
      map_cap = std::max(gmap, 2 * map_cap);
      d_map.alloc(map_cap);
      drop_graphs();
    }
    if (nlist > list_cap) {
      if (h_list) HIPCHECK(hipHostFree(h_list));
      list_cap = std::max(nlist, 2 * list_cap);
      HIPCHECK(hipHostMalloc((void **)&h_list, sizeof(int) * (2 * list_cap + 1),
                             hipHostMallocDefault));
      d_list.alloc(2 * list_cap + 1);
      drop_graphs();
    }
  }

  // Uploads the slots of `slots` (padded to glist with `pad`) and their
  // first logical blocks, and expands the block map on the device.
  void upload_batch(const std::vector<std::pair<int, int>> &slots, size_t glist, int pad,
                    size_t nmap, size_t gmap) {
    int *pref = h_list + list_cap;
    int acc = 0;
    for (size_t a = 0; a < slots.size(); a++) {
      h_list[a] = slots[a].first;
      pref[a] = acc;
      acc += slots[a].second;
    }
    pref[slots.size()] = acc;
    for (size_t a = slots.size(); a < glist; a++) h_list[a] = pad;
    HIPCHECK(hipMemcpyAsync(d_list.p, h_list, sizeof(int) * glist, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_list.p + list_cap, pref, sizeof(int) * (slots.size() + 1),
                            hipMemcpyHostToDevice, st));
    batch_map((int)slots.size(), d_list.p, d_list.p + list_cap, (int)nmap, (int)gmap, d_map.p,
              st);
  }

  // One CG chunk for a (map size, list size) bucket, captured once as a HIP
  // graph of kernel nodes only and replayed for every batch of that size.
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  void drop_graphs() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  static int bucket(size_t v) {
    size_t b = 64;
    while (b < v) b <<= 1;
    return (int)b;
  }
  void launch_chunk(int nmap, int nlist, int chunk) {
    auto key = std::make_pair(nmap, nlist);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      hipGraph_t g;
      HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      batch_iterations(d_map.p, nmap, d_list.p, nlist, d_desc.p, d_scal.p, chunk, st);
      HIPCHECK(hipStreamEndCapture(st, &g));
      hipGraphExec_t ex;
      HIPCHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      HIPCHECK(hipGraphDestroy(g));
      it = graphs.emplace(key, ex).first;
    }
    HIPCHECK(hipGraphLaunch(it->second, st));
  }

  void run_batched(double timeout) {
    const int chunk = o.chunk > 0 ? (o.chunk + REPLACE_EVERY - 1) / REPLACE_EVERY * REPLACE_EVERY
                                  : 2 * REPLACE_EVERY;
    std::vector<int> order(reqs.size());
    for (size_t i = 0; i < reqs.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return reqs[a].r.arrival < reqs[b].r.arrival;
    });
    size_t next = 0;  // first request (in arrival order) not yet admitted
    while (next < order.size() && reqs[order[next]].state != Q_WAIT) next++;
    std::vector<int> active;
    for (size_t i = 0; i < reqs.size(); i++)
      if (reqs[i].state == Q_FWD || reqs[i].state == Q_ADJ) active.push_back((int)i);
    // requests whose last work (outputs, loss) is enqueued, and an event
    // behind that work: they complete once it has passed -- waited for
    // while the next chunk runs, so the host bookkeeping overlaps it
    std::vector<int> done, failed;
    struct Ev {
      hipEvent_t e = nullptr;
      ~Ev() {
        if (e) (void)hipEventDestroy(e);
      }
    } ev;
    HIPCHECK(hipEventCreateWithFlags(&ev.e, hipEventDisableTiming));
    hipEvent_t ev_done = ev.e;
    auto settle = [&]() {
      for (int i : done) complete(reqs[i]);
      for (int i : failed) {
        Req &q = reqs[i];
        std::vector<int>().swap(q.h_obs);
        q.res.done = 1;
        q.res.t_done = now_s(t0);
        q.state = Q_DONE;
        release(q);
      }
      done.clear();
      failed.clear();
    };
    auto flush = [&]() {
      if (done.empty() && failed.empty()) return;
      HIPCHECK(hipEventSynchronize(ev_done));
      settle();
    };
    for (;;) {
      double t = now_s(t0);
      if (t > timeout) break;
      // admit arrived requests, FIFO, while memory and slots allow
      while (next < order.size() && (o.max_active <= 0 || (int)active.size() < o.max_active)) {
        Req &q = reqs[order[next]];
        if (q.r.arrival > t) break;
        if (!admit(q, t)) break;
        if (q.state != Q_DONE) active.push_back(order[next]);
        next++;
      }
      peak_active = std::max(peak_active, (int)active.size());
      if (active.empty()) {
        if (!done.empty() || !failed.empty()) {  // their slots may admit the next ones
          flush();
          continue;
        }
        if (next >= order.size()) break;
        double wait = reqs[order[next]].r.arrival - now_s(t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
        continue;
      }
      // block map over the active systems; with graphs, both grids are padded
      // to power-of-two buckets with entries that do no work (map: system -1,
      // list: the never-running dummy slot)
      size_t nmap = 0;
      for (int i : active) nmap += topos[reqs[i].r.topo].nblk;
      const size_t gmap = use_graph ? (size_t)bucket(nmap) : nmap;
      const size_t glist = use_graph ? (size_t)bucket(active.size()) : active.size();
      ensure_map(gmap, glist);
      std::vector<std::pair<int, int>> slots(active.size());
      for (size_t a = 0; a < active.size(); a++) {
        const Req &q = reqs[active[a]];
        slots[a] = std::make_pair(q.slot, topos[q.r.topo].nblk);
      }
      upload_batch(slots, glist, max_slots, nmap, gmap);  // padding: the dummy slot
      if (use_graph)
        launch_chunk((int)gmap, (int)glist, chunk);
      else
        batch_iterations(d_map.p, (int)nmap, d_list.p, (int)active.size(), d_desc.p, d_scal.p,
                         chunk, st);
      steps += 1;
      iters += chunk;
      batch_sum += (int64_t)active.size();
      HIPCHECK(hipMemcpyAsync(h_scal, d_scal.p, sizeof(BScal) * max_slots,
                              hipMemcpyDeviceToHost, st));
      flush();  // the previous chunk's finished requests, while this one runs
      HIPCHECK(hipStreamSynchronize(st));
      // phase changes; requests that finished or failed complete once their
      // enqueued work has passed
      std::vector<int> still;
      for (int i : active) {
        Req &q = reqs[i];
        const BScal sc = h_scal[q.slot];
        if (sc.status == 0) {
          still.push_back(i);
          continue;
        }
        if (sc.status != 1) {
          if (q.state == Q_FWD) q.res.iters_fwd = sc.it;
          else q.res.iters_adj = sc.it;
          q.res.status = code_of(sc.status);
          failed.push_back(i);
          continue;
        }
        if (q.state == Q_FWD) {
          q.res.relres_fwd = sc.bb > 0 ? std::sqrt(sc.rr / sc.bb) : 0.0;
          to_adjoint(q, sc);
          still.push_back(i);
        } else {
          q.res.relres_adj = sc.bb > 0 ? std::sqrt(sc.rr / sc.bb) : 0.0;
          finish(q, sc);
          done.push_back(i);
        }
      }
      if (!done.empty() || !failed.empty()) HIPCHECK(hipEventRecord(ev_done, st));
      active.swap(still);
    }
    flush();
    wall = now_s(t0);
  }

  // ------------------------------------------------------------ sequential
  void run_seq(double timeout) {
    std::vector<int> order(reqs.size());
    for (size_t i = 0; i < reqs.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return reqs[a].r.arrival < reqs[b].r.arrival;
    });
    for (int i : order) {
      Req &q = reqs[i];
      if (q.state != Q_WAIT) continue;
      double t = now_s(t0);
      if (t > timeout) break;
      if (q.r.arrival > t)
        std::this_thread::sleep_for(std::chrono::duration<double>(q.r.arrival - t));
      t = now_s(t0);
      System &s = *topos[q.r.topo].sys;
      q.nobs = nobs_of(q);
      size_t bytes = layout(s, nullptr, nullptr, q.nobs);
      size_t off;
      if (bytes > arena.cap || !arena.alloc(bytes, off)) {
        q.res.status = ADMFEM_ERR_BUDGET;
        q.res.done = 1;
        q.state = Q_DONE;
        continue;
      }
      q.off = off;
      q.res.device_bytes = (double)bytes;
      layout(s, &q, arena.base + off, q.nobs);
      q.slot = 0;
      q.res.t_admit = t;
      peak_active = std::max(peak_active, 1);
      try {
        serve_seq(q, s);
      } catch (const Error &e) {
        (void)hipStreamSynchronize(st);
        set_error(e.msg);
        q.res.status = e.code == ADMFEM_ERR_HIP ? ADMFEM_ERR_HIP : ADMFEM_ERR_ARG;
      }
      q.res.t_done = now_s(t0);
      q.res.done = 1;
      q.state = Q_DONE;
      arena.release(q.off, bytes);
    }
    wall = now_s(t0);
  }

  // One request on the FP64 single-system path, on the topology's own solver
  // buffers (the request's arena range holds its inputs and outputs).
  void serve_seq(Req &q, System &s) {
    inputs(q);
    admfem_problem p = problem(q);
    launch_assemble_to(s, q.xy, p, SellOut{s.d_sell_val.p, s.d_diag.p, s.rhs.p},
                       st);
    admfem_solve_opts so{ADMFEM_PREC_FP64, o.max_iter, o.rtol, 0, 0.0};
    admfem_solve_stats s0{}, s1{};
    cg_solve(s, s.rhs.p, q.u64, so, s0, st);
    launch_loss(s, q.u64, q.obsi, q.obsv, q.nobs, q.b, q.loss2, st);
    cg_solve(s, q.b, q.lam64, so, s1, st);
    q.res.iters_fwd = s0.iterations;
    q.res.iters_adj = s1.iterations;
    q.res.relres_fwd = s0.rel_residual;
    q.res.relres_adj = s1.rel_residual;
    iters += s0.iterations + s1.iterations;
    steps += 1;
    batch_sum += 1;
    q.res.status = s0.status != ADMFEM_OK ? s0.status : s1.status;
    if (q.res.status != ADMFEM_OK) {
      HIPCHECK(hipStreamSynchronize(st));
      return;
    }
    const Mesh &m = *s.mesh;
    launch_kappa_gradient_to(s, q.xy, q.u64, q.lam64, p.lap_coef, q.grad, st);
    if (q.r.u_out) {
      launch_scatter(s.d_uiperm(), q.u64, q.stage, s.n, st);
      HIPCHECK(hipMemcpyAsync(q.r.u_out, q.stage, sizeof(double) * s.n, hipMemcpyDeviceToHost,
                              st));
    }
    if (q.r.grad_out)
      HIPCHECK(hipMemcpyAsync(q.r.grad_out, q.grad, sizeof(double) * m.ngauss,
                              hipMemcpyDeviceToHost, st));
    double l2[2];
    HIPCHECK(hipMemcpyAsync(l2, q.loss2, sizeof(l2), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    q.res.loss = l2[0];
    if (l2[1] != 0.0) q.res.status = ADMFEM_ERR_ARG;  // bad or repeated obs index
  }
};

// Instrumentation: the batched CG step kernel over a batch of synthetic
// Poisson systems.  The scalars it reads (alpha = beta = 0, it, status) stay
// as batch_init left them, so every launch reads the same buffers, does its
// full work and writes the same values to the other buffers.
double server_spmv_repeat(Server *sv, const int *topos, int nsys, int repeat,
                          hipStream_t st, double *seconds) {
  REQUIRE(nsys <= (int)sv->free_slots.size(), ADMFEM_ERR_ARG, "more systems than free slots");
  for (int i = 0; i < nsys; i++)
    REQUIRE(topos[i] >= 0 && topos[i] < (int)sv->topos.size(), ADMFEM_ERR_ARG,
            "bad topology id");
  std::vector<Req> qs(nsys);
  int admitted = 0;
  double bytes = 0.0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto cleanup = [&] {
    (void)hipStreamSynchronize(st);
    (void)hipStreamSynchronize(sv->st);
    for (int i = 0; i < admitted; i++) sv->release(qs[i]);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  try {
    size_t nmap = 0;
    for (int i = 0; i < nsys; i++) {
      Req &q = qs[i];
      q.r = admfem_request{};
      q.r.topo = topos[i];
      q.r.source = ADMFEM_SRC_SYNTH;
      q.r.seed = (unsigned long long)i;
      const bool ok = sv->admit(q, 0.0);
      REQUIRE(ok && q.state == Q_FWD, ADMFEM_ERR_BUDGET, "batch does not fit the budget");
      admitted++;
      nmap += sv->topos[q.r.topo].nblk;
      const System &s = *sv->topos[q.r.topo].sys;
      // one CG iteration of the canonical CSR form: the SpMV's
      // nnz (s + 4) + 4 (n + 1) + 2 n s plus 6 n s of vector updates (s = 4)
      bytes += (double)s.nnz * 8.0 + 4.0 * (s.n + 1) + 8.0 * s.n * 4.0;
    }
    sv->ensure_map(nmap, (size_t)nsys);
    std::vector<std::pair<int, int>> slots(nsys);
    for (int i = 0; i < nsys; i++)
      slots[i] = std::make_pair(qs[i].slot, sv->topos[qs[i].r.topo].nblk);
    sv->upload_batch(slots, (size_t)nsys, sv->max_slots, nmap, nmap);
    HIPCHECK(hipStreamSynchronize(sv->st));
    HIPCHECK(hipEventCreate(&ev[0]));
    HIPCHECK(hipEventCreate(&ev[1]));
    HIPCHECK(hipEventRecord(ev[0], st));
    batch_spmv_repeat(sv->d_map.p, (int)nmap, sv->d_desc.p, sv->d_scal.p, repeat, st);
    HIPCHECK(hipEventRecord(ev[1], st));
    HIPCHECK(hipEventSynchronize(ev[1]));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    *seconds = repeat > 0 ? 1e-3 * ms / repeat : 0.0;
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  return bytes;
}

// ---------------------------------------------------------------- C-ABI glue
Server *server_create(const admfem_server_opts &o) {
  REQUIRE(o.hbm_budget_bytes > 0, ADMFEM_ERR_ARG, "budget must be positive");
  REQUIRE(o.rtol > 0 && o.max_iter > 0, ADMFEM_ERR_ARG, "bad rtol / max_iter");
  REQUIRE(o.mode == ADMFEM_MODE_BATCHED || o.mode == ADMFEM_MODE_SEQ_FP64, ADMFEM_ERR_ARG,
          "bad mode");
  Server *sv = new Server();
  try {
    sv->o = o;
    HIPCHECK(hipStreamCreateWithFlags(&sv->st, hipStreamNonBlocking));
    sv->arena.init((size_t)o.hbm_budget_bytes);
    sv->max_slots = o.max_active > 0 ? o.max_active : 1024;
    for (int i = sv->max_slots - 1; i >= 0; i--) sv->free_slots.push_back(i);
    sv->use_graph = o.use_graph != 0;
    // one extra slot: the dummy the padded graph launches point at; its
    // status is never "running", so every kernel skips it
    sv->d_desc.alloc(sv->max_slots + 1);
    sv->d_scal.alloc(sv->max_slots + 1);
    HIPCHECK(hipMemset(sv->d_desc.p, 0, sv->d_desc.bytes()));
    HIPCHECK(hipMemset(sv->d_scal.p, 0, sv->d_scal.bytes()));
    BScal dummy{};
    dummy.status = 99;
    HIPCHECK(hipMemcpy(sv->d_scal.p + sv->max_slots, &dummy, sizeof(BScal),
                       hipMemcpyHostToDevice));
    HIPCHECK(hipHostMalloc((void **)&sv->h_scal, sizeof(BScal) * sv->max_slots,
                           hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&sv->h_loss, 2 * sizeof(double) * sv->max_slots,
                           hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void **)&sv->h_desc, sizeof(BSys) * sv->max_slots,
                           hipHostMallocDefault));
  } catch (...) {
    delete sv;
    throw;
  }
  return sv;
}

void server_free(Server *sv) { delete sv; }

int server_add_topology(Server *sv, System *s, double jitter) {
  Topo T;
  T.sys = s;
  T.jitter = jitter;
  T.nblk = (s->npad + BATCH_NT - 1) / BATCH_NT;
  T.bytes = Server::layout(*s, nullptr, nullptr, std::max(1, s->n / 100));
  for (int u = 0; u < s->n; u++)
    if (s->row_bd[s->uperm()[u]] < 0) T.interior.push_back(u);
  REQUIRE(!T.interior.empty(), ADMFEM_ERR_ARG, "topology has no interior node");
  sv->topos.push_back(std::move(T));
  return (int)sv->topos.size() - 1;
}

int64_t server_submit(Server *sv, const admfem_request &r) {
  REQUIRE(r.topo >= 0 && r.topo < (int)sv->topos.size(), ADMFEM_ERR_ARG, "bad topology id");
  REQUIRE(r.kind == 0 || r.kind == 1, ADMFEM_ERR_ARG, "bad request kind");
  REQUIRE(r.source >= 0 && r.source <= 2, ADMFEM_ERR_ARG, "bad request source");
  REQUIRE(r.kind == 0 || r.dt > 0, ADMFEM_ERR_ARG, "heat step needs dt > 0");
  REQUIRE(r.nobs >= 0, ADMFEM_ERR_ARG, "bad nobs");
  if (r.source != ADMFEM_SRC_SYNTH) {
    // every input the request reads is checked here, before it can hold a
    // slot or arena memory
    REQUIRE(r.kappa && r.f, ADMFEM_ERR_ARG, "request needs kappa and f");
    REQUIRE(r.kind == 0 || r.u_prev, ADMFEM_ERR_ARG, "heat step needs u_prev");
    REQUIRE(r.nobs == 0 || (r.obs_idx && r.obs_val), ADMFEM_ERR_ARG,
            "observations need obs_idx and obs_val");
    if (r.source == ADMFEM_SRC_HOST && r.nobs > 0) {
      // host indices: range and distinctness checked now (device indices
      // are checked by the loss kernel, which fails the request)
      const int n = sv->topos[r.topo].sys->n;
      std::vector<int> idx(r.obs_idx, r.obs_idx + r.nobs);
      for (int v : idx)
        REQUIRE(v >= 0 && v < n, ADMFEM_ERR_ARG, "observation index out of range [0, n)");
      std::sort(idx.begin(), idx.end());
      REQUIRE(std::adjacent_find(idx.begin(), idx.end()) == idx.end(), ADMFEM_ERR_ARG,
              "observation indices must be distinct");
    }
  }
  Req q;
  q.r = r;
  q.res.t_arrival = r.arrival;
  sv->reqs.push_back(q);
  return (int64_t)sv->reqs.size() - 1;
}

void server_run(Server *sv, double timeout) {
  if (!sv->started) {
    sv->started = true;
    sv->t0 = std::chrono::steady_clock::now();
  }
  if (sv->o.mode == ADMFEM_MODE_BATCHED) sv->run_batched(timeout);
  else sv->run_seq(timeout);
}

void server_result(Server *sv, int64_t id, admfem_result &out) {
  REQUIRE(id >= 0 && id < (int64_t)sv->reqs.size(), ADMFEM_ERR_ARG, "bad request id");
  out = sv->reqs[id].res;
}

void server_stats(Server *sv, double *s) {
  int done = 0, failed = 0;
  for (auto &q : sv->reqs)
    if (q.res.done) {
      done++;
      if (q.res.status != ADMFEM_OK) failed++;
    }
  s[0] = done;
  s[1] = sv->wall;
  s[2] = (double)sv->arena.peak;
  s[3] = sv->peak_active;
  s[4] = (double)sv->steps;
  s[5] = (double)sv->iters;
  s[6] = failed;
  s[7] = sv->steps ? (double)sv->batch_sum / sv->steps : 0.0;
}

}  // namespace admfem
