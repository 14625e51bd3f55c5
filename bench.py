"""Benchmark: forward+adjoint solves/s at 1e-8 relative error on one MI355X,
with p50/p99 latency, against the plain FP64 GPU path.

Workload (the serving form): a stream of 4096 forward+adjoint requests.  Each
request is P1 Poisson (K u = F) or an implicit heat step ((M + dt K) u =
M u_prev + dt F, dt ~ logU[1e-4, 1e-1]) with probability 1/2, on a unit-square
mesh of side s = round(exp(U[ln 32, ln 512])) quads (2 s^2 triangles), interior
nodes jittered by U[-0.2h, 0.2h]^2; kappa = exp(GRF: 32 random Fourier modes,
sigma 1, length 0.1) and f ~ U[-1, 1] at the Gauss points, homogeneous
Dirichlet on the boundary, L = sum (u - u_obs)^2 over a seeded 1% of the
interior nodes; each request returns u, L and dL/dkappa.  Requests are served
by the continuous-batching engine (FP32-storage Jacobi-PCG with FP64 residual
replacement, every system's true FP64 relative residual driven below rtol)
under a 64 GB HBM admission budget.  Inputs are generated on the device from
each request's seed; mesh topologies are resident before timing starts.

One step = one pass of the whole stream through a fresh engine with every
request queued at t = 0 (throughput at capacity).  After the timed steps:
  * load: the same stream with Poisson arrivals at 80% of the measured
    capacity -> p50 / p99 latency (arrival to result);
  * fp64_plain: the plain FP64 GPU path (one request at a time, FP64
    Jacobi-PCG) on the first requests of the stream -> its solves/s and
    latency; the batched engine on the same subset beside it;
  * accuracy: requests of the stream with host-side inputs, served by both
    paths and compared with the CPU oracle's direct solve (u and dL/dkappa);
  * roofline: the batched SpMV kernel over the stream's first batch, timed
    with HIP events on the stream it runs on;
  * precision_form: one 2048x2048-quad system (8.4M triangles), mixed vs plain
    FP64 forward+adjoint, errors against an FP64 CG solve to 1e-13, the
    single-system SpMV kernels and the assembly timed with HIP events;
  * operators: each reference operator's kernel through its C-ABI entry
    point (P1 at 2048x2048, P2 at 1024x1024 quads: the COO Laplace op and
    its gradient, the load vector, the maps to Gauss points, FemGrad, the
    matrix-free Laplace term; Dirichlet on a COO and the general K\\F op at
    512x512), HIP events, against algorithmic bytes;
  * cpu_baseline: the CPU oracle (C restatement of the reference loops + SciPy
    SuperLU) on the first requests of the stream, one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "adfem.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEEDS = dict(stream=7, arrivals=8, mesh=20240601, kappa=11, kappa_obs=12, f=13, obs=14,
             host=5000)
BUDGET = 64e9
MAX_ACTIVE = 1024  # concurrently admitted requests (engine slots); --max-active


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# --------------------------------------------------------------- the stream
def make_stream(n):
    rng = np.random.default_rng(SEEDS["stream"])
    sides = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), n))).astype(int)
    heat = rng.random(n) < 0.5
    dts = np.exp(rng.uniform(np.log(1e-4), np.log(1e-1), n))
    return sides, heat, dts


def make_topologies(sides):
    import admfem
    from admfem import meshgen
    systems = {}
    for s in sorted(set(sides.tolist())):
        c, e = meshgen.jittered_square(s, seed=0, jitter=0.0, permute=False)
        mesh = admfem.Mesh(c, e, zero_based=True)
        systems[s] = admfem.DirichletSystem(mesh, admfem.bcnode(mesh))
    return systems


def serve(systems, sides, heat, dts, idx, arrivals, mode, rtol, want=False):
    """Serves requests ``idx`` of the stream through a fresh engine; returns
    (wall seconds of run(), results, stats)."""
    import torch
    from admfem.server import Server
    sv = Server(BUDGET, mode=mode, rtol=rtol, max_active=MAX_ACTIVE)
    tid = {s: sv.add_topology(S, 0.2 / s) for s, S in systems.items()}
    ids = [sv.submit(tid[int(sides[i])], kind="heat" if heat[i] else "poisson",
                     dt=float(dts[i]), seed=1000 + int(i), arrival=float(a), want_u=want,
                     want_grad=want) for i, a in zip(idx, arrivals)]
    torch.cuda.synchronize()
    t = time.perf_counter()
    sv.run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    rs = [sv.result(r) for r in ids]
    st = sv.stats()
    sv.close()
    return wall, rs, st


def cg_iteration_bytes(S):
    """Algorithmic bytes of one CG iteration of a system in FP32 storage:
    the canonical CSR SpMV, nnz (4 + 4) + 4 (n + 1) + 2 n 4, plus the
    iteration's vector updates, 6 n 4."""
    return S.nnz * 8 + 4 * (S.n + 1) + 8 * S.n * 4


def summary(n, wall, rs, st, req_sides=None, systems=None):
    lat = np.array([r["latency"] for r in rs])
    svc = np.array([r["t_done"] - r["t_admit"] for r in rs])
    extra = {}
    if req_sides is not None:
        # latency per side bucket, and the algorithmic bytes of every CG
        # iteration the pass ran (FP32 step kernels; replacements not counted)
        extra["p99_ms_by_side"] = {}
        for lo, hi in SIDE_BUCKETS:
            sel = (req_sides >= lo) & (req_sides < hi)
            if sel.any():
                key = f"[{lo},{min(hi, 512)}]" if hi > 512 else f"[{lo},{hi})"
                extra["p99_ms_by_side"][key] = round(1e3 * float(np.percentile(lat[sel], 99)), 2)
        if systems is not None:
            extra["cg_step_bytes"] = float(sum(
                (r["iters_fwd"] + r["iters_adj"]) * cg_iteration_bytes(systems[int(sd)])
                for r, sd in zip(rs, req_sides)))
            extra["cg_step_GBs_of_wall"] = round(extra["cg_step_bytes"] / wall / 1e9, 1)
    return dict(requests=n, failed=int(sum(r["status"] != 0 for r in rs)),
                wall_s=round(wall, 3), solves_per_s=round(n / max(wall, 1e-9), 2),
                p50_ms=round(1e3 * float(np.percentile(lat, 50)), 2),
                p99_ms=round(1e3 * float(np.percentile(lat, 99)), 2),
                service_p50_ms=round(1e3 * float(np.percentile(svc, 50)), 2),
                service_p99_ms=round(1e3 * float(np.percentile(svc, 99)), 2),
                max_relres=float(max(max(r["relres_fwd"], r["relres_adj"]) for r in rs)),
                mean_iters=round(float(np.mean([r["iters_fwd"] + r["iters_adj"] for r in rs])), 1),
                peak_active=int(st["peak_active"]), peak_gb=round(st["peak_bytes"] / 1e9, 3),
                mean_batch=round(st["mean_batch"], 1), **extra)


# ------------------------------------------------- host-side request inputs
def _meshgen():
    """The host mesh generator module, loaded from its file: the CPU
    baseline's worker processes use it without importing the package (and
    with it torch)."""
    import importlib.util
    mod = sys.modules.get("_bench_meshgen")
    if mod is None:
        spec = importlib.util.spec_from_file_location(
            "_bench_meshgen", os.path.join(ROOT, "adfem.jl_amd", "admfem", "meshgen.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        sys.modules["_bench_meshgen"] = mod
    return mod


def host_request(side, heat, dt, seed):
    """Inputs of one request of the stream generated on the host (same
    distributions as the engine's device generator, own random numbers)."""
    meshgen = _meshgen()
    rng = np.random.default_rng(seed)
    c, e = meshgen.jittered_square(side, seed=seed, jitter=0.2, permute=False)
    g = meshgen.gauss_points(c, e)
    kappa = np.exp(meshgen.grf(g, seed + 1))
    f = rng.uniform(-1.0, 1.0, g.shape[0])
    up = meshgen.grf(c, seed + 2) if heat else None
    n = c.shape[0]
    j, i = np.divmod(np.arange(n), side + 1)
    inner = np.flatnonzero((i > 0) & (i < side) & (j > 0) & (j < side))
    obs = np.sort(rng.choice(inner, size=max(1, inner.size // 100), replace=False))
    return dict(side=side, coords=c, elems=e, kappa=kappa, f=f, u_prev=up,
                dt=dt if heat else None, obs=obs.astype(np.int32), obs_val=np.zeros(obs.size))


def oracle_solve(P):
    from oracle import oracle as O
    om = O.OracleMesh(P["coords"], P["elems"])
    bd = om.bcnode()
    obs = P["obs"]

    def dl(u):
        g = np.zeros_like(u)
        g[obs] = 2.0 * (u[obs] - P["obs_val"])
        return g
    kw = {} if P["dt"] is None else dict(dt=P["dt"], u_prev=P["u_prev"])
    return O.forward_adjoint(om, P["kappa"], P["f"], bd, np.zeros(bd.size), dl, **kw)


SIDE_BUCKETS = ((32, 128), (128, 256), (256, 513))


def accuracy(systems, sides, heat, dts, rtol, per_bucket):
    """Host-input requests of the stream through both GPU paths against the
    CPU oracle's direct solve: the first ``per_bucket`` requests of each side
    bucket plus the stream's largest request.  The batched path and the plain
    FP64 path run at the benchmark's rtol (the throughput comparison); the
    plain FP64 path runs again at rtol 1e-12 (its parity bar, 1e-10 against
    the reference's kernels plus a direct solve)."""
    from admfem.server import Server
    pick = []
    for lo, hi in SIDE_BUCKETS:
        pick += [i for i in range(len(sides)) if lo <= sides[i] < hi][:per_bucket]
    big = int(np.argmax(sides))
    if big not in pick:
        pick.append(big)
    Ps = [host_request(int(sides[i]), bool(heat[i]), float(dts[i]), SEEDS["host"] + int(i))
          for i in pick]
    runs = {"batched": ("batched", rtol), "fp64_plain": ("seq_fp64", rtol),
            "fp64_parity": ("seq_fp64", 1e-12)}
    out = {}
    for name, (mode, rt) in runs.items():
        sv = Server(BUDGET, mode=mode, rtol=rt)
        tid = {s: sv.add_topology(S, 0.0) for s, S in systems.items()}
        ids = [sv.submit(tid[P["side"]], kind="heat" if P["dt"] else "poisson",
                         dt=P["dt"] or 0.0, coords=P["coords"], kappa=P["kappa"], f=P["f"],
                         u_prev=P["u_prev"], obs_idx=P["obs"], obs_val=P["obs_val"])
               for P in Ps]
        sv.run()
        out[name] = [sv.result(r) for r in ids]
        sv.close()
    buckets = {f"[{lo},{min(hi, 512)}]" if hi > 512 else f"[{lo},{hi})": (lo, hi)
               for lo, hi in SIDE_BUCKETS}
    per = {b: {n: dict(u=0.0, grad=0.0) for n in runs} for b in buckets}
    for k, P in enumerate(Ps):
        ref = oracle_solve(P)
        b = next(b for b, (lo, hi) in buckets.items() if lo <= P["side"] < hi)
        for n in runs:
            r = out[n][k]
            assert r["status"] == 0, (n, k, r["status"])
            e = per[b][n]
            e["u"] = max(e["u"], relerr(r["u"], ref["u"]))
            e["grad"] = max(e["grad"], relerr(r["grad"], ref["grad_kappa"]))
    worst = {n: dict(u=max(per[b][n]["u"] for b in per), grad=max(per[b][n]["grad"] for b in per))
             for n in runs}
    return dict(requests=len(Ps), sides=[P["side"] for P in Ps],
                reference="CPU oracle: C restatement of the reference ops + SciPy SuperLU",
                rtol={n: rt for n, (_, rt) in runs.items()},
                max_rel_err_u=worst["batched"]["u"], max_rel_err_grad=worst["batched"]["grad"],
                fp64_plain_max_rel_err_u=worst["fp64_plain"]["u"],
                fp64_plain_max_rel_err_grad=worst["fp64_plain"]["grad"],
                fp64_parity_max_rel_err_u=worst["fp64_parity"]["u"],
                fp64_parity_max_rel_err_grad=worst["fp64_parity"]["grad"],
                per_side_bucket=per)


def relerr(a, b):
    a = a.double().cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.double().cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ------------------------------------------------------------------ roofline
def hip_time(fn, repeat):
    """Average seconds per call of fn() over ``repeat`` calls, from HIP events
    recorded on torch's current stream (the stream the library is handed)."""
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / repeat


def roof(nbytes, avg_s, kernel, **extra):
    gbs = nbytes / avg_s / 1e9
    d = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
             frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None, kernel=kernel,
             avg_us=round(avg_s * 1e6, 2), algorithmic_bytes=int(nbytes))
    d.update(extra)
    return d


PMC_SUMMARY = os.path.join(ROOT, "profiles", "bk_step_stream_first1024.md")


def measured_traffic(nsys):
    """HBM bytes per bk_step launch over the stream's first 1024 requests,
    from the committed PMC summary (FETCH_SIZE + WRITE_SIZE, each scaled by
    a copy of known size); None for another batch or without the file."""
    if nsys != 1024 or not os.path.exists(PMC_SUMMARY):
        return None
    for ln in open(PMC_SUMMARY):
        if ln.startswith("traffic_bytes_per_launch"):
            return int(ln.split("=")[1])
    return None


def copy_rate(mib=1024):
    """GB/s of a streaming elementwise kernel (read + write of ``mib`` MiB,
    torch's vectorised elementwise kernel), HIP events: the bandwidth a
    streaming kernel reaches on this GPU."""
    import torch
    src = torch.ones(mib << 17, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    avg = hip_time(lambda: [torch.mul(src, 1.0, out=dst) for _ in range(5)], 5)
    return round(2 * src.numel() * 8 / avg / 1e9, 1)


def batched_spmv_roofline(systems, sides, nsys, repeat=200):
    """bk_step (the batched CG step: SpMV with the iteration's vector updates
    fused in) over the stream's first nsys requests.  The library records
    the HIP events on torch's current stream around the kernel launches only
    (the batch is admitted and assembled beforehand in the same call)."""
    from admfem.server import Server
    sv = Server(BUDGET, mode="batched")
    tid = {s: sv.add_topology(S, 0.2 / s) for s, S in systems.items()}
    topos = [tid[int(s)] for s in sides[:nsys]]
    sv.spmv_repeat(topos, 5)
    nbytes, avg = sv.spmv_repeat(topos, repeat)
    sv.close()
    spmv = sum(systems[int(s)].nnz * 8 + 4 * (systems[int(s)].n + 1) + 8 * systems[int(s)].n
               for s in sides[:nsys])
    r = roof(nbytes, avg, "bk_step", batch=f"the stream's first {nsys} requests, one "
             "Poisson system each, one launch over all of them",
             format="SELL-64, FP32 scaled off-diagonals, FP64 arithmetic; one CG iteration: "
                    "r -= alpha q, y += alpha p, p = r + beta p, q = A p, four dot products",
             bytes_note="canonical CSR CG iteration, nnz*(4+4) + 4(n+1) + 2n*4 for the SpMV "
                        "plus 6n*4 for the vector updates, summed over the systems",
             spmv_only_bytes=int(spmv))
    t = measured_traffic(nsys)
    if t is not None:
        r["traffic"] = t
        r["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy " \
                              "(profiles/bk_step_stream_first1024.md)"
    return r


# ------------------------------------------------------------ precision form
def precision_form(m, rtol, kernels_only=False):
    import ctypes
    import torch
    import admfem
    from admfem import _lib, meshgen
    t0 = time.perf_counter()
    coords, elems = meshgen.jittered_square(m, seed=SEEDS["mesh"])
    torch.cuda.synchronize()
    t_m = time.perf_counter()
    mesh = admfem.Mesh(coords, elems, zero_based=True)
    t_b = time.perf_counter()
    bd = admfem.bcnode(mesh)
    t_s = time.perf_counter()
    S = admfem.DirichletSystem(mesh, bd)
    torch.cuda.synchronize()
    t_e = time.perf_counter()
    lib_setup = dict(mesh=round(t_b - t_m, 3), bcnode=round(t_s - t_b, 3),
                     system=round(t_e - t_s, 3), total=round(t_e - t_m, 3))
    g = admfem.gauss_nodes(mesh)
    kappa = torch.as_tensor(np.exp(meshgen.grf(g, SEEDS["kappa"])), device="cuda")
    kappa2 = torch.as_tensor(np.exp(meshgen.grf(g, SEEDS["kappa_obs"])), device="cuda")
    f = torch.as_tensor(np.random.default_rng(SEEDS["f"]).uniform(-1, 1, mesh.ngauss),
                        device="cuda")
    interior = np.setdiff1d(np.arange(mesh.nnode), bd - 1)
    obs = torch.as_tensor(np.sort(np.random.default_rng(SEEDS["obs"]).choice(
        interior, size=mesh.nnode // 100, replace=False)).astype(np.int32), device="cuda")
    S.assemble(kappa=kappa2, f=f)
    res = {}
    if not kernels_only:
        u2, _ = S.solve(precision="fp64", rtol=1e-12)
        uobs = u2[obs.long()].contiguous()
        ref = S.forward_adjoint(obs, uobs, precision="fp64", rtol=1e-13, kappa=kappa, f=f)
    setup = time.perf_counter() - t0
    log(f"precision form m={m}: setup {setup:.1f}s")
    for prec in () if kernels_only else ("mixed", "fp64"):
        S.forward_adjoint(obs, uobs, precision=prec, rtol=rtol, kappa=kappa, f=f, want_lam=False)
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = S.forward_adjoint(obs, uobs, precision=prec, rtol=rtol, kappa=kappa, f=f,
                                want_lam=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        res[prec] = dict(seconds=round(dt, 4), solves_per_s=round(1.0 / dt, 4),
                         rel_err_u=relerr(out["u"], ref["u"]),
                         rel_err_grad=relerr(out["grad_kappa"], ref["grad_kappa"]),
                         iterations=[out["stats"][0]["iterations"],
                                     out["stats"][1]["iterations"]])
    L = _lib.lib()
    rl = {}
    for prec, code in (("mixed", _lib.PREC_MIXED), ("fp64", _lib.PREC_FP64)):
        nb = ctypes.c_double()
        _lib.check(L.admfem_system_spmv_repeat(S._h, code, 10, ctypes.byref(nb),
                                               _lib.stream_ptr()))
        avg = hip_time(lambda: _lib.check(L.admfem_system_spmv_repeat(
            S._h, code, 200, None, _lib.stream_ptr())), 200)
        rl[prec] = roof(nb.value, avg, "k_mstep (FP32, scaled)" if prec == "mixed" else "k_step (FP64)",
                        bytes_note="one canonical CSR CG iteration: SpMV + 6n*s of vector updates")
    # assembly: fused K, F and Dirichlet into CSR values + rhs, then the SELL copy
    nnz, N, ne, ng = S.nnz, mesh.nnode, mesh.nelem, mesh.ngauss
    b_asm = ng * 8 + 12 * ne + 16 * N + nnz * 8
    b_fused = b_asm + ng * 8 + N * 8
    avg = hip_time(lambda: [S.assemble(kappa=kappa, f=f) for _ in range(20)], 20)
    rl["assembly"] = roof(b_fused, avg, "assemble_kernel (fused K, F, Dirichlet -> SELL)",
                          bytes_note="kappa, f, connectivity, coordinates read; values (once, "
                                     "FP64) and rhs written: the kernel computes K and F",
                          k_only_bytes=int(b_asm),
                          k_only_frac=round(b_asm / avg / 1e9 / HBM_PEAK_GBS, 4),
                          k_only_note="kappa, connectivity, coordinates read, the matrix "
                                      "values written (ngauss*8 + 12*nelem + 16*N + nnz*8): "
                                      "the count for assembling K alone")
    # adjoint gradient dL/dkappa at every Gauss point (the API call: u and
    # lam permuted to the internal node order, then kappa_grad_kernel)
    ug = torch.rand(N, dtype=torch.float64, device="cuda")
    lg = torch.rand(N, dtype=torch.float64, device="cuda")
    S.kappa_gradient(ug, lg)
    avg = hip_time(lambda: [S.kappa_gradient(ug, lg) for _ in range(20)], 20)
    b_grad = 12 * ne + 16 * N + 2 * N * 8 + ng * 8
    rl["kappa_gradient"] = roof(b_grad, avg,
                                "kappa_grad_kernel (+ permuting u, lam to the internal order)",
                                bytes_note="connectivity, coordinates, u and lam read once, "
                                           "dL/dkappa written")
    out = dict(workload=f"{m}x{m}-quad jittered node-permuted mesh ({ne} triangles, {N} "
                        f"nodes, nnz {nnz}), Poisson, kappa=exp(GRF), f~U[-1,1], 1% observed "
                        f"nodes, rtol {rtol:g}; errors vs FP64 CG to 1e-13",
               setup_s=round(setup, 1), roofline=rl,
               library_setup_s=dict(lib_setup, note="Mesh (device topology), bcnode, "
                                    "DirichletSystem (post-BC pattern, SELL layout, assembly "
                                    "blocks) for the node-permuted mesh; setup_s adds the host "
                                    "input generation and the reference solves"))
    if res:
        out.update(mixed=res["mixed"], fp64_plain=res["fp64"],
                   speedup_mixed_vs_fp64=round(res["fp64"]["seconds"] /
                                               res["mixed"]["seconds"], 3))
    return out


# ------------------------------------------------------ operator kernels
def ops_roofline(m1=2048, m2=1024, solve_m=512, repeat=5):
    """Each reference operator's kernel(s) through its C-ABI entry point on
    preallocated device buffers (HIP events on the stream the library is
    handed), against algorithmic bytes: inputs read once, outputs written
    once, with 12 B of connectivity per element and 16 B of coordinates per
    node where the op needs geometry.  P1 at m1 x m1 quads, P2 at m2 x m2;
    the general K\\F op on the Dirichlet-eliminated Poisson COO at solve_m."""
    import ctypes
    import torch
    import admfem
    from admfem import _lib, meshgen
    L, st = _lib.lib(), _lib.stream_ptr
    res = {}

    def run(name, fn, nbytes, **extra):
        fn()
        avg = hip_time(lambda: [fn() for _ in range(repeat)], repeat)
        res[name] = roof(nbytes, avg, name, **extra)

    for deg, m in ((1, m1), (2, m2)):
        c, e = meshgen.jittered_square(m, seed=SEEDS["mesh"])
        mesh = admfem.Mesh(c, e, zero_based=True, degree=deg)
        h, ng, nd, ne, N = mesh.handle, mesh.ngauss, mesh.elem_ndof, mesh.nelem, mesh.ndof
        nt = nd * nd * ng
        geo = 12 * ne + 16 * mesh.nnode
        tag = f"P{deg} {m}x{m}"
        g = torch.rand(ng, dtype=torch.float64, device="cuda") + 0.5
        u = torch.rand(N, dtype=torch.float64, device="cuda")
        ind = torch.empty((nt, 2), dtype=torch.int64, device="cuda")
        vv = torch.empty(nt, dtype=torch.float64, device="cuda")
        P = _lib.ptr
        run(f"{tag} laplace COO (FemLaplaceScalar)",
            lambda: _lib.check(L.admfem_fem_laplace_scalar(h, P(g), P(ind), P(vv), st())),
            nt * 24 + ng * 8 + geo, triplets=nt)
        gk = torch.empty(ng, dtype=torch.float64, device="cuda")
        run(f"{tag} laplace COO grad (FemLaplaceScalarGrad)",
            lambda: _lib.check(L.admfem_fem_laplace_scalar_grad(h, P(vv), P(gk), st())),
            nt * 8 + ng * 8 + geo)
        if deg == 1:  # the Dirichlet op on this COO, above the 256 MB Infinity Cache
            bd1 = admfem.bcnode(mesh)
            nb1 = bd1.size
            bdt1 = torch.as_tensor(bd1.astype(np.int64), device="cuda")
            bdv1 = torch.zeros(nb1, dtype=torch.float64, device="cuda")
            rhs1 = torch.rand(N, dtype=torch.float64, device="cuda")
            oi1 = torch.empty((nt + nb1, 2), dtype=torch.int64, device="cuda")
            ov1 = torch.empty(nt + nb1, dtype=torch.float64, device="cuda")
            orhs1 = torch.empty(N, dtype=torch.float64, device="cuda")
            cnt1 = ctypes.c_int64()

            def dirichlet1():
                _lib.check(L.admfem_impose_dirichlet(P(ind), P(vv), nt, P(bdt1), P(bdv1), nb1,
                                                     P(rhs1), N, P(oi1), P(ov1), P(orhs1),
                                                     ctypes.byref(cnt1), st()))
            dirichlet1()
            run(f"{tag} impose_Dirichlet_boundary_conditions (COO)", dirichlet1,
                nt * 24 + cnt1.value * 24 + 2 * N * 8 + nb1 * 16, triplets_in=nt,
                triplets_out=int(cnt1.value))
            del oi1, ov1, orhs1, rhs1
        del ind, vv
        rhs = torch.empty(N, dtype=torch.float64, device="cuda")
        run(f"{tag} source (FemSourceScalar)",
            lambda: _lib.check(L.admfem_fem_source_scalar(h, P(g), P(rhs), st())),
            ng * 8 + N * 8 + geo)
        run(f"{tag} source grad (FemSourceScalarGrad)",
            lambda: _lib.check(L.admfem_fem_source_scalar_grad(h, P(u), P(gk), st())),
            N * 8 + ng * 8 + geo)
        run(f"{tag} dof_to_gauss_points (DofToGaussPoints)",
            lambda: _lib.check(L.admfem_dof_to_gauss_points(h, P(u), P(gk), st())),
            N * 8 + ng * 8 + 4 * nd * ne)
        run(f"{tag} dof_to_gauss_points grad",
            lambda: _lib.check(L.admfem_dof_to_gauss_points_grad(h, P(g), P(rhs), st())),
            ng * 8 + N * 8 + 4 * nd * ne)
        gg = torch.empty(2 * ng, dtype=torch.float64, device="cuda")
        run(f"{tag} eval_grad_on_gauss_pts1 (FemGrad)",
            lambda: _lib.check(L.admfem_fem_grad(h, P(u), P(gg), st())),
            N * 8 + 2 * ng * 8 + geo)
        run(f"{tag} laplace term, matrix-free K(nu) u (ComputeLaplaceTerm)",
            lambda: _lib.check(L.admfem_fem_laplace_term(h, P(u), P(g), P(rhs), st())),
            N * 8 + ng * 8 + N * 8 + geo)
        run(f"{tag} laplace term grad (ComputeLaplaceTermGrad)",
            lambda: _lib.check(L.admfem_fem_laplace_term_grad(h, P(u), P(u), P(g), P(rhs),
                                                               P(gk), st())),
            2 * N * 8 + ng * 8 + N * 8 + ng * 8 + geo)
        if deg == 2:  # the fused P2 system: assembly into SELL, adjoint gradient
            S = admfem.DirichletSystem(mesh, admfem.bcnode(mesh))
            f2 = torch.rand(ng, dtype=torch.float64, device="cuda") - 0.5
            run(f"{tag} fused system assembly, K, F, Dirichlet -> SELL (assemble_p2_kernel)",
                lambda: S.assemble(kappa=g, f=f2),
                2 * ng * 8 + geo + 24 * ne + S.nnz * 8 + 2 * N * 8, nnz=S.nnz)
            run(f"{tag} fused system dL/dkappa (permute u, lam + kappa_grad_p2_kernel)",
                lambda: S.kappa_gradient(u, u), 2 * N * 8 + geo + 24 * ne + ng * 8)
            S.close()
            del f2
        mesh.close()
        del g, u, gk, gg, rhs
        torch.cuda.empty_cache()
    # Dirichlet on the COO, then the general sparse solve (ADCME SparseSolver)
    c, e = meshgen.jittered_square(solve_m, seed=SEEDS["mesh"])
    mesh = admfem.Mesh(c, e, zero_based=True)
    kap = torch.rand(mesh.ngauss, dtype=torch.float64, device="cuda") + 0.5
    f = torch.rand(mesh.ngauss, dtype=torch.float64, device="cuda") - 0.5
    A = admfem.compute_fem_laplace_matrix1(kap, mesh)
    F = admfem.compute_fem_source_term1(f, mesh)
    bd = admfem.bcnode(mesh)
    sN, N, nb = A.values.numel(), mesh.nnode, bd.size
    bdt = torch.as_tensor(bd.astype(np.int64), device="cuda")
    bdv = torch.zeros(nb, dtype=torch.float64, device="cuda")
    oi = torch.empty((sN + nb, 2), dtype=torch.int64, device="cuda")
    ov = torch.empty(sN + nb, dtype=torch.float64, device="cuda")
    orhs = torch.empty(N, dtype=torch.float64, device="cuda")
    cnt = ctypes.c_int64()
    ind = A.indices.contiguous()
    P = _lib.ptr

    def dirichlet():
        _lib.check(L.admfem_impose_dirichlet(P(ind), P(A.values), sN, P(bdt), P(bdv), nb, P(F),
                                             N, P(oi), P(ov), P(orhs), ctypes.byref(cnt), st()))
    dirichlet()
    run(f"P1 {solve_m}x{solve_m} impose_Dirichlet_boundary_conditions (COO)", dirichlet,
        sN * 24 + cnt.value * 24 + 2 * N * 8 + nb * 16, triplets_in=sN,
        triplets_out=int(cnt.value))
    K = admfem.SparseTensor(oi[:cnt.value], ov[:cnt.value], N)
    out = {}
    for prec in ("fp64", "mixed"):
        admfem.solve(K, orhs, precision=prec, rtol=1e-10)
        torch.cuda.synchronize()
        t = time.perf_counter()
        uu = admfem.solve(K, orhs, precision=prec, rtol=1e-10)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        S = admfem.DirichletSystem(mesh, bd).assemble(kappa=kap, f=f)
        u0, _ = S.solve(precision="fp64", rtol=1e-13)
        out[prec] = dict(seconds=round(dt, 4), rel_err_u=relerr(uu, u0))
    res[f"P1 {solve_m}x{solve_m} K\\F general op (COO in, rtol 1e-10)"] = out
    mesh.close()
    return res


# -------------------------------------------------------------- CPU baseline
def cpu_baseline(sides, heat, dts, budget_s=20.0):
    """The CPU oracle on the first requests of the stream, pinned to one
    core, until ``budget_s`` of CPU work has been done."""
    old = os.sched_getaffinity(0)
    done, spent, nodes = 0, 0.0, 0
    try:
        os.sched_setaffinity(0, {min(old)})
        for i in range(len(sides)):
            P = host_request(int(sides[i]), bool(heat[i]), float(dts[i]), SEEDS["host"] + i)
            t = time.perf_counter()
            oracle_solve(P)
            spent += time.perf_counter() - t
            done += 1
            nodes += (int(sides[i]) + 1) ** 2
            if spent >= budget_s:
                break
    finally:
        os.sched_setaffinity(0, old)
    return dict(value=round(done / spent, 4), unit="solves/s", cores=1, kind="port",
                sample=f"the stream's first {done} requests (sides {sides[:done].min()}-"
                       f"{sides[:done].max()}, {nodes} nodes in all), host-generated inputs, "
                       f"each: mesh precompute, COO emission, Dirichlet, CSR, SuperLU "
                       f"factor+solve, adjoint solve, gradient chain; {spent:.1f} s on 1 core")


def cpu_worker(spec):
    """One process of the all-cores CPU baseline: requests k, k + W, k + 2W,
    ... of the stream through the oracle on core ``core`` until ``budget``
    seconds of work; prints its count and seconds as JSON."""
    k, W, budget, core, n = (float(v) if i == 2 else int(v)
                             for i, v in enumerate(spec.split(",")))
    os.sched_setaffinity(0, {core})
    sides, heat, dts = make_stream(n)
    done, spent, nodes = 0, 0.0, 0
    for i in range(k, n, W):
        P = host_request(int(sides[i]), bool(heat[i]), float(dts[i]), SEEDS["host"] + i)
        t = time.perf_counter()
        oracle_solve(P)
        spent += time.perf_counter() - t
        done += 1
        nodes += (int(sides[i]) + 1) ** 2
        if spent >= budget:
            break
    print(json.dumps(dict(done=done, spent=spent, nodes=nodes)), flush=True)


def cpu_baseline_all_cores(n, budget_s=20.0, max_workers=16):
    """The CPU oracle on every core this process may use (at most
    ``max_workers``), one worker process per core, each on its own share of
    the stream; the aggregate rate is the sum of the workers' rates, measured
    while all of them run."""
    import subprocess
    cores = sorted(os.sched_getaffinity(0))[:max_workers]
    W = len(cores)
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               f"{k},{W},{budget_s},{cores[k]},{n}"],
                              stdout=subprocess.PIPE, text=True) for k in range(W)]
    outs = []
    for p in procs:
        o, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"cpu worker failed with {p.returncode}")
        outs.append(json.loads(o.strip().splitlines()[-1]))
    rate = sum(o["done"] / o["spent"] for o in outs)
    done = sum(o["done"] for o in outs)
    return dict(value=round(rate, 3), unit="solves/s", cores=W, kind="port",
                sample=f"{W} worker processes, one per core, worker k on requests k, k+{W}, ... "
                       f"of the stream until {budget_s:g} s of work each ({done} requests, "
                       f"{sum(o['nodes'] for o in outs)} nodes in all); rate = sum of the "
                       f"workers' rates")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed passes of the stream")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before them")
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--rtol", type=float, default=1e-9,
                    help="true relative residual of every solve")
    ap.add_argument("--fp64-requests", type=int, default=512)
    ap.add_argument("--accuracy-per-bucket", type=int, default=4,
                    help="accuracy requests per side bucket (plus the largest request)")
    ap.add_argument("--roofline-systems", type=int, default=1024)
    ap.add_argument("--precision-m", type=int, default=2048, help="0: skip")
    ap.add_argument("--precision-rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="set up the stream and time only the batched SpMV kernel")
    ap.add_argument("--precision-only", action="store_true",
                    help="run only the single-system precision form (for profiling)")
    ap.add_argument("--ops-only", action="store_true",
                    help="time only the reference operators' kernels (P1, P2, Dirichlet, K\\F)")
    ap.add_argument("--no-ops", action="store_true", help="skip the operator-kernel leg")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the timed passes of the stream at capacity (for profiling)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="with --precision-only: time the SpMV and assembly kernels, no solves")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="with --roofline-only or --precision-only: first copy a MIB-MiB "
                         "buffer once (a plain copy kernel of known size for scaling memory "
                         "counters)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--max-active", type=int, default=MAX_ACTIVE,
                    help="engine slots: requests admitted at once (memory permitting)")
    a = ap.parse_args()
    globals()["MAX_ACTIVE"] = a.max_active
    if a.cpu_worker:
        cpu_worker(a.cpu_worker)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}): the serving engine "
              "and its kernels run on one GPU; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no HIP device"
    if a.calibrate_copy > 0 and (a.roofline_only or a.precision_only):
        src = torch.ones(a.calibrate_copy << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {a.calibrate_copy} MiB read, {a.calibrate_copy} MiB written")
        del src, dst
    if a.precision_only:
        print(json.dumps(precision_form(a.precision_m, a.precision_rtol, a.kernels_only)),
              flush=True)
        return
    if a.ops_only:
        print(json.dumps(ops_roofline()), flush=True)
        return
    n = a.requests
    sides, heat, dts = make_stream(n)
    t0 = time.perf_counter()
    systems = make_topologies(sides)
    setup_s = time.perf_counter() - t0
    log(f"{len(systems)} topologies in {setup_s:.1f}s")
    nroof = min(a.roofline_systems, n)
    if a.roofline_only:
        print(json.dumps(batched_spmv_roofline(systems, sides, nroof)), flush=True)
        return
    idx = np.arange(n)
    zeros = np.zeros(n)
    for w in range(a.warmup):
        wall, rs, st = serve(systems, sides, heat, dts, idx, zeros, "batched", a.rtol)
        log(f"warmup pass {w}: {n / wall:.1f} solves/s")
    walls, last = [], None
    for k in range(a.steps):
        wall, rs, st = serve(systems, sides, heat, dts, idx, zeros, "batched", a.rtol)
        walls.append(wall)
        last = summary(n, wall, rs, st, sides, systems)
        log(f"step {k}: {n / wall:.1f} solves/s, failed {last['failed']}")
    if a.serve_only:
        print(json.dumps(last), flush=True)
        return
    total = float(np.sum(walls))
    value = n * a.steps / total
    rate = 0.8 * value
    arr = np.cumsum(np.random.default_rng(SEEDS["arrivals"]).exponential(1.0 / rate, n))
    wall, rs, st = serve(systems, sides, heat, dts, idx, arr, "batched", a.rtol)
    load = summary(n, wall, rs, st, sides)
    load["arrival_rate_per_s"] = round(rate, 2)
    log(f"load: p50 {load['p50_ms']} ms p99 {load['p99_ms']} ms")
    nf = min(n, a.fp64_requests)
    sub = np.arange(nf)
    wall, rs, st = serve(systems, sides, heat, dts, sub, np.zeros(nf), "seq_fp64", a.rtol)
    fp64 = summary(nf, wall, rs, st, sides[sub])
    wall, rs, st = serve(systems, sides, heat, dts, sub, np.zeros(nf), "batched", a.rtol)
    bsub = summary(nf, wall, rs, st, sides[sub])
    log(f"fp64 plain: {fp64['solves_per_s']} solves/s; batched same subset "
        f"{bsub['solves_per_s']}")
    # latency under load, like for like: each path on the same requests with
    # Poisson arrivals at 80% of its own capacity on them
    lat = {}
    for name, mode, cap in (("fp64_plain", "seq_fp64", fp64["solves_per_s"]),
                            ("batched", "batched", bsub["solves_per_s"])):
        r80 = 0.8 * cap
        arr_s = np.cumsum(np.random.default_rng(SEEDS["arrivals"]).exponential(1.0 / r80, nf))
        wall, rs, st = serve(systems, sides, heat, dts, sub, arr_s, mode, a.rtol)
        lat[name] = summary(nf, wall, rs, st, sides[sub])
        lat[name]["arrival_rate_per_s"] = round(r80, 2)
        log(f"{name} at 80% of its capacity: p50 {lat[name]['p50_ms']} ms "
            f"p99 {lat[name]['p99_ms']} ms")
    acc = accuracy(systems, sides, heat, dts, a.rtol, a.accuracy_per_bucket)
    log(f"accuracy: u {acc['max_rel_err_u']:.2e} grad {acc['max_rel_err_grad']:.2e}")
    rl = batched_spmv_roofline(systems, sides, nroof)
    rl["copy_GBs"] = copy_rate()
    line = {
        "metric": "forward+adjoint solves/s at 1e-8 rel err; p50/p99 latency vs plain FP64 GPU",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * total / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": f"serving: {n} forward+adjoint requests per step, Poisson or implicit "
                        f"heat step (p=1/2, dt~logU[1e-4,1e-1]), side round(exp(U[ln32,ln512])) "
                        f"quads, jittered unit square, kappa=exp(GRF), f~U[-1,1], 1% observed "
                        f"interior nodes, all queued at t=0, {BUDGET / 1e9:g} GB HBM budget",
            "topologies": len(systems),
            "max_active": MAX_ACTIVE,
            "rtol": a.rtol,
            "precision": "FP32 storage, FP64 arithmetic inner CG; FP64 residual replacement "
                         "and solution",
            "outputs": "timed passes keep u and dL/dkappa on the device (no host copies); "
                       "the accuracy leg copies them back",
        },
        "p50_ms": load["p50_ms"],
        "p99_ms": load["p99_ms"],
        "capacity": last,
        "load": load,
        "fp64_plain": fp64,
        "batched_same_subset": bsub,
        "speedup_vs_fp64": round(value / max(fp64["solves_per_s"], 1e-9), 3),
        "speedup_vs_fp64_same_subset": round(bsub["solves_per_s"] /
                                             max(fp64["solves_per_s"], 1e-9), 3),
        "latency_under_load_same_subset": lat,
        "accuracy": acc,
        "roofline": rl,
    }
    if a.precision_m > 0:
        line["precision_form"] = precision_form(a.precision_m, a.precision_rtol)
    if not a.no_ops:
        log("operator kernels")
        try:
            line["operators"] = ops_roofline()
        except Exception as exc:  # reported, never hides the headline line
            line["operators"] = {"error": repr(exc)}
    if not a.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(sides, heat, dts)
        log("cpu baseline, all cores")
        try:
            line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(n)
        except Exception as exc:  # reported, never hides the headline line
            line["cpu_baseline_all_cores"] = {"error": repr(exc)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
