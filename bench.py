"""Benchmark: forward+adjoint solves/s of the P1 Poisson path on one MI355X.

Workload (one "step"): on a 1024x1024-quad jittered, node-permuted unit-square
mesh (2,097,152 triangles, 1,050,625 nodes), kappa = exp(GRF, 32 Fourier
modes, sigma 1, length 0.1) and f ~ U[-1,1] at the 3 Gauss points of every
triangle, homogeneous Dirichlet on the boundary:
  assemble K, F -> impose Dirichlet -> u = K^-1 F -> L = sum (u - u_obs)^2
  over a seeded 1% of interior nodes -> lam = K^-1 dL/du -> dL/dkappa.
The product path is the mixed-precision one (FP32 Jacobi-PCG with FP64
residual replacement and FP64 solution accumulation); the plain FP64 path is
timed beside it, and both report their relative error against an FP64
reference (FP64 CG to a 1e-13 true residual; at the start of the run also
checked against the CPU direct solve of the oracle when the cpu_baseline leg
runs).  Inputs are resident in HBM before timing starts.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--m 1024]
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
SEEDS = dict(mesh=20240601, kappa=11, kappa_obs=12, f=13, obs=14)


def build_problem(m):
    import torch
    import admfem
    from admfem import meshgen
    coords, elems = meshgen.jittered_square(m, seed=SEEDS["mesh"])
    mesh = admfem.Mesh(coords, elems, zero_based=True)
    g = admfem.gauss_nodes(mesh)
    kappa = np.exp(meshgen.grf(g, SEEDS["kappa"]))
    kappa_obs = np.exp(meshgen.grf(g, SEEDS["kappa_obs"]))
    f = np.random.default_rng(SEEDS["f"]).uniform(-1.0, 1.0, mesh.ngauss)
    bd = admfem.bcnode(mesh)
    interior = np.setdiff1d(np.arange(mesh.nnode), bd - 1)
    obs = np.sort(np.random.default_rng(SEEDS["obs"]).choice(
        interior, size=max(1, mesh.nnode // 100), replace=False))
    S = admfem.DirichletSystem(mesh, bd)
    dev = dict(kappa=torch.as_tensor(kappa, device="cuda"),
               kappa_obs=torch.as_tensor(kappa_obs, device="cuda"),
               f=torch.as_tensor(f, device="cuda"),
               obs=torch.as_tensor(obs.astype(np.int32), device="cuda"))
    # u_obs: the state of a second seeded kappa (FP64 solve)
    S.assemble(kappa=dev["kappa_obs"], f=dev["f"])
    u2, _ = S.solve(precision="fp64", rtol=1e-12)
    dev["uobs"] = u2[dev["obs"].long()].contiguous()
    host = dict(coords=coords, elems=elems, kappa=kappa, f=f, bd=bd, obs=obs,
                uobs=dev["uobs"].cpu().numpy())
    return mesh, S, dev, host


def run_steps(S, dev, precision, rtol, steps, warmup):
    import torch
    out = None
    for _ in range(warmup):
        out = S.forward_adjoint(dev["obs"], dev["uobs"], precision=precision, rtol=rtol,
                                kappa=dev["kappa"], f=dev["f"], want_lam=False)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        out = S.forward_adjoint(dev["obs"], dev["uobs"], precision=precision, rtol=rtol,
                                kappa=dev["kappa"], f=dev["f"], want_lam=False)
        lat.append(time.perf_counter() - a)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return out, total, np.array(lat)


def relerr(a, b):
    a = a.double().cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.double().cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def spmv_roofline(S, precision, repeat=200):
    """Average duration of the solver's SpMV kernel, timed with HIP events on
    the stream it is launched on (torch's current stream, which the library
    is handed), and its algorithmic-byte rate against HBM peak."""
    import ctypes
    import torch
    from admfem import _lib
    L = _lib.lib()
    prec = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
    nbytes = ctypes.c_double()
    _lib.check(L.admfem_system_spmv_repeat(S._h, prec, 20, ctypes.byref(nbytes), _lib.stream_ptr()))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    _lib.check(L.admfem_system_spmv_repeat(S._h, prec, repeat, None, _lib.stream_ptr()))
    e1.record()
    torch.cuda.synchronize()
    avg_s = e0.elapsed_time(e1) / 1e3 / repeat
    gbs = nbytes.value / avg_s / 1e9
    return dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None,
                kernel=f"k_spmv<{'float' if prec == _lib.PREC_MIXED else 'double'}>",
                avg_us=round(avg_s * 1e6, 2), algorithmic_bytes=int(nbytes.value))


def serving(nreq, fp64_reqs, seed=7):
    """Serving workload: forward+adjoint requests, each Poisson or an implicit
    heat step (M + dt K, dt ~ logU[1e-4, 1e-1]) with p = 1/2, on a jittered
    unit-square mesh whose side is log-uniform in [32, 512] quads (rounded to
    a multiple of 16), kappa = exp(GRF), f ~ U[-1, 1], FP32 inner solves with
    FP64 residual replacement to a 1e-8 true relative residual, 64 GB HBM
    admission budget.  Three runs: capacity (all requests queued at t = 0),
    load (Poisson arrivals at 80% of the measured capacity: p50 / p99
    latency), and the plain FP64 path serving the first requests one at a
    time (its solves/s)."""
    import admfem
    from admfem import meshgen
    from admfem.server import Server
    rng = np.random.default_rng(seed)
    sides = (np.round(np.exp(rng.uniform(np.log(32), np.log(512), nreq)) / 16) * 16)
    sides = np.clip(sides, 32, 512).astype(int)
    heat = rng.random(nreq) < 0.5
    dts = np.exp(rng.uniform(np.log(1e-4), np.log(1e-1), nreq))
    t0 = time.perf_counter()
    systems = {}
    for s in sorted(set(sides.tolist())):
        c, e = meshgen.jittered_square(s, seed=0, jitter=0.0, permute=False)
        mesh = admfem.Mesh(c, e, zero_based=True)
        systems[s] = admfem.DirichletSystem(mesh, admfem.bcnode(mesh))
    setup_s = time.perf_counter() - t0

    def run(mode, n, arrivals):
        sv = Server(64e9, mode=mode, rtol=1e-8)
        tid = {s: sv.add_topology(S, 0.2 / s) for s, S in systems.items()}
        ids = [sv.submit(tid[int(sides[i])], kind="heat" if heat[i] else "poisson",
                         dt=float(dts[i]), seed=1000 + i, arrival=float(arrivals[i]),
                         want_u=False, want_grad=False) for i in range(n)]
        t = time.perf_counter()
        sv.run()
        wall = time.perf_counter() - t
        rs = [sv.result(i) for i in ids]
        bad = sum(r["status"] != 0 for r in rs)
        lat = np.array([r["latency"] for r in rs])
        return dict(requests=n, failed=bad, wall_s=round(wall, 3),
                    solves_per_s=round(n / max(wall, 1e-9), 2),
                    p50_ms=round(1e3 * float(np.percentile(lat, 50)), 2),
                    p99_ms=round(1e3 * float(np.percentile(lat, 99)), 2),
                    peak_active=int(sv.stats()["peak_active"]),
                    peak_gb=round(sv.stats()["peak_bytes"] / 1e9, 3),
                    mean_batch=round(sv.stats()["mean_batch"], 1))

    cap = run("batched", nreq, np.zeros(nreq))
    rate = 0.8 * cap["solves_per_s"]
    arr = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, nreq))
    load = run("batched", nreq, arr)
    load["arrival_rate_per_s"] = round(rate, 2)
    nf = min(nreq, fp64_reqs)
    fp64 = run("seq_fp64", nf, np.zeros(nf))
    cap_sub = run("batched", nf, np.zeros(nf))
    return dict(workload=f"{nreq} forward+adjoint requests, Poisson or heat (p=1/2), side "
                         f"logU[32,512] (x16), jittered square, kappa=exp(GRF), rtol 1e-8, "
                         f"64 GB budget", topologies=len(systems), setup_s=round(setup_s, 1),
                capacity=cap, load=load, fp64_plain_subset=fp64, batched_subset=cap_sub,
                speedup_vs_fp64_plain=round(cap_sub["solves_per_s"] /
                                            max(fp64["solves_per_s"], 1e-9), 2))


def cpu_baseline(host, m):
    """One forward+adjoint of the same workload with the CPU oracle (serial C
    restatement of the reference loops + SciPy SuperLU), pinned to one core."""
    from oracle import oracle as O
    cores = 1
    old = os.sched_getaffinity(0)
    try:
        os.sched_setaffinity(0, {min(old)})
        t0 = time.perf_counter()
        om = O.OracleMesh(host["coords"], host["elems"])
        obs, uobs = host["obs"], host["uobs"]

        def dloss(u):
            g = np.zeros_like(u)
            g[obs] = 2.0 * (u[obs] - uobs)
            return g

        res = O.forward_adjoint(om, host["kappa"], host["f"], host["bd"],
                                np.zeros(host["bd"].size), dloss)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    return dict(value=round(1.0 / dt, 5), unit="solves/s", cores=cores, kind="port",
                sample=f"one forward+adjoint at m={m} (mesh precompute, COO emission, "
                       f"Dirichlet, CSR, SuperLU factor+solve, adjoint solve, gradient chain), "
                       f"{dt:.1f} s on 1 core"), res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=1024)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=512,
                    help="serving workload size (0: skip; 4096 for the full stream)")
    ap.add_argument("--serving-fp64-requests", type=int, default=48,
                    help="requests served by the plain FP64 path for the comparison")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}): this build is "
              "single-GPU; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no HIP device"
    m = a.m
    print(f"[bench] building m={m} problem", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    mesh, S, dev, host = build_problem(m)
    print(f"[bench] setup {time.perf_counter() - t0:.1f}s; reference solve", file=sys.stderr,
          flush=True)
    ref = S.forward_adjoint(dev["obs"], dev["uobs"], precision="fp64", rtol=1e-13,
                            kappa=dev["kappa"], f=dev["f"])
    torch.cuda.synchronize()
    print(f"[bench] mixed steps={a.steps}", file=sys.stderr, flush=True)
    out, total, lat = run_steps(S, dev, "mixed", a.rtol, a.steps, a.warmup)
    err_u = relerr(out["u"], ref["u"])
    err_g = relerr(out["grad_kappa"], ref["grad_kappa"])
    print(f"[bench] fp64 steps={a.fp64_steps}", file=sys.stderr, flush=True)
    out64, total64, lat64 = run_steps(S, dev, "fp64", a.rtol, a.fp64_steps, 1)
    err_u64 = relerr(out64["u"], ref["u"])
    err_g64 = relerr(out64["grad_kappa"], ref["grad_kappa"])
    roof = spmv_roofline(S, "mixed")
    roof64 = spmv_roofline(S, "fp64")
    value = a.steps / total
    value64 = a.fp64_steps / total64
    line = {
        "metric": "forward+adjoint solves/s at 1e-8 rel err",
        "value": round(value, 4),
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
            "workload": f"P1 Poisson forward+adjoint, {m}x{m}-quad jittered node-permuted "
                        f"mesh ({mesh.nelem} triangles, {mesh.nnode} nodes), kappa=exp(GRF), "
                        f"f~U[-1,1], 1% observed nodes, Jacobi-PCG rtol {a.rtol:g}",
            "precision": "FP32 inner CG + FP64 residual replacement",
            "nnz": S.nnz,
        },
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "rel_err_u": err_u,
        "rel_err_grad": err_g,
        "iterations": [out["stats"][0]["iterations"], out["stats"][1]["iterations"]],
        "fp64_plain": {
            "value": round(value64, 4),
            "ms_per_step": round(1e3 * total64 / a.fp64_steps, 3),
            "p50_ms": round(1e3 * float(np.percentile(lat64, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat64, 99)), 3),
            "rel_err_u": err_u64,
            "rel_err_grad": err_g64,
            "iterations": [out64["stats"][0]["iterations"], out64["stats"][1]["iterations"]],
            "roofline": roof64,
        },
        "speedup_vs_fp64": round(value / value64, 3),
        "roofline": roof,
    }
    if a.serving_requests > 0:
        print(f"[bench] serving {a.serving_requests} requests", file=sys.stderr, flush=True)
        line["serving"] = serving(a.serving_requests, a.serving_fp64_requests)
    if not a.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb, cres = cpu_baseline(host, m)
        line["cpu_baseline"] = cb
        line["rel_err_u_vs_cpu_direct"] = relerr(out["u"], cres["u"])
        line["rel_err_grad_vs_cpu_direct"] = relerr(out["grad_kappa"], cres["grad_kappa"])
        line["fp64_plain"]["rel_err_u_vs_cpu_direct"] = relerr(out64["u"], cres["u"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
