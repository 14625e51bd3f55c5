"""Serving engine on the GPU: results against the CPU oracle, the same
result alone and in a mix, budget admission, failure statuses."""
import numpy as np
import pytest

import admfem
from admfem import meshgen
from admfem.server import Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def topo(m, seed=0):
    c, e = meshgen.jittered_square(m, seed=seed, jitter=0.0, permute=False)
    mesh = admfem.Mesh(c, e, zero_based=True)
    S = admfem.DirichletSystem(mesh, admfem.bcnode(mesh))
    return c, e, S


def host_problem(c, e, seed, heat=False):
    rng = np.random.default_rng(seed)
    m = int(round(np.sqrt(c.shape[0]))) - 1
    h = 1.0 / m
    cc = c.copy()
    inner = (c[:, 0] > 0) & (c[:, 0] < 1) & (c[:, 1] > 0) & (c[:, 1] < 1)
    cc[inner] += rng.uniform(-0.2 * h, 0.2 * h, size=(inner.sum(), 2))
    om = O.OracleMesh(cc, e)
    kap = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), seed + 1))
    f = rng.uniform(-1, 1, om.ngauss)
    bd = om.bcnode()
    interior = np.setdiff1d(np.arange(om.nnode), bd - 1)
    obs = rng.choice(interior, size=max(3, interior.size // 20), replace=False).astype(np.int32)
    uobs = rng.standard_normal(obs.size) * 1e-3
    up = np.sin(np.pi * cc[:, 0]) * np.sin(np.pi * cc[:, 1]) if heat else None
    dt = 0.01 if heat else None
    return dict(coords=cc, om=om, kappa=kap, f=f, bd=bd, obs=obs, uobs=uobs, u_prev=up, dt=dt)


def oracle_ref(P):
    def dl(u):
        g = np.zeros_like(u)
        g[P["obs"]] = 2 * (u[P["obs"]] - P["uobs"])
        return g
    kw = {} if P["dt"] is None else dict(dt=P["dt"], u_prev=P["u_prev"])
    return O.forward_adjoint(P["om"], P["kappa"], P["f"], P["bd"], np.zeros(P["bd"].size), dl,
                             **kw)


def submit(sv, tid, P, arrival=0.0):
    return sv.submit(tid, kind="heat" if P["dt"] else "poisson", dt=P["dt"] or 0.0,
                     arrival=arrival, coords=P["coords"], kappa=P["kappa"], f=P["f"],
                     u_prev=P["u_prev"], obs_idx=P["obs"], obs_val=P["uobs"])


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("mode", ["batched", "seq_fp64"])
def test_server_matches_oracle(mode):
    print(f"server vs oracle, mode={mode}", flush=True)
    sv = Server(2 << 30, mode=mode, rtol=1e-11)
    cases = []
    for m, heat, seed in ((12, False, 1), (20, True, 2), (16, False, 3), (9, True, 4)):
        c, e, S = topo(m)
        tid = sv.add_topology(S, jitter=0.2 / m)
        P = host_problem(c, e, seed, heat)
        cases.append((submit(sv, tid, P), P, S))
    sv.run()
    for rid, P, S in cases:
        r = sv.result(rid)
        assert r["done"] and r["status"] == 0, r
        ref = oracle_ref(P)
        assert rel(r["u"], ref["u"]) < 1e-8
        assert rel(r["grad"], ref["grad_kappa"]) < 1e-8
        assert r["loss"] == pytest.approx(((ref["u"][P["obs"]] - P["uobs"]) ** 2).sum(), rel=1e-8)


def test_result_same_alone_and_in_mix():
    c, e, S = topo(24)
    c2, e2, S2 = topo(40)
    P = host_problem(c, e, 7)
    alone = Server(1 << 30, rtol=1e-10)
    t = alone.add_topology(S, 0.2 / 24)
    ra = submit(alone, t, P)
    alone.run()
    mix = Server(1 << 30, rtol=1e-10)
    t1 = mix.add_topology(S, 0.2 / 24)
    t2 = mix.add_topology(S2, 0.2 / 40)
    ids = [mix.submit(t2, seed=s, arrival=0.0) for s in range(3)]
    rm = submit(mix, t1, P, arrival=0.0)
    ids += [mix.submit(t1 if s % 2 else t2, kind="heat" if s % 3 == 0 else "poisson", dt=0.05,
                       seed=100 + s, arrival=0.0) for s in range(6)]
    mix.run()
    a, b = alone.result(ra), mix.result(rm)
    assert a["status"] == 0 and b["status"] == 0
    assert mix.stats()["peak_active"] > 1
    np.testing.assert_array_equal(a["u"], b["u"])       # bit-identical
    np.testing.assert_array_equal(a["grad"], b["grad"])
    assert all(mix.result(i)["status"] == 0 for i in ids)


def test_budget_admission_and_rejection():
    c, e, S = topo(32)
    probe = Server(1 << 30)
    t = probe.add_topology(S, 0.2 / 32)
    rid = probe.submit(t, seed=1)
    probe.run()
    need = probe.result(rid)["device_bytes"]
    # room for exactly one request at a time: all complete, never two at once
    one = Server(1.5 * need)
    t = one.add_topology(S, 0.2 / 32)
    ids = [one.submit(t, seed=s) for s in range(4)]
    one.run()
    bad = [(i, one.result(i)["status"], one.result(i)["iters_fwd"], one.result(i)["iters_adj"])
           for i in ids if one.result(i)["status"] != 0]
    assert not bad, bad
    assert one.stats()["peak_active"] == 1
    # a request that can never fit is rejected, not hung
    tiny = Server(need / 4)
    t = tiny.add_topology(S, 0.2 / 32)
    rid = tiny.submit(t, seed=3)
    tiny.run(timeout=60)
    assert tiny.result(rid)["status"] == admfem._lib.ERR_BUDGET


def test_failure_statuses():
    c, e, S = topo(16)
    P = host_problem(c, e, 5)
    bad = dict(P)
    bad["kappa"] = P["kappa"].copy()
    bad["kappa"][7] = np.nan
    sv = Server(1 << 30, rtol=1e-10)
    t = sv.add_topology(S, 0.2 / 16)
    rn = submit(sv, t, bad)
    sv.run(timeout=60)
    assert sv.result(rn)["status"] == admfem._lib.ERR_NAN
    few = Server(1 << 30, rtol=1e-12, max_iter=3)
    t = few.add_topology(S, 0.2 / 16)
    rf = submit(few, t, P)
    few.run(timeout=60)
    assert few.result(rf)["status"] == admfem._lib.ERR_NOT_CONVERGED


def test_largest_stream_requests_match_oracle():
    """The serving stream's largest sides: a Poisson request at side 480 and
    a heat step at side 511, batched together at the benchmark's rtol 1e-9,
    each within 1e-8 of the oracle's direct solve (u and dL/dkappa)."""
    cases = []
    sv = Server(8 << 30, rtol=1e-9)
    for m, heat, seed in ((480, False, 61), (511, True, 62)):
        c, e, S = topo(m)
        P = host_problem(c, e, seed, heat=heat)
        if heat:
            P["dt"] = 0.05
        t = sv.add_topology(S, 0.2 / m)
        cases.append((m, P, submit(sv, t, P)))
    sv.run(timeout=120)
    for m, P, rid in cases:
        r = sv.result(rid)
        assert r["status"] == 0, (m, r["status"])
        ref = oracle_ref(P)
        eu, eg = rel(r["u"], ref["u"]), rel(r["grad"], ref["grad_kappa"])
        print(f"side {m} {'heat' if P['dt'] else 'poisson'}: rel err u {eu:.2e} grad {eg:.2e}, "
              f"iterations {r['iters_fwd']}/{r['iters_adj']}", flush=True)
        assert eu < 1e-8 and eg < 1e-8


@pytest.mark.parametrize("mode", ["batched", "seq_fp64"])
def test_bad_requests_fail_alone(mode):
    """A host request without kappa, or with an observation index outside
    [0, n) or repeated, is refused at submit; a device request with such an
    index is admitted and finishes with ERR_ARG.  Either way it holds no slot
    or memory afterwards: the valid requests around it complete and match the
    oracle."""
    import torch
    c, e, S = topo(16)
    P = host_problem(c, e, 9)
    ref = oracle_ref(P)
    sv = Server(1 << 30, mode=mode, rtol=1e-11, max_active=2)
    t = sv.add_topology(S, 0.2 / 16)
    for kw in (dict(kappa=None), dict(obs=np.append(P["obs"], S.n).astype(np.int32)),
               dict(obs=np.append(P["obs"], P["obs"][0]).astype(np.int32))):
        Q = dict(P)
        Q.update(kw)
        with pytest.raises(admfem.AdmfemError) as ei:
            submit(sv, t, Q)
        assert ei.value.code == admfem._lib.ERR_ARG
    dev = lambda a: torch.as_tensor(a, device="cuda")  # noqa: E731
    ids = []
    for k, bad_obs in enumerate((None, np.append(P["obs"], -1), None,
                                 np.append(P["obs"], P["obs"][1]), None)):
        o = P["obs"] if bad_obs is None else bad_obs.astype(np.int32)
        uo = np.append(P["uobs"], np.zeros(o.size - P["obs"].size))
        ids.append((bad_obs is None, sv.submit(
            t, coords=dev(P["coords"]), kappa=dev(P["kappa"]), f=dev(P["f"]),
            obs_idx=dev(o), obs_val=dev(uo))))
    sv.run(timeout=120)
    for ok, rid in ids:
        r = sv.result(rid)
        assert r["done"]
        if ok:
            assert r["status"] == 0
            assert rel(r["u"], ref["u"]) < 1e-8 and rel(r["grad"], ref["grad_kappa"]) < 1e-8
        else:
            assert r["status"] == admfem._lib.ERR_ARG


def test_poisson_arrivals_complete_with_latency():
    tops = [topo(m) for m in (16, 32, 48)]
    sv = Server(4 << 30, rtol=1e-8)
    tids = [sv.add_topology(S, 0.2 / m) for (c, e, S), m in zip(tops, (16, 32, 48))]
    rng = np.random.default_rng(0)
    arr = np.cumsum(rng.exponential(0.002, size=30))
    ids = [sv.submit(tids[i % 3], kind="heat" if i % 2 else "poisson", dt=0.01, seed=i,
                     arrival=float(a)) for i, a in enumerate(arr)]
    sv.run()
    rs = [sv.result(i) for i in ids]
    bad = [(i, r["status"], r["done"], r["iters_fwd"], r["iters_adj"]) for i, r in enumerate(rs)
           if r["status"] != 0 or not r["done"]]
    assert not bad, bad
    assert all(r["t_admit"] >= r["t_arrival"] - 1e-9 and r["t_done"] >= r["t_admit"] for r in rs)
    assert all(r["relres_fwd"] <= 1e-8 and r["relres_adj"] <= 1e-8 for r in rs)


def test_graph_replay_matches_eager():
    """Chunks replayed from captured HIP graphs (padded grids, dummy slot)
    give bit-identical results to eager launches, across many replays with
    systems joining and leaving."""
    tops = [topo(m) for m in (12, 20, 28)]
    out = []
    for g in (False, True):
        sv = Server(2 << 30, rtol=1e-10, use_graph=g, chunk=4)
        tids = [sv.add_topology(S, 0.2 / m) for (c, e, S), m in zip(tops, (12, 20, 28))]
        ids = [sv.submit(tids[i % 3], kind="heat" if i % 2 else "poisson", dt=0.03, seed=50 + i,
                         arrival=0.0) for i in range(9)]
        sv.run()
        out.append([sv.result(i) for i in ids])
    for a, b in zip(*out):
        assert a["status"] == 0 and b["status"] == 0
        np.testing.assert_array_equal(a["u"], b["u"])
        np.testing.assert_array_equal(a["grad"], b["grad"])
