"""Writes the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

c1_m8.npz   Mesh(8,8,1/8) (the reference's rectangle generator, version 1)
            with the README case's coefficients kappa = sin(x)(1+y^2)+1 and
            f = 1e5(x+y) at the Gauss points (README.md:48-49): mesh
            quadrature, Laplace COO, source vector, ImposeDirichlet output,
            the direct solve u, and dL/dkappa of L = sum_obs u^2.
c1_m64.npz  the same case on Mesh(64,64,1/64) (8192 triangles, 4225 nodes):
            u, dL/dkappa and the post-Dirichlet CSR pattern (inputs follow
            from the generator and the coefficient formulas).
heat_m6.npz a jittered 6x6 mesh, implicit heat step A = M + dt K,
            rhs = M u_prev + dt F: the Dirichlet output, u and dL/dkappa.
twoholes_large.npz  the reference's twoholes_large.stl mesh (1303 nodes,
            2432 triangles, two holes) with the Poisson example's
            coefficients: mesh, bcnode, u and dL/dkappa.  Reads the reference
            checkout, so it runs where /root/reference exists.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "adfem.jl_amd"))
sys.path.insert(0, ROOT)

from admfem import meshgen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def readme_case(m):
    c, e = meshgen.rectangle(m, m, 1.0 / m)
    om = O.OracleMesh(c, e)
    x, y = om.gx, om.gy
    kappa = np.sin(x) * (1 + y ** 2) + 1.0
    f = 1e5 * (x + y)
    return c, e, om, kappa, f


def obs_nodes(om, every):
    bd = om.bcnode()
    interior = np.setdiff1d(np.arange(om.nnode), bd - 1)
    return bd, interior[::every].astype(np.int64)


def chain(om, kappa, f, bd, obs, **kw):
    def dl(u):
        g = np.zeros_like(u)
        g[obs] = 2.0 * u[obs]
        return g
    return O.forward_adjoint(om, kappa, f, bd, np.zeros(bd.size), dl, **kw)


def c1(m, full):
    c, e, om, kappa, f = readme_case(m)
    bd, obs = obs_nodes(om, 7)
    r = chain(om, kappa, f, bd, obs)
    A = r["A_bc"].tocsr()
    out = dict(m=m, bd=bd, obs=obs, u=r["u"], grad_kappa=r["grad_kappa"],
               indptr=A.indptr.astype(np.int32), indices=A.indices.astype(np.int32))
    if full:  # inputs too (at m=64 they follow from the generator and formulas)
        out.update(coords=c, elems=e, kappa=kappa, f=f)
        Ki, Kv = om.laplace_coo(kappa)
        F = om.source(f)
        oi, ov, orhs = O.dirichlet(Ki, Kv, bd, F, np.zeros(bd.size))
        out.update(elems_oriented=om.elems, gx=om.gx, gy=om.gy, w=om.w, area=om.area,
                   K_indices=Ki, K_values=Kv, F=F, D_indices=oi, D_values=ov, D_rhs=orhs,
                   lam=r["lam"])
    return out


def heat(m, seed):
    c, e = meshgen.jittered_square(m, seed=seed)
    om = O.OracleMesh(c, e)
    rng = np.random.default_rng(seed)
    kappa = np.exp(meshgen.grf(np.stack([om.gx, om.gy], 1), seed + 1))
    f = rng.uniform(-1, 1, om.ngauss)
    up = np.sin(np.pi * c[:, 0]) * np.sin(np.pi * c[:, 1])
    bd, obs = obs_nodes(om, 3)
    dt = 0.05
    r = chain(om, kappa, f, bd, obs, dt=dt, u_prev=up)
    return dict(m=m, coords=c, elems=e, kappa=kappa, f=f, u_prev=up, dt=dt, bd=bd, obs=obs,
                u=r["u"], grad_kappa=r["grad_kappa"], A_rhs=r["rhs_bc"])


def twoholes():
    """The mesh of docs/src/codes/Poisson/inverse.jl (deps/MFEM/MeshData/
    twoholes_large.stl, read as Mesh(filename) does) with that example's
    coefficients kappa = sin(x)(1+y^2)+1, f = 1e5(x+y), homogeneous
    Dirichlet on bcnode; u and dL/dkappa of L = sum u^2."""
    from admfem.meshfile import read_mesh
    c, e = read_mesh("/root/reference/deps/MFEM/MeshData/twoholes_large.stl")
    om = O.OracleMesh(c, e)
    kappa = np.sin(om.gx) * (1 + om.gy ** 2) + 1.0
    f = 1e5 * (om.gx + om.gy)
    bd = om.bcnode()
    r = O.forward_adjoint(om, kappa, f, bd, np.zeros(bd.size), lambda u: 2 * u)
    return dict(coords=c, elems=e, bd=bd, u=r["u"], grad_kappa=r["grad_kappa"])


def main():
    np.savez_compressed(os.path.join(HERE, "c1_m8.npz"), **c1(8, True))
    np.savez_compressed(os.path.join(HERE, "c1_m64.npz"), **c1(64, False))
    np.savez_compressed(os.path.join(HERE, "heat_m6.npz"), **heat(6, 3))
    np.savez_compressed(os.path.join(HERE, "twoholes_large.npz"), **twoholes())


if __name__ == "__main__":
    main()
