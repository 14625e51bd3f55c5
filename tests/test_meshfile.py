"""Mesh files (Mesh(filename), src/MFEM/MUtils.jl:33-50): the readers of
admfem.meshfile on files written here in each format, and on the reference's
own twoholes_large.stl through the committed fixture."""
import os
import struct

import numpy as np
import pytest
import scipy.io

from admfem import meshgen
from admfem.meshfile import read_mesh
from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mesh():
    return meshgen.jittered_square(5, seed=4, permute=True)


def _write_stl_ascii(path, c, e):
    with open(path, "w") as fh:
        fh.write("solid test\r\n")
        for t in e:
            fh.write("facet normal 0 0 1\r\n  outer loop\r\n")
            for v in t:
                fh.write(f"    vertex {float(c[v, 0])!r} {float(c[v, 1])!r} 0\r\n")
            fh.write("  endloop\r\nendfacet\r\n")
        fh.write("endsolid test\r\n")


def _first_appearance(c, e):
    """Reference numbering: vertices in order of first appearance in the
    facet list."""
    seen, order = {}, []
    for v in e.reshape(-1):
        if v not in seen:
            seen[v] = len(order)
            order.append(v)
    return c[order], np.vectorize(seen.get)(e)


def test_stl_ascii_and_binary(tmp_path):
    c, e = _mesh()
    want_c, want_e = _first_appearance(c, e)
    p = tmp_path / "m.stl"
    _write_stl_ascii(p, c, e)
    rc, re_ = read_mesh(str(p))
    np.testing.assert_array_equal(rc, want_c)
    np.testing.assert_array_equal(re_, want_e)
    c32 = c.astype(np.float32).astype(np.float64)     # binary STL holds float32
    b = tmp_path / "b.stl"
    with open(b, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", e.shape[0]))
        for t in e:
            fh.write(struct.pack("<3f", 0, 0, 1))
            for v in t:
                fh.write(struct.pack("<3f", c32[v, 0], c32[v, 1], 0.0))
            fh.write(struct.pack("<H", 0))
    bc, be = read_mesh(str(b))
    wc, we = _first_appearance(c32, e)
    np.testing.assert_array_equal(bc, wc)
    np.testing.assert_array_equal(be, we)


def test_msh4_triangles_and_mat(tmp_path):
    c, e = _mesh()
    p = tmp_path / "m.msh"
    n = c.shape[0]
    with open(p, "w") as fh:
        fh.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        fh.write(f"$Nodes\n1 {n} 1 {n}\n2 1 0 {n}\n")
        fh.write("".join(f"{k + 1}\n" for k in range(n)))
        fh.write("".join(f"{float(x)!r} {float(y)!r} 0\n" for x, y in c))
        fh.write("$EndNodes\n")
        fh.write(f"$Elements\n2 {e.shape[0] + 1} 1 {e.shape[0] + 1}\n")
        fh.write("1 1 1 1\n1 1 2\n")                  # a line block first, skipped
        fh.write(f"2 1 2 {e.shape[0]}\n")
        fh.write("".join(f"{k + 2} {a + 1} {b + 1} {d + 1}\n" for k, (a, b, d) in enumerate(e)))
        fh.write("$EndElements\n")
    mc, me = read_mesh(str(p))
    np.testing.assert_array_equal(mc, c)
    np.testing.assert_array_equal(me, e)
    m = tmp_path / "m.mat"
    scipy.io.savemat(m, {"nodes": c, "elems": e + 1})
    mc, me = read_mesh(str(m))
    np.testing.assert_array_equal(mc, c)
    np.testing.assert_array_equal(me, e)


def test_twoholes_fixture_is_a_valid_mesh_and_reproduced_by_oracle():
    d = np.load(os.path.join(G, "twoholes_large.npz"))
    c, e = d["coords"], d["elems"]
    assert c.shape == (1303, 2) and e.shape == (2432, 3)
    om = O.OracleMesh(c, e)
    assert (om.area > 0).all()
    # Euler: V - E + F = 1 - holes for a planar mesh with two holes
    assert om.nnode - om.edges().shape[0] + om.nelem == 1 - 2
    np.testing.assert_array_equal(np.sort(om.bcnode()), np.sort(d["bd"]))
    kappa = np.sin(om.gx) * (1 + om.gy ** 2) + 1.0
    f = 1e5 * (om.gx + om.gy)

    def dl(u):
        return 2 * u
    r = O.forward_adjoint(om, kappa, f, d["bd"], np.zeros(d["bd"].size), dl)
    np.testing.assert_allclose(r["u"], d["u"], rtol=1e-12, atol=1e-12 * np.abs(d["u"]).max())
    g0 = d["grad_kappa"]
    np.testing.assert_allclose(r["grad_kappa"], g0, rtol=1e-9, atol=1e-12 * np.abs(g0).max())
