"""Mesh files -> (coords, elems): the reader behind ``Mesh(filename)``
(src/MFEM/MUtils.jl:33-50).

The reference reads MATLAB ``.mat`` files (``nodes``, ``elems`` 1-based) with
MAT.jl and everything else with the Python package meshio (4.2), keeping the
x, y columns of the points and the triangles.  meshio is not part of this
build, so the formats the reference's own data and examples use are read
here directly:

* ``.stl`` (ASCII or binary): the facets' vertices, identical points merged
  in order of first appearance;
* ``.msh`` (Gmsh 4.1 ASCII): the nodes and every 3-node triangle block;
* ``.mat`` (MATLAB v5 and earlier, through SciPy): ``nodes`` and ``elems``.

Returns coords (N x 2, float64) and elems (nelem x 3, int64, 0-based).
"""
import os
import struct

import numpy as np


def read_mesh(filename, file_format=None):
    fmt = (file_format or os.path.splitext(filename)[1].lstrip(".")).lower()
    if fmt == "stl":
        return _read_stl(filename)
    if fmt in ("msh", "gmsh"):
        return _read_msh(filename)
    if fmt == "mat":
        return _read_mat(filename)
    raise ValueError(f"unsupported mesh file format {fmt!r} (stl, msh, mat)")


def _merge_points(pts):
    """Identical points merged, kept in order of first appearance."""
    _, first, inv = np.unique(pts, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")      # unique groups by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return pts[first[order]], rank[inv.reshape(-1)]


def _read_stl(filename):
    with open(filename, "rb") as fh:
        data = fh.read()
    tris = None
    head = data[:5].lower()
    if head == b"solid" and b"facet" in data[:1024]:
        words = data.split()
        vals = []
        for i, w in enumerate(words):
            if w == b"vertex":
                vals.append((float(words[i + 1]), float(words[i + 2]), float(words[i + 3])))
        tris = np.asarray(vals, dtype=np.float64).reshape(-1, 3, 3)
    else:
        if len(data) < 84:
            raise ValueError("truncated binary STL")
        (nf,) = struct.unpack_from("<I", data, 80)
        rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
        if len(data) < 84 + nf * rec.itemsize:
            raise ValueError("truncated binary STL")
        tris = np.frombuffer(data, dtype=rec, count=nf, offset=84)["v"].astype(np.float64)
    if tris.size == 0:
        raise ValueError("No triangles found in the mesh file.")
    pts, idx = _merge_points(tris.reshape(-1, 3))
    return np.ascontiguousarray(pts[:, :2]), idx.reshape(-1, 3).astype(np.int64)


def _read_msh(filename):
    with open(filename) as fh:
        lines = fh.read().split("\n")
    pos = {ln.strip(): k for k, ln in enumerate(lines) if ln.startswith("$")}
    if "$MeshFormat" not in pos or not lines[pos["$MeshFormat"] + 1].startswith("4"):
        raise ValueError("only Gmsh 4 ASCII .msh files are read")
    k = pos["$Nodes"] + 1
    nblocks, nnodes = (int(v) for v in lines[k].split()[:2])
    k += 1
    tags, xyz = [], []
    for _ in range(nblocks):
        _, _, _, n = (int(v) for v in lines[k].split())
        k += 1
        tags.extend(int(lines[k + j]) for j in range(n))
        xyz.extend([float(v) for v in lines[k + n + j].split()] for j in range(n))
        k += 2 * n
    tags = np.asarray(tags, dtype=np.int64)
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    index = np.full(tags.max() + 1, -1, dtype=np.int64)
    index[tags] = np.arange(tags.size)
    k = pos["$Elements"] + 1
    nblocks = int(lines[k].split()[0])
    k += 1
    tri = []
    for _ in range(nblocks):
        _, _, etype, n = (int(v) for v in lines[k].split())
        k += 1
        if etype == 2:  # 3-node triangle
            for j in range(n):
                tri.append([int(v) for v in lines[k + j].split()[1:4]])
        k += n
    if not tri:
        raise ValueError("No triangles found in the mesh file.")
    return np.ascontiguousarray(xyz[:, :2]), index[np.asarray(tri, dtype=np.int64)]


def _read_mat(filename):
    import scipy.io
    try:
        d = scipy.io.loadmat(filename)
    except NotImplementedError as err:  # v7.3 = HDF5
        raise ValueError("MATLAB v7.3 (HDF5) files are not read; save with -v7") from err
    nodes = np.asarray(d["nodes"], dtype=np.float64)[:, :2]
    elems = np.asarray(d["elems"], dtype=np.int64) - 1
    return np.ascontiguousarray(nodes), elems
