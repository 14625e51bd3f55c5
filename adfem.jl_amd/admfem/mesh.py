"""`Mesh` and its getters -- the Python mirror of src/MFEM/MFEM.jl.

Node and element ids follow the reference's Julia API: 1-based in ``elems``,
``conn``, ``edges`` and ``bcnode`` results.  The mesh lives on the GPU
(libadmfem handle); getters copy to host numpy arrays as the reference's
ccall getters do (deps/MFEM/API.cpp:17-65).
"""
import ctypes
import os

import numpy as np

from . import _lib
from .meshfile import read_mesh
from .meshgen import rectangle

P1, P2 = 1, 2


class Mesh:
    """P1 or P2 triangle mesh on the GPU.

    ``Mesh(coords, elems, order=-1, degree=1, lorder=-1)``: coords N x 2,
    elems nelem x 3 **1-based** (src/MFEM/MFEM.jl:60-119).
    ``Mesh(m, n, h; version=1)``: the structured generator
    (src/MFEM/MFEM.jl:134-171).
    ``Mesh(filename, file_format=None)``: a mesh file (src/MFEM/MUtils.jl:
    33-50; formats in admfem.meshfile).
    After construction ``elems`` holds the orientation-fixed connectivity, as
    the reference's mfem_get_element_to_vertices overwrite does
    (src/MFEM/MFEM.jl:106).
    """

    def __init__(self, *args, order=-1, degree=1, lorder=-1, version=1, seed=None,
                 zero_based=False, file_format=None):
        if len(args) == 1 and isinstance(args[0], (str, os.PathLike)):
            coords, elems = read_mesh(os.fspath(args[0]), file_format)
            zero_based = True
            if order == -1:
                order = 2  # MUtils.jl:33 defaults order = 2
        elif len(args) == 3 and all(isinstance(a, (int, np.integer)) for a in args[:2]):
            m, n, h = args
            coords, elems = rectangle(int(m), int(n), float(h), version=version, seed=seed)
            zero_based = True
        elif len(args) in (2, 3, 4, 5):
            coords, elems = args[0], args[1]
            if len(args) > 2:
                order = args[2]
            if len(args) > 3:
                degree = args[3]
            if len(args) > 4:
                lorder = args[4]
        else:
            raise TypeError("Mesh(coords, elems[, order, degree, lorder]) or Mesh(m, n, h)")
        if degree in ("P1", "P2"):
            degree = int(degree[1])
        if degree not in (1, 2):
            raise NotImplementedError("only degree 1 (P1) and 2 (P2) elements are built")
        coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 2)
        e0 = np.asarray(elems, dtype=np.int64).reshape(-1, 3)
        if not zero_based:
            e0 = e0 - 1
        e32 = np.ascontiguousarray(e0, dtype=np.int32)
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.admfem_mesh_create(_lib.ptr(coords), coords.shape[0], _lib.ptr(e32),
                                        e32.shape[0], int(order), int(degree), ctypes.byref(h)))
        self._h = h
        info = np.zeros(8, dtype=np.int64)
        _lib.check(L.admfem_mesh_info(h, _lib.ptr(info)))
        self.nnode, self.nelem, self.ngauss, self.ndof, self.elem_ndof, self.nedge = (
            int(v) for v in info[:6])
        self.nnz_full = int(info[6])
        self.nq = int(info[7])
        self.order = (2 if degree == 1 else 4) if order == -1 else int(order)
        self.lorder = 6 if lorder == -1 else int(lorder)
        self.elem_type = P1 if degree == 1 else P2
        self.degree = degree
        self.nodes = coords.copy()
        self._elems = self._conn = self._edges = None

    # The element, dof and edge tables (the Julia struct's fields) are
    # copied from the library on first access.
    @property
    def elems(self):
        """nelem x 3, oriented, 1-based."""
        if self._elems is None:
            el = np.zeros((self.nelem, 3), dtype=np.int32)
            _lib.check(_lib.lib().admfem_mesh_elems(self._h, _lib.ptr(el)))
            self._elems = el.astype(np.int64) + 1
        return self._elems

    @property
    def conn(self):
        """nelem x elem_ndof, 1-based (P1: the vertices; P2: + edge dofs)."""
        if self._conn is None:
            dofs = np.zeros((self.nelem, self.elem_ndof), dtype=np.int32)
            _lib.check(_lib.lib().admfem_mesh_dofs(self._h, _lib.ptr(dofs)))
            self._conn = dofs.astype(np.int64) + 1
        return self._conn

    @property
    def edges(self):
        """nedge x 2, 1-based, MFEM numbering."""
        if self._edges is None:
            ed = np.zeros(2 * self.nedge, dtype=np.int64)
            _lib.check(_lib.lib().admfem_mesh_edges(self._h, _lib.ptr(ed)))
            self._edges = ed.reshape(2, self.nedge).T.copy()
        return self._edges

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().admfem_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):  # src/MFEM/MFEM.jl:195-197
        return int(self.conn.max())

    def pattern(self):
        """Pre-BC CSR pattern (row_ptr, col), 0-based user ids."""
        rp = np.zeros(self.nnode + 1, dtype=np.int32)
        col = np.zeros(self.nnz_full, dtype=np.int32)
        _lib.check(_lib.lib().admfem_mesh_pattern(self._h, _lib.ptr(rp), _lib.ptr(col)))
        return rp, col


def get_ngauss(mesh):
    """src/MFEM/MFEM.jl:207-209"""
    return mesh.ngauss


def gauss_nodes(mesh):
    """src/MFEM/MFEM.jl:226-232: ngauss x 2 Gauss point coordinates."""
    x = np.zeros(mesh.ngauss)
    y = np.zeros(mesh.ngauss)
    _lib.check(_lib.lib().admfem_mesh_gauss(mesh.handle, _lib.ptr(x), _lib.ptr(y)))
    return np.stack([x, y], 1)


def gauss_weights(mesh):
    """src/MFEM/MCore.jl:770-774"""
    w = np.zeros(mesh.ngauss)
    _lib.check(_lib.lib().admfem_mesh_weights(mesh.handle, _lib.ptr(w)))
    return w


def get_area(mesh):
    """src/MFEM/MFEM.jl:216-220"""
    a = np.zeros(mesh.nelem)
    _lib.check(_lib.lib().admfem_mesh_area(mesh.handle, _lib.ptr(a)))
    return a


def fem_nodes(mesh):
    return mesh.nodes


def bcnode(mesh):
    """Boundary dofs, 1-based (src/MFEM/MCore.jl:439-450): the boundary
    vertices ascending (the reference returns a Set's order), then for P2
    nnode + the ids of the boundary edges, ascending."""
    out = np.zeros(mesh.ndof, dtype=np.int64)
    cnt = ctypes.c_int()
    _lib.check(_lib.lib().admfem_mesh_bcnode(mesh.handle, _lib.ptr(out), ctypes.byref(cnt)))
    return out[:cnt.value].copy()


def bcedge(mesh):
    """Boundary edges (k x 2, 1-based, min first): edges used by one element
    (src/MFEM/MCore.jl:385-406)."""
    e = mesh.elems
    pairs = np.concatenate([e[:, [0, 1]], e[:, [2, 1]], e[:, [0, 2]]])
    pairs = np.sort(pairs, axis=1)
    uniq, cnt = np.unique(pairs, axis=0, return_counts=True)
    return uniq[cnt % 2 == 1]
