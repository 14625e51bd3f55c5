"""CPU checks of the C-ABI boundary: the library loads and exports every
symbol include/admfem.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "admfem.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)
    skip = {"if", "defined", "sizeof", "return"}
    out = []
    for n in names:
        if n in skip or n.isupper():
            continue
        out.append(n)
    return sorted(set(out))


def test_header_declares_reference_symbols():
    names = declared_functions()
    for ref in ("init_nnfem_mesh", "mfem_get_ngauss", "mfem_get_gauss", "mfem_get_gauss_weights",
                "mfem_get_area", "mfem_get_elem_ndof", "mfem_get_ndof", "mfem_get_connectivity",
                "mfem_get_element_to_vertices", "FemLaplaceScalar_forward_Julia",
                "FemSourceScalar_forward_Julia"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    from admfem import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from admfem import _lib, server
    bound = set(_lib.exported_symbols()) | set(server.exported_symbols())
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    from admfem import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_product_never_imports_oracle():
    pkg = os.path.join(ROOT, "adfem.jl_amd")
    for d, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(d, f), errors="replace").read()
                assert "oracle" not in re.sub(r"(#|//).*", "", txt).lower() or f == "build.py", f


@pytest.mark.parametrize("bad", [{"degree": 3}, {"degree": "BDM1"}])
def test_mesh_rejects_unsupported_degree_without_gpu(bad):
    from admfem import Mesh
    with pytest.raises(NotImplementedError):
        Mesh(2, 2, 0.5, **bad)
