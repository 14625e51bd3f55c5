"""The allocation-free adjoint-gradient entry point
(admfem_system_kappa_gradient_noalloc) against the allocating one
(admfem_system_kappa_gradient), through the C-ABI: two calls in a row, the
second on a smaller system, each bitwise equal to the existing entry point's
output (same kernels, same inputs, a different staging buffer); then a
system beyond the scratch limit is refused with ERR_UNSUPPORTED before any
device work."""
import ctypes

import numpy as np
import pytest
import torch

import admfem
from admfem import _lib, meshgen

pytestmark = pytest.mark.gpu


def system(m, seed):
    c, e = meshgen.jittered_square(m, seed=seed)
    mesh = admfem.Mesh(c, e, zero_based=True)
    return admfem.DirichletSystem(mesh, admfem.bcnode(mesh))


def grads(S, seed):
    rng = np.random.default_rng(seed)
    u = torch.as_tensor(rng.standard_normal(S.n), device="cuda")
    lam = torch.as_tensor(rng.standard_normal(S.n), device="cuda")
    L = _lib.lib()
    out = []
    for fn in (L.admfem_system_kappa_gradient, L.admfem_system_kappa_gradient_noalloc):
        g = torch.full((S.mesh.ngauss,), np.nan, dtype=torch.float64, device="cuda")
        _lib.check(fn(S._h, _lib.ptr(u), _lib.ptr(lam), ctypes.c_double(0.7), _lib.ptr(g),
                      _lib.stream_ptr()))
        torch.cuda.synchronize()
        out.append(g.cpu().numpy())
    return out


def test_noalloc_gradient_matches_allocating_entry_point():
    print("no-alloc gradient: 300x300 then 40x40", flush=True)
    big, small = system(300, 1), system(40, 2)
    for S, seed in ((big, 3), (small, 4)):
        ref, got = grads(S, seed)
        assert np.isfinite(ref).all()
        np.testing.assert_array_equal(got, ref)


def test_noalloc_gradient_refuses_systems_beyond_its_scratch():
    print("no-alloc gradient: 1500x1500 (2,253,001 nodes) refused", flush=True)
    S = system(1500, 5)
    assert S.n > 2_097_152
    u = torch.zeros(S.n, dtype=torch.float64, device="cuda")
    g = torch.zeros(S.mesh.ngauss, dtype=torch.float64, device="cuda")
    rc = _lib.lib().admfem_system_kappa_gradient_noalloc(
        S._h, _lib.ptr(u), _lib.ptr(u), ctypes.c_double(1.0), _lib.ptr(g), _lib.stream_ptr())
    assert rc == _lib.ERR_UNSUPPORTED
    assert "2,097,152" in _lib.lib().admfem_last_error().decode()
