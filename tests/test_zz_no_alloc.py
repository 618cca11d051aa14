"""The allocation-free entry point tfm_internal_force_static against
tfm_internal_force on the same inputs: two calls in a row on one GPU, the
second on a smaller mesh, both bitwise equal to the entry point with a work
buffer (the same kernels run on the same data), and the size limit."""

import ctypes as C

import pytest
import torch

from torchfem import Solid
from torchfem import _lib as L
from torchfem.base import _MatHolder
from torchfem.materials import IsotropicElasticity3D
from torchfem.mesh import cube_hexa

pytestmark = pytest.mark.gpu
DT = torch.float64


def _forces(model, u, static):
    pat = model._pattern
    kind, p0, p1, sc = model._material_params()
    h = _MatHolder(kind, p0, p1, sc)
    nq = pat.ndof * pat.dim
    f = torch.empty(pat.n_dofs, dtype=DT, device="cuda")
    gm = torch.empty(pat.n_elem, nq, dtype=DT, device="cuda")
    qm = torch.empty(pat.n_elem, nq, dtype=DT, device="cuda")
    if static:
        rc = L.lib().tfm_internal_force_static(pat.stream(), C.byref(pat.mesh), C.byref(h.mat),
                                               L.ptr(u), None, L.ptr(f), L.ptr(gm), L.ptr(qm),
                                               None, None)
    else:
        work = torch.empty(pat.n_elem * pat.npe * pat.ndof, dtype=DT, device="cuda")
        rc = L.lib().tfm_internal_force(pat.stream(), C.byref(pat.mesh), C.byref(h.mat),
                                        L.ptr(u), None, L.ptr(f), L.ptr(gm), L.ptr(qm), None,
                                        None, L.ptr(work))
    L.check(rc)
    torch.cuda.synchronize()
    return f, gm, qm


def test_static_internal_force_matches_the_work_buffer_variant():
    for n, seed in ((10, 0), (4, 1)):  # the second call on a smaller mesh
        nodes, el = cube_hexa(n + 1, n + 1, n + 1)
        E = 500.0 + 1000.0 * torch.rand(len(el), generator=torch.Generator().manual_seed(seed), dtype=DT)
        model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E.cuda(), 0.3))
        u = torch.randn(model.n_dofs, generator=torch.Generator().manual_seed(seed + 7),
                        dtype=DT).cuda()
        ref = _forces(model, u, static=False)
        got = _forces(model, u, static=True)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        assert float(ref[0].abs().max()) > 0.0


def test_static_internal_force_rejects_meshes_beyond_its_buffer():
    n = 82  # 82^3 Hexa1 elements * 8 nodes * 3 DOFs > 4 Mi doubles
    nodes, el = cube_hexa(n + 1, n + 1, n + 1)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    assert model.n_elem * 8 * 3 > 4 * 1024 * 1024
    u = torch.zeros(model.n_dofs, dtype=DT, device="cuda")
    with pytest.raises(L.HIPError, match="static work buffer"):
        _forces(model, u, static=True)
