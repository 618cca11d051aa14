"""GPU parity of the individual HIP kernels against the CPU oracle.

Integer outputs (pattern indices) must match bit for bit; FP64 kernels to
1e-12 relative (only the summation order differs from the oracle's einsum /
index_add_); FP32 matrix values must be the exact rounding of the FP64 ones.
"""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle as O
from torchfem import Planar, PlanarHeat, Solid, SolidHeat
from torchfem import _lib as L
from torchfem.base import _MatHolder, _System
from torchfem.materials import (IsotropicConductivity2D, IsotropicConductivity3D,
                                IsotropicElasticity3D, IsotropicElasticityPlaneStress)

pytestmark = pytest.mark.gpu
DT = torch.float64


def _case(name, general=False):
    """(product model, oracle model) pairs on small meshes of every element type."""
    g = torch.Generator().manual_seed(1)
    if name == "Quad1":
        nodes, el = O.rect_quad(5, 4, 2.0, 1.0)
        E = 800.0 + 400.0 * torch.rand(len(el), generator=g, dtype=DT)
        t = 0.5 + torch.rand(len(el), generator=g, dtype=DT)
        mat = IsotropicElasticityPlaneStress(E, 0.3)
        Cfull = O.plane_stress_C(E, torch.tensor(0.3))
        if general:
            Cfull = Cfull.clone()
            Cfull[:, 0, 0, 0, 0] *= 1.5
            Cfull[:, 0, 1, 0, 1] *= 0.7
            Cfull[:, 1, 0, 1, 0] *= 0.7
            Cfull[:, 0, 1, 1, 0] *= 0.7
            Cfull[:, 1, 0, 0, 1] *= 0.7
            mat = IsotropicElasticityPlaneStress(E, 0.3).vectorize(len(el))
            mat.C = Cfull.clone()
        m = Planar(nodes.cuda(), el.cuda(), mat, thickness=t.cuda())
        o = O.OracleModel(nodes, el, O.Quad1, "planar", Cfull, thickness=t)
        return m, o
    if name == "Quad1-heat":
        nodes, el = O.rect_quad(6, 5, 2.0, 1.0)
        k = torch.exp(0.5 * torch.randn(len(el), generator=g, dtype=DT))
        t = 0.5 + torch.rand(len(el), generator=g, dtype=DT)
        m = PlanarHeat(nodes.cuda(), el.cuda(), IsotropicConductivity2D(k), thickness=t.cuda())
        o = O.OracleModel(nodes, el, O.Quad1, "heat", O.iso_KAPPA(k, dim=2), thickness=t)
        return m, o
    if name == "Tetra1-heat":
        nodes, el = O.cube_tetra6(3, jitter=0.1, seed=0)
        k = torch.exp(0.5 * torch.randn(len(el), generator=g, dtype=DT))
        mat = IsotropicConductivity3D(k)
        KAPPA = O.iso_KAPPA(k)
        if general:
            # an anisotropic SPD conductivity exercises the general layout
            R = 0.3 * torch.randn(len(el), 3, 3, generator=g, dtype=DT)
            KAPPA = KAPPA + R @ R.transpose(1, 2)
            mat = IsotropicConductivity3D(k)
            mat.KAPPA = KAPPA.clone()
        m = SolidHeat(nodes.cuda(), el.cuda(), mat)
        o = O.OracleModel(nodes, el, O.Tetra1, "heat", KAPPA)
        return m, o
    if name == "Tetra1":
        nodes, el = O.cube_tetra6(2, jitter=0.1, seed=3)
    elif name == "Hexa1":
        nodes, el = O.cube_hexa(4, 3, 5)
        nodes = nodes + 0.05 * torch.sin(3.0 * nodes[:, [1, 2, 0]])  # distorted elements
    else:
        nodes, el = O.linear_to_quadratic(*O.cube_hexa(3, 3, 4))
        nodes = nodes + 0.03 * torch.sin(3.0 * nodes[:, [1, 2, 0]])
    E = 800.0 + 400.0 * torch.rand(len(el), generator=g, dtype=DT)
    mat = IsotropicElasticity3D(E, 0.3)
    Cfull = O.iso3d_C(E, torch.tensor(0.3))
    if general:
        # an anisotropic tangent exercises the general layout
        Cfull = Cfull.clone()
        Cfull[:, 0, 0, 0, 0] *= 1.7
        Cfull[:, 1, 2, 1, 2] *= 0.6
        Cfull[:, 2, 1, 2, 1] *= 0.6
        Cfull[:, 1, 2, 2, 1] *= 0.6
        Cfull[:, 2, 1, 1, 2] *= 0.6
        mat = IsotropicElasticity3D(E, 0.3).vectorize(len(el))
        mat.C = Cfull.clone()
    m = Solid(nodes.cuda(), el.cuda(), mat)
    o = O.OracleModel(nodes, el, O.ELEMENTS[name], "solid", Cfull)
    return m, o


CASES = [("Quad1", False), ("Quad1", True), ("Quad1-heat", False), ("Tetra1", False),
         ("Tetra1", True), ("Tetra1-heat", False), ("Tetra1-heat", True), ("Hexa1", False),
         ("Hexa1", True), ("Hexa2", False), ("Hexa2", True)]


@pytest.mark.parametrize("name,general", CASES)
def test_device_pattern_is_bit_exact(name, general):
    m, o = _case(name, general)
    assert torch.equal(m.glob_idx.cpu(), o.glob_idx)
    assert torch.equal(m.k_map.cpu(), o.k_map)
    assert torch.equal(m.diag_map.cpu(), o.diag_map)
    assert torch.equal(m.idx.cpu().long(), o.idx)


@pytest.mark.parametrize("name,general", CASES)
def test_element_matrices(name, general):
    m, o = _case(name, general)
    k = m.k0().cpu()
    ko = o.k0()
    assert torch.allclose(k, ko, rtol=1e-12, atol=1e-12 * ko.abs().max())


def _constraints(m, o):
    con = torch.zeros(o.n_nod, o.ndof, dtype=torch.bool)
    con[o.nodes[:, 0] == o.nodes[:, 0].min()] = True
    con[3, 0] = True
    m.constraints = con.cuda()
    return con


@pytest.mark.parametrize("name,general", CASES)
def test_fused_assembly_matches_oracle(name, general):
    m, o = _case(name, general)
    con = _constraints(m, o)
    kind, p0, p1, sc = m._material_params()
    s64 = _System(m, kind, p0, p1, sc, "fp64")
    s32 = _System(m, kind, p0, p1, sc, "mixed")
    cidx = torch.nonzero(con.ravel()).ravel()
    ref = o.assemble_matrix(o.k0(), cidx)
    v = s64.val64.cpu()
    assert torch.allclose(v, ref, rtol=1e-12, atol=1e-12 * ref.abs().max())
    # the FP32 copy is the rounding of the FP64 sums off the nodal diagonal
    # blocks; a diagonal block holds the row sums of the FP64 values (the
    # difference form of the FP32 product)
    rows, cols = m.glob_idx.cpu()
    nd = o.ndof
    diag_blk = (rows // nd) == (cols // nd)
    v32 = s32.val32.cpu()
    assert torch.equal(v32[~diag_blk], v[~diag_blk].float())
    R = torch.zeros(o.n_dofs, nd, dtype=DT).index_put_((rows, cols % nd), v, accumulate=True)
    expect = R[rows[diag_blk], cols[diag_blk] % nd]
    assert torch.allclose(v32[diag_blk].double(), expect, rtol=2.0**-23,
                          atol=1e-12 * float(v.abs().max()))
    # a system for four or more right-hand sides keeps the plain rounding
    s32p = _System(m, kind, p0, p1, sc, "mixed", nrhs=4)
    assert torch.equal(s32p.val32.cpu(), v.float())
    # bitwise reproducible
    s64b = _System(m, kind, p0, p1, sc, "fp64")
    assert torch.equal(s64b.val64, s64.val64)
    # nodal diagonal blocks are the matrix' own diagonal blocks
    nd = o.ndof
    Kd = o.csr(v).toarray()
    blocks = np.stack([Kd[n * nd:(n + 1) * nd, n * nd:(n + 1) * nd] for n in range(o.n_nod)])
    assert np.array_equal(s64.diag.cpu().numpy().reshape(-1, nd, nd), blocks)
    # assemble_matrix(k) from element matrices gives the same values
    K = m.assemble_matrix(m.k0(), cidx.cuda())
    assert torch.allclose(K._values().cpu(), v, rtol=1e-13, atol=1e-13 * ref.abs().max())


@pytest.mark.parametrize("name", ["Quad1", "Quad1-heat", "Tetra1", "Tetra1-heat", "Hexa1",
                                  "Hexa2"])
def test_tiled_and_per_node_assembly_agree(name, monkeypatch):
    """The node-tiled kernel (column lanes, geometric moments) and the
    one-node-per-block kernel assemble the same matrix: FP64 values and nodal
    diagonal blocks to rounding, both bitwise repeatable, with Dirichlet rows
    and a node count that leaves the last tile partly empty."""
    m, o = _case(name)
    _constraints(m, o)
    kind, p0, p1, sc = m._material_params()
    tiled = _System(m, kind, p0, p1, sc, "mixed")
    tiled2 = _System(m, kind, p0, p1, sc, "mixed")
    monkeypatch.setenv("TFM_ASSEMBLE_ROWS", "1")
    rows = _System(m, kind, p0, p1, sc, "mixed")
    rows2 = _System(m, kind, p0, p1, sc, "mixed")
    monkeypatch.delenv("TFM_ASSEMBLE_ROWS")
    for a, b in ((tiled, tiled2), (rows, rows2)):
        assert torch.equal(a.val64, b.val64) and torch.equal(a.val32, b.val32)
        assert torch.equal(a.diag, b.diag)
    scale = float(rows.val64.abs().max())
    assert float((tiled.val64 - rows.val64).abs().max()) < 1e-14 * scale
    assert float((tiled.diag - rows.diag).abs().max()) < 1e-14 * scale
    assert float((tiled.val32.double() - rows.val32.double()).abs().max()) < 2e-7 * scale


def test_tiled_assembly_with_orphan_nodes(monkeypatch):
    """Nodes of no element keep only their diagonal (base.py:89-90): 1 on a
    constrained DOF, 0 otherwise, in both assembly kernels and the oracle.
    The orphans sit in the middle of the numbering, inside a tile."""
    nodes, el = O.cube_hexa(4, 3, 3)
    n0 = len(nodes)
    extra = torch.tensor([[5.0, 5.0, 5.0], [6.0, 6.0, 6.0]], dtype=DT)
    # renumber so that the two orphans become nodes 5 and 6
    perm = torch.cat([torch.arange(5), torch.tensor([n0, n0 + 1]), torch.arange(5, n0)])
    allnodes = torch.cat([nodes, extra])[perm]
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(len(perm))
    el2 = inv[el]
    E = 800.0 + 400.0 * torch.rand(len(el2), generator=torch.Generator().manual_seed(2), dtype=DT)
    m = Solid(allnodes.cuda(), el2.cuda(), IsotropicElasticity3D(E, 0.3))
    o = O.OracleModel(allnodes, el2, O.Hexa1, "solid", O.iso3d_C(E, torch.tensor(0.3)))
    con = torch.zeros(len(allnodes), 3, dtype=torch.bool)
    con[allnodes[:, 0] == 0.0] = True
    con[5, 1] = True  # one constrained DOF of an orphan
    m.constraints = con.cuda()
    kind, p0, p1, sc = m._material_params()
    tiled = _System(m, kind, p0, p1, sc, "mixed")
    monkeypatch.setenv("TFM_ASSEMBLE_ROWS", "1")
    rows = _System(m, kind, p0, p1, sc, "mixed")
    monkeypatch.delenv("TFM_ASSEMBLE_ROWS")
    ref = o.assemble_matrix(o.k0(), torch.nonzero(con.ravel()).ravel())
    for sysm in (tiled, rows):
        v = sysm.val64.cpu()
        assert torch.allclose(v, ref, rtol=1e-12, atol=1e-12 * ref.abs().max())
        D = sysm.diag.cpu().reshape(-1, 3, 3)
        assert torch.equal(D[5], torch.diag(torch.tensor([0.0, 1.0, 0.0], dtype=DT)))
        assert torch.equal(D[6], torch.zeros(3, 3, dtype=DT))


@pytest.mark.parametrize("name,general", CASES)
def test_spmv(name, general):
    m, o = _case(name, general)
    _constraints(m, o)
    kind, p0, p1, sc = m._material_params()
    s64 = _System(m, kind, p0, p1, sc, "fp64")
    s32 = _System(m, kind, p0, p1, sc, "mixed")
    A = o.csr(s64.val64.cpu())
    x = torch.randn(o.n_dofs, generator=torch.Generator().manual_seed(2), dtype=DT)
    ref = A @ x.numpy()
    pat = m._pattern
    for prec, vals, xx, tol in [(0, s64.val64, x.cuda(), 1e-13), (1, s32.val32, x.float().cuda(), 3e-6)]:
        y = torch.empty_like(xx)
        L.check(L.lib().tfm_spmv(pat.stream(), C.byref(pat.mesh), prec, L.ptr(vals), L.ptr(xx), L.ptr(y)))
        err = np.abs(y.double().cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < tol, (prec, err)


@pytest.mark.parametrize("name,general", CASES)
def test_internal_force_and_fluxes(name, general):
    m, o = _case(name, general)
    u = 1e-3 * torch.randn(o.n_dofs, generator=torch.Generator().manual_seed(4), dtype=DT)
    kind, p0, p1, sc = m._material_params()
    h = _MatHolder(kind, p0, p1, sc)
    F, fe, gm, qm, gg, qg = m._forces(h.mat, u.cuda(), means=True, per_gp=True)
    _, fo, grado, fluxo = o.integrate(u, need_k=False)
    Fo = o.assemble_rhs(fo)
    assert torch.allclose(F.cpu(), Fo, rtol=1e-12, atol=1e-12 * Fo.abs().max())
    assert torch.allclose(fe.cpu().reshape(fo.shape), fo, rtol=1e-12, atol=1e-12 * fo.abs().max())
    qo = fluxo.reshape(qg.shape)
    assert torch.allclose(qg.cpu(), qo, rtol=1e-12, atol=1e-12 * qo.abs().max())
    assert torch.allclose(qm.cpu(), qo.mean(0), rtol=1e-12, atol=1e-12 * qo.abs().max())
    Ho = grado.reshape(gg.shape)
    if kind in (L.MAT_ISO, L.MAT_GENERAL):  # mechanics grad = I + H
        Ho = Ho - torch.eye(o.n_dim, dtype=DT).reshape(-1)
    assert torch.allclose(gg.cpu(), Ho, rtol=1e-12, atol=1e-14)
    # K u from the assembled (unconstrained) matrix equals the matrix-free force
    K = o.csr(o.assemble_matrix(o.k0(), torch.zeros(0, dtype=torch.long)))
    Ku = K @ u.numpy()
    assert np.allclose(F.cpu().numpy(), Ku, rtol=1e-11, atol=1e-12 * np.abs(Ku).max())
    # assemble_rhs API
    F2 = m.assemble_rhs(fe.reshape(m.n_elem, -1))
    assert torch.equal(F2, F)


@pytest.mark.parametrize("name,general", CASES)
def test_sensitivity_matches_autograd(name, general):
    """lam^T dK/dp u per element against torch autograd of the oracle's
    element matrices with respect to the material tangent and thickness."""
    m, o = _case(name, general)
    g = torch.Generator().manual_seed(5)
    u = torch.randn(o.n_dofs, generator=g, dtype=DT)
    lam = torch.randn(o.n_dofs, generator=g, dtype=DT)
    kind, p0, p1, sc = m._material_params()
    h = _MatHolder(kind, p0, p1, sc)
    from torchfem.base import _SensView
    out = _SensView(m, h).sensitivity(u.cuda(), lam.cuda(), -1.0)[0].cpu()
    # autograd reference through the oracle's residual lam^T f(u; p)
    mat = o.mat.clone().requires_grad_(True)
    tk = o.thickness.clone().requires_grad_(True) if o.thickness is not None else None
    oo = O.OracleModel(o.nodes, o.elements, o.etype, o.kind, mat, thickness=tk, pattern=o.pattern)
    _, f, _, _ = oo.integrate(u, need_k=False)
    val = -torch.inner(oo.assemble_rhs(f), lam)
    grads = torch.autograd.grad(val, [mat] + ([tk] if tk is not None else []))
    gC = grads[0]
    E = o.n_elem
    d = o.n_dim
    if kind == L.MAT_ISO:
        I = torch.eye(d, dtype=DT)
        I4 = torch.einsum("ij,kl->ijkl", I, I)
        I4S = torch.einsum("ik,jl->ijkl", I, I) + torch.einsum("il,jk->ijkl", I, I)
        ref0 = torch.einsum("eijkl,ijkl->e", gC, I4)
        ref1 = torch.einsum("eijkl,ijkl->e", gC, I4S)
        assert torch.allclose(out[:, 0], ref0, rtol=1e-10, atol=1e-12 * ref0.abs().max())
        assert torch.allclose(out[:, 1], ref1, rtol=1e-10, atol=1e-12 * ref1.abs().max())
    elif kind == L.MAT_GENERAL:
        ref = gC.reshape(E, -1)
        assert torch.allclose(out[:, : d**4], ref, rtol=1e-10, atol=1e-12 * ref.abs().max())
    elif kind == L.MAT_HEAT_ISO:
        ref0 = torch.einsum("eij,ij->e", gC, torch.eye(d, dtype=DT))
        assert torch.allclose(out[:, 0], ref0, rtol=1e-10, atol=1e-12 * ref0.abs().max())
    elif kind == L.MAT_HEAT_GENERAL:
        ref = gC.reshape(E, -1)
        assert torch.allclose(out[:, : d**2], ref, rtol=1e-10, atol=1e-12 * ref.abs().max())
    if tk is not None:
        assert torch.allclose(out[:, -1], grads[1], rtol=1e-10, atol=1e-12 * grads[1].abs().max())


def _pattern_meshes():
    g = torch.Generator().manual_seed(3)
    out = {}
    nodes, el = O.rect_quad(7, 5)
    out["Quad1"] = (nodes, el, 2)
    nodes, el = O.cube_tetra6(4, jitter=0.1)
    out["Tetra1-heat"] = (nodes, el, 1)
    nodes, el = O.cube_hexa(6, 5, 4)
    # elements in random order: the atomics fill each node's incidence list
    # in an arbitrary order, the sort must restore the ascending one
    out["Hexa1-shuffled"] = (nodes, el[torch.randperm(len(el), generator=g)], 3)
    nodes, el = O.linear_to_quadratic(*O.cube_hexa(4, 4, 3))
    out["Hexa2"] = (nodes, el, 3)
    # two nodes that belong to no element keep only their diagonal
    nodes, el = O.cube_hexa(3, 3, 3)
    nodes = torch.cat([nodes, torch.tensor([[5.0, 5.0, 5.0], [6.0, 6.0, 6.0]], dtype=DT)])
    out["Hexa1-orphans"] = (nodes, el, 3)
    return out


@pytest.mark.parametrize("name", list(_pattern_meshes()))
def test_device_pattern_build_equals_the_host_builder(name):
    """tfm_pattern_device_count/fill reproduce tfm_pattern_create's arrays bit
    for bit (base.py:72-122 restated on the device)."""
    from torchfem._pattern import build_device_pattern, build_host_pattern

    nodes, el, ndof = _pattern_meshes()[name]
    host = build_host_pattern(el, len(nodes), ndof)
    dev = build_device_pattern(el.to(torch.int32).cuda().contiguous(), len(nodes), ndof,
                               torch.device("cuda", 0))
    n_adj = int(host["adj_ptr"][-1])
    for key in ("ne_ptr", "ne_elem", "ne_loc", "adj_ptr", "val_ptr"):
        assert np.array_equal(dev[key].cpu().numpy(), host[key]), key
    assert np.array_equal(dev["adj"].cpu().numpy()[:n_adj], host["adj"])
    assert np.array_equal(dev["epos"].cpu().numpy().view(np.uint16), host["epos"])
    for key in ("nnz", "max_deg", "max_inc"):
        assert dev[key] == host[key], key
    rows = np.repeat(np.arange(len(nodes)), np.diff(host["adj_ptr"]))
    assert dev["adj16"] is not None
    assert np.array_equal(dev["adj16"].cpu().numpy()[:n_adj], (host["adj"] - rows).astype(np.int16))


def test_device_pattern_rejects_bad_connectivity():
    from torchfem._pattern import build_device_pattern

    el = torch.tensor([[0, 1, 2, 7]], dtype=torch.int32).cuda()
    with pytest.raises(L.HIPError, match="outside"):
        build_device_pattern(el, 4, 1, torch.device("cuda", 0))


def test_sixteen_bit_neighbour_offsets():
    """The optional adj16 array holds adj[k] - row node exactly when every
    offset fits in 16 bits, and is absent otherwise."""
    from torchfem._pattern import DevicePattern

    dev = torch.device("cuda", 0)
    nodes, el = O.cube_hexa(6, 6, 6)
    dp = DevicePattern(nodes, el, 3, "Hexa1", dev)
    assert dp.adj16 is not None and dp.mesh.adj16
    rows = torch.repeat_interleave(torch.arange(dp.n_nod, device=dev), torch.diff(dp.adj_ptr.long()))
    assert torch.equal(rows + dp.adj16.long()[: len(rows)], dp.adj.long()[: len(rows)])
    # two elements whose nodes are 40,000 apart: offsets beyond int16
    far = torch.tensor([[0, 1, 2, 3, 40000, 40001, 40002, 40003]])
    pts = torch.zeros(40004, 3, dtype=torch.float64)
    dp2 = DevicePattern(pts, far, 3, "Hexa1", dev)
    assert dp2.adj16 is None and not dp2.mesh.adj16
