"""CPU tests of the host side: the C-ABI library loads and exports every
entry point declared in include/tfm.h, and the host pattern builder
reproduces the reference's index arrays bit for bit (no GPU needed)."""

import re
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from torchfem import _lib as L
from torchfem._pattern import build_host_pattern

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "tfm.h").read_text()
    declared = set(re.findall(r"\b(tfm_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 19
    lib = L.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert lib.tfm_version() == 1


def test_signature_table_covers_the_header():
    header = (ROOT / "include" / "tfm.h").read_text()
    declared = set(re.findall(r"\b(tfm_[a-z0-9_]+)\s*\(", header))
    assert declared == set(L._SIGNATURES)


def test_pattern_builder_rejects_bad_connectivity():
    el = torch.tensor([[0, 1, 2, 7]])
    with pytest.raises(L.HIPError, match="out of range"):
        build_host_pattern(el, 4, 2)


def _expand(hp, elements, ndof):
    """The documented node-block layout expanded on the host: entry (a, p, b)
    of node n is row n*ndof+a, column adj[p]*ndof+b at val_ptr[n]+a*ndof*deg+p*ndof+b."""
    n_nod = len(hp["adj_ptr"]) - 1
    rows = np.empty(hp["nnz"], np.int64)
    cols = np.empty(hp["nnz"], np.int64)
    rowstart = np.empty(n_nod * ndof, np.int64)
    for n in range(n_nod):
        a0, a1 = hp["adj_ptr"][n], hp["adj_ptr"][n + 1]
        deg = a1 - a0
        vb = hp["val_ptr"][n]
        if deg == 0:
            for a in range(ndof):
                rows[vb + a] = cols[vb + a] = n * ndof + a
                rowstart[n * ndof + a] = vb + a
            continue
        adj = hp["adj"][a0:a1]
        for a in range(ndof):
            rowstart[n * ndof + a] = vb + a * ndof * deg
            q = vb + a * ndof * deg + np.arange(ndof * deg)
            rows[q] = n * ndof + a
            cols[q] = (adj[:, None] * ndof + np.arange(ndof)[None, :]).ravel()
    el = elements.numpy()
    E, npe = el.shape
    epos = hp["epos"].reshape(E, npe, npe).astype(np.int64)
    kmap = (rowstart[(el[:, :, None] * ndof + np.arange(ndof)).reshape(E, npe * ndof)][:, :, None]
            + (np.repeat(epos, ndof, axis=1).repeat(ndof, axis=2) * ndof)
            + np.tile(np.arange(ndof), npe)[None, None, :])
    diag = np.empty(n_nod * ndof, np.int64)
    for n in range(n_nod):
        a0, a1 = hp["adj_ptr"][n], hp["adj_ptr"][n + 1]
        deg = a1 - a0
        for a in range(ndof):
            if deg == 0:
                diag[n * ndof + a] = rowstart[n * ndof + a]
            else:
                p = int(np.searchsorted(hp["adj"][a0:a1], n))
                diag[n * ndof + a] = rowstart[n * ndof + a] + p * ndof + a
    return rows, cols, kmap.reshape(-1), diag


@pytest.mark.parametrize(
    "mesh,ndof",
    [
        (lambda: O.cube_hexa(4, 3, 5), 3),
        (lambda: O.rect_quad(6, 4), 2),
        (lambda: O.cube_tetra6(3), 1),
        (lambda: O.linear_to_quadratic(*O.cube_hexa(3, 3, 3)), 3),
        (lambda: (torch.cat([O.rect_quad(3, 3)[0], torch.tensor([[5.0, 5.0]])]), O.rect_quad(3, 3)[1]), 2),
    ],
)
def test_host_pattern_reproduces_reference_indices(mesh, ndof):
    nodes, el = mesh()
    hp = build_host_pattern(el, len(nodes), ndof)
    _, glob_idx, k_map, diag_map = O.build_pattern(el, len(nodes), ndof)
    rows, cols, kmap, diag = _expand(hp, el, ndof)
    assert np.array_equal(rows, glob_idx[0].numpy())
    assert np.array_equal(cols, glob_idx[1].numpy())
    assert np.array_equal(kmap, k_map.numpy().astype(np.int64))
    assert np.array_equal(diag, diag_map.numpy().astype(np.int64))
    # incidence lists are ascending by element id (the gather order)
    for n in range(len(nodes)):
        s = hp["ne_elem"][hp["ne_ptr"][n]:hp["ne_ptr"][n + 1]]
        assert np.all(np.diff(s) >= 0)


def test_hexa_cube_64_pattern_size_closed_form():
    """At the benchmark size the host pattern has 9 (3n-2)^3 entries (n = 65)."""
    nodes, el = O.cube_hexa(65, 65, 65)
    hp = build_host_pattern(el, len(nodes), 3)
    assert hp["nnz"] == O.hexa_nnz_closed_form(65) == 64_701_513
    assert hp["max_deg"] == 27 and hp["max_inc"] == 8


def test_hyperelastic_material_matches_oracle():
    """Hyperelastic3D's stress and tangent (torch.func on the host side of the
    finite-strain path) against the oracle's hyper_step, per-element
    parameters kept differentiable through vectorize."""
    from torchfem.materials import Hyperelastic3D

    g = torch.Generator().manual_seed(7)
    F = torch.eye(3, dtype=torch.float64) + 0.2 * torch.randn(2, 4, 3, 3, generator=g,
                                                               dtype=torch.float64)
    base = torch.tensor([350.0, 800.0], dtype=torch.float64, requires_grad=True)
    mat = Hyperelastic3D(O.neo_hookean_psi, base).vectorize(4)
    assert mat.params.shape == (4, 2) and mat.is_vectorized
    P, A = mat.stress_tangent(F)
    P_o, A_o = O.hyper_step(O.neo_hookean_psi, F, base.detach())
    assert torch.allclose(P, P_o, rtol=1e-13, atol=1e-10)
    assert torch.allclose(A, A_o, rtol=1e-13, atol=1e-10)
    assert torch.allclose(mat.stress(F), P_o, rtol=1e-13, atol=1e-10)
    # P is linear in (mu, lam): dP/dparams . params == P
    (gp,) = torch.autograd.grad(mat.stress(F).sum(), base)
    assert abs(float(gp @ base.detach()) - float(P_o.sum())) < 1e-9 * float(P_o.abs().sum())


def test_plastic_material_matches_oracle():
    """IsotropicPlasticity3D.step (the host side of the incremental path)
    against the oracle's point-by-point J2 restatement, from a pre-stressed,
    pre-strained state, batched over [points, elements] with a nonlinear
    hardening law."""
    from torchfem.materials import IsotropicPlasticity3D

    def sf(q):
        return 200.0 + 150.0 * (1.0 - torch.exp(-20.0 * q))

    def sfp(q):
        return 3000.0 * torch.exp(-20.0 * q)

    g = torch.Generator().manual_seed(3)
    mag = torch.logspace(-5, -0.5, 5, dtype=torch.float64)[None, :, None, None]
    H = mag * torch.randn(2, 5, 3, 3, generator=g, dtype=torch.float64)
    S = 20.0 * torch.randn(2, 5, 3, 3, generator=g, dtype=torch.float64)
    S = 0.5 * (S + S.transpose(-1, -2))
    q = 0.02 * torch.rand(2, 5, 1, generator=g, dtype=torch.float64)
    mat = IsotropicPlasticity3D(1000.0, 0.3, sf, sfp, tolerance=1e-12, max_iter=50).vectorize(5)
    P, st, A = mat.step(H, None, S, q, None, None, 0)
    P_o, st_o, A_o = O.j2_step(1000.0, 0.3, sf, sfp, tol=1e-12, max_iter=50)(H, None, S, q)
    assert bool((st > q).any()) and bool((st == q).any())  # some points yield, some not
    assert torch.allclose(P, P_o, rtol=1e-11, atol=1e-9)
    assert torch.allclose(st, st_o, rtol=1e-11, atol=1e-14)
    assert torch.allclose(A, A_o, rtol=1e-10, atol=1e-8)
