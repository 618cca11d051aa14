"""Edge cases of the GPU path against the CPU oracle: zero loads (with and
without an initial guess), single-element models, fully constrained models
and a batch of load cases that holds a zero case."""

import pytest
import torch

import oracle as O
from torchfem import Planar, Solid, SolidHeat
from torchfem.elements import linear_to_quadratic
from torchfem.materials import (IsotropicConductivity3D, IsotropicElasticity3D,
                                IsotropicElasticityPlaneStress)
from torchfem.mesh import cube_hexa, cube_tetra6, rect_quad
from torchfem.sparse import sparse_solve

pytestmark = pytest.mark.gpu
DT = torch.float64
PRECISIONS = ["fp64", "mixed"]


def rel(a, b):
    a = a.detach().cpu()
    b = b.detach().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_zero_load_gives_zero_fields_and_gradient(precision):
    """No forces and no prescribed displacements: u, f, sigma and the
    gradient of sum(u) with respect to E are exactly zero (the reference's
    Newton loop stops at once, base.py:760-770), and nothing raises."""
    nodes, el = cube_hexa(5, 5, 5)
    E = torch.full((len(el),), 1000.0, dtype=DT, device="cuda", requires_grad=True)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E, 0.3))
    model.constraints[model.nodes[:, 2] == 0.0, :] = True
    u, f, sigma, _, _ = model.solve(differentiable_parameters=E, precision=precision)
    (g,) = torch.autograd.grad(u.sum(), E)
    assert float(u.detach().abs().max()) == 0.0
    assert float(f.detach().abs().max()) == 0.0
    assert float(sigma.detach().abs().max()) == 0.0
    assert float(g.abs().max()) == 0.0


@pytest.mark.parametrize("method", ["cg", "minres"])
def test_zero_rhs_returns_zero_despite_initial_guess(method):
    """b = 0 gives x = 0 whatever x0 is, as scipy.sparse.linalg.cg/minres
    return when ||b|| = 0 (the solvers behind sparse.py:490-512)."""
    torch.manual_seed(0)
    A = torch.randn(8, 8, dtype=DT)
    A = A @ A.T + 8 * torch.eye(8, dtype=DT)
    idx = torch.cartesian_prod(torch.arange(8), torch.arange(8)).T
    As = torch.sparse_coo_tensor(idx, A.flatten(), A.shape).coalesce().cuda()
    x, _ = sparse_solve(As, torch.zeros(8, dtype=DT, device="cuda"), method=method,
                        x0=torch.randn(8, dtype=DT).cuda())
    assert float(x.abs().max()) == 0.0


def _single(etype):
    """One element of each type, clamped on its faces (edges) at coordinate 0
    along the returned axes, loaded at the node farthest from the origin.
    The 20-node element under its 2x2x2 rule (elements.py:1383-1406) has
    zero-energy modes that one clamped face leaves singular, so it is
    clamped on three faces."""
    if etype == "Quad1":
        nodes, el = rect_quad(2, 2, 2.0, 1.0)
        return nodes, el, O.Quad1, "planar", (0,)
    if etype == "Tetra1":
        nodes = torch.tensor([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.2, 0.1, 0.9]], dtype=DT)
        return nodes, torch.tensor([[0, 1, 2, 3]]), O.Tetra1, "solid", (2,)
    nodes, el = cube_hexa(2, 2, 2, 1.0, 2.0, 0.5)
    if etype == "Hexa2":
        nodes, el = linear_to_quadratic(nodes, el)
        return nodes, el, O.Hexa2, "solid", (0, 1, 2)
    return nodes, el, O.Hexa1, "solid", (0,)


@pytest.mark.parametrize("etype", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_single_element_matches_oracle(etype, precision):
    nodes, el, oet, kind, axes = _single(etype)
    if kind == "planar":
        model = Planar(nodes.cuda(), el.cuda(), IsotropicElasticityPlaneStress(1000.0, 0.3))
        C = O.plane_stress_C(torch.full((1,), 1000.0, dtype=DT), torch.tensor(0.3))
        o = O.OracleModel(nodes, el, oet, kind, C, thickness=torch.ones(1, dtype=DT))
    else:
        model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
        o = O.OracleModel(nodes, el, oet, kind, O.iso3d_C(torch.full((1,), 1000.0, dtype=DT), torch.tensor(0.3)))
    assert model.etype.__name__ == etype
    X = model.nodes
    for axis in axes:
        model.constraints[X[:, axis] == 0.0, :] = True
    far = int(torch.argmax(X.sum(dim=1)))
    model.forces[far, :] = torch.linspace(-1.0, 0.5, model.n_dof_per_node, dtype=DT, device="cuda")
    u, f, sigma, Fg, _ = model.solve(precision=precision, stol=1e-12)
    uo, fo, so, go = o.solve(model.constraints.cpu(), model.forces.cpu(),
                             torch.zeros_like(model.forces).cpu())
    assert rel(u, uo) < 1e-9
    assert rel(f, fo) < 1e-9
    assert rel(sigma, so) < 1e-9
    assert rel(Fg, go) < 1e-12


@pytest.mark.parametrize("precision", PRECISIONS)
def test_fully_constrained_model_returns_prescribed_field(precision):
    """Every DOF prescribed: the system is the identity, u equals the
    prescribed field and f holds the internal forces K u (base.py:844-855)."""
    nodes, el = cube_tetra6(3, jitter=0.1, seed=2)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    model.constraints[:, :] = True
    g = torch.Generator().manual_seed(3)
    model.displacements = 1e-3 * torch.randn(len(nodes), 3, generator=g, dtype=DT).cuda()
    u, f, sigma, _, _ = model.solve(precision=precision)
    o = O.OracleModel(nodes, el, O.Tetra1, "solid",
                      O.iso3d_C(torch.full((len(el),), 1000.0, dtype=DT), torch.tensor(0.3)))
    uo, fo, so, _ = o.solve(model.constraints.cpu(), model.forces.cpu(), model.displacements.cpu())
    assert torch.equal(u.cpu(), model.displacements.cpu())
    assert rel(f, fo) < 1e-12
    assert rel(sigma, so) < 1e-12


@pytest.mark.parametrize("precision", PRECISIONS)
def test_batch_with_a_zero_load_case(precision):
    """Three heat load cases, the middle one zero: its temperatures and its
    share of the gradient are zero, the others match single solves."""
    nodes, el = cube_hexa(7, 7, 7)
    kap = torch.full((len(el),), 2.0, dtype=DT, device="cuda", requires_grad=True)
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(kap))
    X = model.nodes
    model.constraints[X[:, 0] == 0.0] = True
    g = torch.Generator().manual_seed(4)
    Fb = torch.randn(3, len(nodes), 1, generator=g, dtype=DT).cuda()
    Fb[:, X[:, 0] == 0.0] = 0.0
    Fb[1] = 0.0
    T, _ = model.solve_load_cases(Fb, differentiable_parameters=kap, precision=precision, stol=1e-10)
    assert float(T[1].detach().abs().max()) == 0.0
    (gb,) = torch.autograd.grad(T[1].sum(), kap, retain_graph=True)
    assert float(gb.abs().max()) == 0.0
    for k in (0, 2):
        model.heat_flux = Fb[k]
        Tk, _, _, _, _ = model.solve(precision="fp64", stol=1e-12)
        assert rel(T[k], Tk) < 1e-8


@pytest.mark.parametrize("ncase", [3, 6])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_batched_tetra_cases_match_single_solves(precision, ncase):
    """Batched solves on a Tetra1 mesh (about 15 neighbours per node: the
    8-lane teams, with four and eight padded rhs) agree with single solves."""
    nodes, el = cube_tetra6(5, jitter=0.1, seed=5)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    X = model.nodes
    model.constraints[X[:, 2] == 0.0, :] = True
    g = torch.Generator().manual_seed(6)
    Fb = torch.randn(ncase, len(nodes), 3, generator=g, dtype=DT).cuda()
    Fb[:, X[:, 2] == 0.0] = 0.0
    u, f = model.solve_load_cases(Fb, precision=precision, stol=1e-10)
    for k in range(ncase):
        model.forces = Fb[k]
        uk, fk, _, _, _ = model.solve(precision="fp64", stol=1e-12)
        assert rel(u[k], uk) < 1e-8
        assert rel(f[k], fk) < 1e-8
