"""The solver's locality order (FEM._solver_view): a quadratic mesh whose
mid-side nodes follow the corners scatters every row's neighbours, so the
PCG runs on a renumbered copy of the pattern. Results, gradients, batched
load cases and the assembled K must not depend on it.
"""

import pytest
import torch

from torchfem import Solid
from torchfem.elements import linear_to_quadratic
from torchfem.materials import IsotropicElasticity3D
from torchfem.mesh import cube_hexa

pytestmark = pytest.mark.gpu
DT = torch.float64


def rel(a, b):
    return float((a - b).norm() / b.norm())


def _model(renumber: bool):
    nodes, el = linear_to_quadratic(*cube_hexa(25, 25, 25))
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    X = m.nodes
    m.constraints[X[:, 2] == 0.0] = True
    top = X[:, 2] == 1.0
    m.forces[top, 0] = 1e-3
    m.forces[top, 2] = -2e-3
    if not renumber:
        m._sv_checked, m._sv = True, None
    return m


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_renumbered_solve_matches_mesh_order(precision):
    a, b = _model(True), _model(False)
    assert a._pattern.adj16 is None and a._solver_view() is not None
    assert a._solver_view().pattern.adj16 is not None
    out = []
    for m in (a, b):
        E = torch.full((m.n_elem,), 1000.0, dtype=DT, device="cuda", requires_grad=True)
        m.material = IsotropicElasticity3D(E, 0.3)
        u, f, sig, _, _ = m.solve(differentiable_parameters=E, stol=1e-11, precision=precision)
        (g,) = torch.autograd.grad((m.forces * u).sum(), E)
        out.append((u.detach(), f.detach(), sig.detach(), g))
    for x, y in zip(out[0], out[1]):
        assert rel(x, y) < 1e-8
    assert torch.equal(a.K._values(), b.K._values())  # K in the model's order either way


def test_renumbered_load_cases_match_mesh_order():
    a, b = _model(True), _model(False)
    g = torch.Generator(device="cuda").manual_seed(0)
    F = 1e-3 * torch.randn(4, a.n_nod, 3, generator=g, device="cuda", dtype=DT)
    F[:, a.constraints.any(dim=1)] = 0.0
    ua, fa = a.solve_load_cases(F, stol=1e-11, precision="mixed")
    ub, fb = b.solve_load_cases(F, stol=1e-11, precision="mixed")
    assert rel(ua, ub) < 1e-8 and rel(fa, fb) < 1e-8
