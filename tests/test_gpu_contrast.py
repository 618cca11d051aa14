"""Density designs with void elements: rho in {1e-3, 1} (E = 1000 rho^3, a
stiffness contrast of 1e9, the lower bound of the reference's bracket
topology example, examples/optimization/solid/bracket.ipynb). FP32 rounding
of the stiff entries exceeds the stiffness of the void elements, so the FP32
operator of a mixed solve is not positive definite on this design and its
iterations break down; the solve finishes in FP64 from the last refined x,
and later solves with the same matrix (the adjoint) run in FP64 directly.
Displacements and compliance sensitivities against the oracle's direct solve
and the closed form -3 rho^2 u_e^T k_e u_e, in both precisions, with the
block-Jacobi and the two-level preconditioner (where the coarse correction
does not reach the tolerance on such a design, block-Jacobi continues from
its iterate)."""

import pytest
import torch

import oracle as O
from torchfem import Solid
from torchfem.materials import IsotropicElasticity3D
from torchfem.mesh import cube_hexa

pytestmark = pytest.mark.gpu
DT = torch.float64


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def _void_design(n, seed=0):
    nodes, el = cube_hexa(n + 1, n + 1, n + 1)
    g = torch.Generator().manual_seed(seed)
    rho = torch.where(torch.rand(len(el), generator=g) < 0.5, 1e-3, 1.0).to(DT)
    return nodes, el, rho


@pytest.fixture(scope="module")
def oracle_24():
    nodes, el, rho = _void_design(24)
    o = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(1000.0 * rho**3, torch.tensor(0.3)))
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 0] == 0.0] = True
    F = torch.zeros(len(nodes), 3, dtype=DT)
    face = nodes[:, 0] == 1.0
    F[face, 2] = -1.0 / int(face.sum())
    uo, _, _, _ = o.solve(con, F, torch.zeros_like(F))
    unit = O.OracleModel(nodes, el, O.Hexa1, "solid",
                         O.iso3d_C(torch.ones(len(el), dtype=DT), torch.tensor(0.3)),
                         pattern=o.pattern)
    go = O.simp_compliance_sensitivity(unit, uo, rho, 1000.0)
    return nodes, el, rho, con, F, uo, go


@pytest.mark.parametrize("preconditioner", ["block_jacobi", "two_level"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_void_design_solve_and_sensitivity(oracle_24, precision, preconditioner):
    nodes, el, rho0, con, F, uo, go = oracle_24
    rho = rho0.clone().cuda().requires_grad_(True)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0 * rho**3, 0.3))
    model.constraints = con.cuda()
    model.forces = F.cuda()
    u, f, _, _, _ = model.solve(differentiable_parameters=rho, precision=precision,
                                preconditioner=preconditioner, stol=1e-10)
    info = model.last_solve_info
    (g,) = torch.autograd.grad(torch.inner(model.forces.ravel(), u.ravel()), rho)
    assert rel(u, uo) < 1e-7
    # the void elements' sensitivities are 1e-6 of the solid ones: compare
    # each group on its own scale
    solid = rho0 > 0.5
    assert rel(g[solid.cuda()], go[solid]) < 1e-7
    assert rel(g[~solid.cuda()], go[~solid]) < 1e-5
    fb = info["forward"]["pcg"][0]["fp64_fallback"]
    if preconditioner == "two_level":
        return
    if precision == "mixed":
        # the FP32 iterations break down on this design and FP64 finishes;
        # the adjoint, with the same matrix, then runs in FP64 from the start
        assert fb > 0
        assert info["adjoint"]["precision"] == "fp64"
    else:
        assert fb == 0
