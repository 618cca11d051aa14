"""Closed-form solutions on the benchmark meshes at their full sizes.

The configuration tests compare the GPU path with itself (mixed against
FP64, homogeneity, finite differences of its own solves), which would not
notice an error of assembly or SpMV that only appears at large sizes (index
widths, 16-bit neighbour offsets, grid-stride spans, the per-XCD split).
Here every mesh of the benchmarks is solved with boundary data whose exact
solution lies in the finite-element space, so the discrete solution equals
it up to the solver tolerance:

* patch tests (tests/test_models.py:54-71 of the reference): a linear field
  prescribed on the whole boundary of a homogeneous body is reproduced in
  the interior and gives the constant stress C:eps (or flux KAPPA grad T);
* layered bodies: with a modulus (conductivity) that varies from one layer
  of elements to the next along x, the uniaxial-strain field u = (w(x), 0, 0)
  with w' = s0 / M(x), M = lambda + 2G, is in equilibrium (sigma_xx = s0 in
  every layer; sigma_yy = sigma_zz = lambda w' depend on x only), and it is
  piecewise linear with kinks on element faces, so the elements reproduce
  it exactly; the same holds for T(x) with T' = q0 / kappa(x).

FP64 solves run to a relative residual of 1e-12 and must match to 1e-10;
mixed-precision solves run to 1e-10 and must match to 1e-8.
"""

import pytest
import torch

from torchfem import Solid, SolidHeat
from torchfem.elements import linear_to_quadratic
from torchfem.materials import IsotropicConductivity3D, IsotropicElasticity3D
from torchfem.mesh import cube_hexa, cube_tetra6

pytestmark = pytest.mark.gpu
DT = torch.float64


@pytest.fixture(autouse=True, params=["block_jacobi", "two_level"])
def preconditioner(request):
    """Every full-size case runs with both preconditioners of the FEM solves
    (the default, "auto", would pick the two-level one at these sizes)."""
    from torchfem.sparse import default_preconditioner, set_default_preconditioner
    old = default_preconditioner()
    set_default_preconditioner(request.param)
    yield request.param
    set_default_preconditioner(old)
CASES = [("fp64", 1e-12, 1e-10), ("mixed", 1e-10, 1e-8)]

G_LIN = torch.tensor([[1.0e-3, 2.0e-4, -1.0e-4], [3.0e-4, -1.0e-3, 5.0e-5],
                      [1.0e-4, 2.0e-4, 7.0e-4]], dtype=DT)
C_LIN = torch.tensor([2.0e-3, -1.0e-3, 5.0e-4], dtype=DT)


def rel(a, b):
    a = a.detach().double()
    b = b.detach().double()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def boundary(nodes):
    mask = torch.zeros(len(nodes), dtype=torch.bool, device=nodes.device)
    for d in range(nodes.shape[1]):
        c = nodes[:, d]
        mask |= (c == c.min()) | (c == c.max())
    return mask


def iso_stress(E, nu, eps):
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    return lam * torch.trace(eps) * torch.eye(3, dtype=DT, device=eps.device) + 2.0 * G * eps


def run_patch(model, E, nu, precision, stol, tol):
    Gm = G_LIN.to(model.device)
    u_exact = model.nodes @ Gm.T + C_LIN.to(model.device)
    bnd = boundary(model.nodes)
    model.constraints = bnd[:, None].repeat(1, 3)
    model.displacements = torch.where(bnd[:, None], u_exact, torch.zeros_like(u_exact))
    u, f, sigma, _, _ = model.solve(precision=precision, stol=stol)
    relres = model.last_solve_info["forward"]["pcg"][-1]["relres"][0]
    assert relres <= stol, relres
    assert rel(u, u_exact) < tol, rel(u, u_exact)
    s_exact = iso_stress(E, nu, 0.5 * (Gm + Gm.T))
    assert rel(sigma, s_exact.expand_as(sigma)) < tol
    # no loads inside: the internal forces at interior nodes are the
    # residual of the solve, far below the boundary reactions
    inner = ~bnd
    assert float(f[inner].abs().max()) < 100 * tol * float(f[bnd].abs().max())


@pytest.mark.parametrize("precision,stol,tol", CASES)
def test_hexa1_64_patch(precision, stol, tol):
    """Config-2 mesh: Hexa1 64^3 (274,625 nodes, 823,875 DOFs)."""
    nodes, el = cube_hexa(65, 65, 65)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    assert model.n_dofs == 823875
    run_patch(model, 1000.0, 0.3, precision, stol, tol)


@pytest.mark.parametrize("precision,stol,tol", CASES)
def test_hexa2_32_patch(precision, stol, tol):
    """Config-5 mesh: Hexa2 32^3 (140,481 nodes, 421,443 DOFs)."""
    nodes, el = linear_to_quadratic(*cube_hexa(33, 33, 33))
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    assert model.n_dofs == 421443
    run_patch(model, 1000.0, 0.3, precision, stol, tol)


@pytest.mark.parametrize("precision,stol,tol", CASES)
def test_tetra1_80_jittered_patch(precision, stol, tol):
    """Config-4 mesh (Kuhn split of 80^3 cells, jittered interior nodes) as an
    elastic solid: the linear field is reproduced on the distorted tets."""
    nodes, el = cube_tetra6(80, jitter=0.1, seed=0)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    run_patch(model, 1000.0, 0.3, precision, stol, tol)


@pytest.mark.parametrize("anisotropic", [False, True])
@pytest.mark.parametrize("precision,stol,tol", CASES)
def test_tetra1_80_jittered_heat_patch(anisotropic, precision, stol, tol):
    """Config-4 mesh (531,441 DOFs): T = a.x + c on the boundary of a body of
    constant conductivity gives that T inside and the flux KAPPA a in every
    element (the reference's flux sign, conductivity.py:90-92)."""
    nodes, el = cube_tetra6(80, jitter=0.1, seed=0)
    assert len(nodes) == 531441
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(1.7))
    if anisotropic:
        kap = torch.tensor([[2.0, 0.3, -0.2], [0.3, 1.5, 0.1], [-0.2, 0.1, 1.2]], dtype=DT)
        model.material.KAPPA = kap.cuda().expand(len(el), 3, 3).contiguous()
    else:
        kap = 1.7 * torch.eye(3, dtype=DT)
    a = torch.tensor([0.7, -0.4, 0.25], dtype=DT, device="cuda")
    T_exact = (model.nodes @ a + 0.1)[:, None]
    bnd = boundary(model.nodes)
    model.constraints = bnd[:, None].clone()
    model.temperatures = torch.where(bnd[:, None], T_exact, torch.zeros_like(T_exact))
    T, f, q, _, _ = model.solve(precision=precision, stol=stol)
    assert rel(T, T_exact) < tol, rel(T, T_exact)
    q_exact = (kap.cuda() @ a).expand(len(el), 3)
    assert rel(q, q_exact) < tol, rel(q, q_exact)
    assert float(f[~bnd].abs().max()) < 100 * tol * float(f[bnd].abs().max())


@pytest.mark.parametrize("precision,stol,tol", CASES)
def test_simp48_layered_modulus(precision, stol, tol):
    """Config-3 mesh (Hexa1 48^3, 352,947 DOFs) with a SIMP modulus
    1000 rho^3, rho ~ U(0.3, 1) drawn per layer of elements along x (seed 0):
    u = (w(x), 0, 0) with w' = s0 / M(x) is exact, every element carries
    sigma_xx = s0, sigma_yy = sigma_zz = lambda_e s0 / M_e."""
    n, nu, s0 = 48, 0.3, -2.0
    nodes, el = cube_hexa(n + 1, n + 1, n + 1)
    rho = 0.3 + 0.7 * torch.rand(n, generator=torch.Generator().manual_seed(0), dtype=DT)
    E_layer = 1000.0 * rho**3
    M_layer = E_layer * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    # layer index of an element from its first node's x (cube_hexa numbers
    # node (i, j, k) as i*Ny*Nz + j*Nz + k, so i = x-layer)
    layer = el[:, 0] // ((n + 1) * (n + 1))
    E = E_layer[layer]
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E.cuda(), nu))
    # w at the layer interfaces x_i = i / n, w(0) = 0
    w_nodes = torch.cat([torch.zeros(1, dtype=DT), torch.cumsum(s0 / M_layer / n, 0)])
    i_node = torch.round(nodes[:, 0] * n).long()
    u_exact = torch.zeros(len(nodes), 3, dtype=DT)
    u_exact[:, 0] = w_nodes[i_node]
    u_exact = u_exact.cuda()
    bnd = boundary(model.nodes)
    model.constraints = bnd[:, None].repeat(1, 3)
    model.displacements = torch.where(bnd[:, None], u_exact, torch.zeros_like(u_exact))
    u, f, sigma, _, _ = model.solve(precision=precision, stol=stol)
    assert rel(u, u_exact) < tol, rel(u, u_exact)
    lam_e = (E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))).cuda()
    M_e = M_layer[layer].cuda()
    s_exact = torch.zeros(len(el), 3, 3, dtype=DT, device="cuda")
    s_exact[:, 0, 0] = s0
    s_exact[:, 1, 1] = lam_e * s0 / M_e
    s_exact[:, 2, 2] = lam_e * s0 / M_e
    assert rel(sigma, s_exact) < tol
    assert float(f[~bnd].abs().max()) < 100 * tol * float(f[bnd].abs().max())


@pytest.mark.parametrize("precision,stol,tol", CASES)
def test_tetra1_80_layered_conductivity(precision, stol, tol):
    """Config-4 mesh without jitter (layers of cells must be planar) with
    kappa = exp(N(0, 0.5^2)) per layer of cells along x (seed 1):
    T' = q0 / kappa(x), the flux is (q0, 0, 0) in every tetrahedron."""
    n, q0 = 80, 1.3
    nodes, el = cube_tetra6(n, jitter=0.0)
    kap_layer = torch.exp(0.5 * torch.randn(n, generator=torch.Generator().manual_seed(1), dtype=DT))
    # the cell of a tetrahedron is that of its node v0 (the cell's corner)
    layer = el[:, 0] // ((n + 1) * (n + 1))
    kap = kap_layer[layer]
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(kap.cuda()))
    T_nodes = torch.cat([torch.zeros(1, dtype=DT), torch.cumsum(q0 / kap_layer / n, 0)])
    T_exact = T_nodes[torch.round(nodes[:, 0] * n).long()][:, None].cuda()
    bnd = boundary(model.nodes)
    model.constraints = bnd[:, None].clone()
    model.temperatures = torch.where(bnd[:, None], T_exact, torch.zeros_like(T_exact))
    T, f, q, _, _ = model.solve(precision=precision, stol=stol)
    assert rel(T, T_exact) < tol, rel(T, T_exact)
    q_exact = torch.zeros(len(el), 3, dtype=DT, device="cuda")
    q_exact[:, 0] = q0
    assert rel(q, q_exact) < tol, rel(q, q_exact)
