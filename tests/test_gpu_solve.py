"""End-to-end GPU parity of FEM.solve() and its adjoint gradients against the
CPU oracle, on reduced versions of the benchmark configurations and on the
reference's own known-answer tests (README golden, patch test)."""

import math

import pytest
import torch

import oracle as O
from torchfem import Planar, Solid, SolidHeat
from torchfem.elements import linear_to_quadratic
from torchfem.materials import (IsotropicConductivity3D, IsotropicElasticity3D,
                                IsotropicElasticityPlaneStress)
from torchfem.mesh import cube_hexa, cube_tetra6, rect_quad

pytestmark = pytest.mark.gpu
DT = torch.float64
PRECISIONS = ["fp64", "mixed"]


def rel(a, b):
    a = a.detach().cpu()
    b = b.detach().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_readme_golden_compliance_gradient(precision):
    """README.md:101-113: dC/dthickness = [-0.0208, -0.0053]."""
    nodes = torch.tensor([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, 1.0]])
    elements = torch.tensor([[0, 1, 4, 3], [1, 2, 5, 4]])
    cant = Planar(nodes.cuda(), elements.cuda(), IsotropicElasticityPlaneStress(E=1000.0, nu=0.3))
    cant.forces[5, 1] = -1.0
    cant.constraints[[0, 3], :] = True
    cant.thickness.requires_grad = True
    u, f, _, _, _ = cant.solve(differentiable_parameters=cant.thickness, precision=precision)
    compliance = torch.inner(f.ravel(), u.ravel())
    (g,) = torch.autograd.grad(compliance, cant.thickness)
    assert torch.allclose(g.cpu(), torch.tensor([-0.0208, -0.0053], dtype=DT), atol=5e-5)


def test_outputs_detached_without_declared_parameters():
    """tests/test_gradients.py:161-176."""
    nodes = torch.tensor([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, 1.0]])
    elements = torch.tensor([[0, 1, 4, 3], [1, 2, 5, 4]])
    cant = Planar(nodes.cuda(), elements.cuda(), IsotropicElasticityPlaneStress(E=1000.0, nu=0.3))
    cant.forces[5, 1] = -1.0
    cant.constraints[[0, 3], :] = True
    rho = torch.ones(2, device="cuda", requires_grad=True)
    cant.thickness = rho**3
    out = cant.solve()
    assert all(not t.requires_grad for t in out)


def _boundary(nodes):
    mask = torch.zeros(len(nodes), dtype=torch.bool, device=nodes.device)
    for d in range(nodes.shape[1]):
        c = nodes[:, d]
        mask |= torch.isclose(c, c.min()) | torch.isclose(c, c.max())
    return mask


@pytest.mark.parametrize("etype", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_patch_test(etype, precision):
    """tests/test_models.py:54-71 on the GPU path."""
    if etype == "Quad1":
        model = Planar(*(t.cuda() for t in rect_quad(4, 4)), IsotropicElasticityPlaneStress(1000.0, 0.3))
        G = torch.tensor([[1.0e-3, 2.0e-4], [3.0e-4, -1.0e-3]], dtype=DT, device="cuda")
    else:
        mesh = {"Tetra1": cube_tetra6(2), "Hexa1": cube_hexa(3, 3, 3),
                "Hexa2": linear_to_quadratic(*cube_hexa(3, 3, 3))}[etype]
        model = Solid(*(t.cuda() for t in mesh), IsotropicElasticity3D(1000.0, 0.3))
        G = torch.tensor([[1.0e-3, 2.0e-4, -1.0e-4], [3.0e-4, -1.0e-3, 5.0e-5],
                          [1.0e-4, 2.0e-4, 7.0e-4]], dtype=DT, device="cuda")
    assert model.etype.__name__ == etype
    u_exact = model.nodes @ G.T
    model.constraints = _boundary(model.nodes)[:, None].repeat(1, model.n_dof_per_node)
    model.displacements = u_exact
    u, _, sigma, _, _ = model.solve(precision=precision, stol=1e-13)
    assert torch.allclose(u, u_exact, atol=1e-12)
    assert torch.allclose(sigma, sigma[0].expand_as(sigma), atol=1e-10)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_planar_cantilever_config1(precision):
    """Plane-stress Quad1 cantilever on the 64 x 32 mesh over [0,2]x[0,1]
    (2145 nodes, 4290 DOFs), E=1000, nu=0.3, clamped at x=0, total tip load -1:
    u, f (reactions) and sigma against the oracle's direct solve."""
    nodes, el = rect_quad(65, 33, 2.0, 1.0)
    model = Planar(nodes.cuda(), el.cuda(), IsotropicElasticityPlaneStress(1000.0, 0.3))
    model.constraints[model.nodes[:, 0] == 0.0, :] = True
    tip = model.nodes[:, 0] == 2.0
    assert int(tip.sum()) == 33
    model.forces[tip, 1] = -1.0 / 33
    u, f, sigma, Fg, _ = model.solve(precision=precision, stol=1e-12)
    o = O.OracleModel(nodes, el, O.Quad1, "planar",
                      O.plane_stress_C(torch.full((len(el),), 1000.0, dtype=DT), torch.tensor(0.3)),
                      thickness=torch.ones(len(el), dtype=DT))
    uo, fo, so, go = o.solve(model.constraints.cpu(), model.forces.cpu(), torch.zeros(len(nodes), 2, dtype=DT))
    assert rel(u, uo) < 1e-9
    assert rel(f, fo) < 1e-9
    assert rel(sigma, so) < 1e-9
    assert rel(Fg, go) < 1e-12
    # reactions balance the applied load
    assert abs(float(f[:, 1].sum())) < 1e-9


def _simp_model(n, seed=0):
    nodes, el = cube_hexa(n + 1, n + 1, n + 1)
    rho = 0.3 + 0.7 * torch.rand(len(el), generator=torch.Generator().manual_seed(seed), dtype=DT)
    return nodes, el, rho


@pytest.mark.parametrize("precision", PRECISIONS)
def test_simp_compliance_gradient(precision):
    """Config 3 reduced to 8^3: rho ~ U(0.3, 1), E = 1000 rho^3, clamped x=0,
    unit -z load on the x=1 face; dC/drho vs the oracle's implicit-function
    gradient, the closed form and central finite differences."""
    n = 8
    nodes, el, rho0 = _simp_model(n)
    rho = rho0.clone().cuda().requires_grad_(True)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0 * rho**3, 0.3))
    model.constraints[model.nodes[:, 0] == 0.0] = True
    right = model.nodes[:, 0] == 1.0
    model.forces[right, 2] = -1.0 / int(right.sum())
    u, f, _, _, _ = model.solve(differentiable_parameters=rho, precision=precision, stol=1e-12)
    comp = torch.inner(model.forces.ravel(), u.ravel())
    (g,) = torch.autograd.grad(comp, rho)

    r = rho0.clone().requires_grad_(True)
    o = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(1000.0 * r**3, torch.tensor(0.3)))
    F = model.forces.cpu()
    con = model.constraints.cpu()
    uo, _, _, _ = o.solve(con, F, torch.zeros_like(F), params=(r,))
    (go,) = torch.autograd.grad(torch.inner(F.ravel(), uo.ravel()), r)
    assert rel(u, uo) < 1e-9
    assert rel(g, go) < 1e-8
    unit = O.OracleModel(nodes, el, O.Hexa1, "solid",
                         O.iso3d_C(torch.ones(len(el), dtype=DT), torch.tensor(0.3)), pattern=o.pattern)
    closed = O.simp_compliance_sensitivity(unit, uo.detach(), rho0, 1000.0)
    assert rel(g, closed) < 1e-8
    # central differences (h = 1e-6, FP64) on a few elements
    h = 1e-6
    for e in [0, 77, 301, len(el) - 1]:
        vals = []
        for s in (1.0, -1.0):
            rr = rho0.clone()
            rr[e] += s * h
            m2 = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0 * rr.cuda() ** 3, 0.3))
            m2.constraints = model.constraints
            m2.forces = model.forces
            u2, _, _, _, _ = m2.solve(precision="fp64", stol=1e-13)
            vals.append(float(torch.inner(model.forces.ravel(), u2.ravel())))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert math.isclose(float(g[e]), fd, rel_tol=1e-5), (e, float(g[e]), fd)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_displacement_driven_cube_and_force_gradient(precision):
    """Config 2 reduced to 10^3: z=0 clamped, u_z=-0.01 on z=1; the published
    benchmark's gradient d(sum u)/d(forces) (benchmarks/cubes.py:27-39)."""
    n = 10
    nodes, el = cube_hexa(n + 1, n + 1, n + 1)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    X = model.nodes
    model.constraints[X[:, 2] == 0.0, :] = True
    model.constraints[X[:, 2] == 1.0, 2] = True
    model.displacements[X[:, 2] == 1.0, 2] = -0.01
    model.forces = torch.zeros_like(X, requires_grad=True)
    u, f, sigma, _, _ = model.solve(differentiable_parameters=model.forces, precision=precision,
                                    stol=1e-12)
    (g,) = torch.autograd.grad(u.sum(), model.forces)
    o = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1000.0, dtype=DT), torch.tensor(0.3)))
    Fo = torch.zeros(len(nodes), 3, dtype=DT, requires_grad=True)
    uo, fo, so, _ = o.solve(model.constraints.cpu(), Fo, model.displacements.cpu(), params=(Fo,))
    (go,) = torch.autograd.grad(uo.sum(), Fo)
    assert rel(u, uo) < 1e-9
    assert rel(f, fo) < 1e-9
    assert rel(sigma, so) < 1e-9
    assert rel(g, go) < 1e-8


@pytest.mark.parametrize("precision", PRECISIONS)
def test_thermal_tetra_mean_temperature_gradient(precision):
    """Config 4 reduced to 8^3 cells (6 tets each, jittered), kappa =
    exp(N(0, 0.5^2)); T=0 at x=0, T=1 at x=1; d(mean T)/d(kappa)."""
    n = 8
    nodes, el = cube_tetra6(n, jitter=0.1, seed=0)
    k0 = torch.exp(0.5 * torch.randn(len(el), generator=torch.Generator().manual_seed(1), dtype=DT))
    kap = k0.clone().cuda().requires_grad_(True)
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(kap))
    X = model.nodes
    model.constraints[X[:, 0] == 0.0] = True
    model.constraints[X[:, 0] == 1.0] = True
    model.temperatures[X[:, 0] == 1.0] = 1.0
    T, f, q, gT, _ = model.solve(differentiable_parameters=kap, precision=precision, stol=1e-12)
    (g,) = torch.autograd.grad(T.mean(), kap)
    kr = k0.clone().requires_grad_(True)
    o = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(kr))
    To, fo, qo, go_, = o.solve(model.constraints.cpu(), torch.zeros(len(nodes), 1, dtype=DT),
                               model.temperatures.cpu(), params=(kr,))
    (gko,) = torch.autograd.grad(To.mean(), kr)
    assert rel(T, To) < 1e-9
    assert rel(f, fo) < 1e-9
    assert rel(q, qo.squeeze(-2)) < 1e-9
    assert rel(g, gko) < 1e-8


@pytest.mark.parametrize("precision", PRECISIONS)
def test_hexa2_traction_and_modulus_gradient(precision):
    """Config 5 reduced to 3^3 Hexa2: clamped z=0, random traction on z=1
    through integrate_surface_load; compliance gradient w.r.t. per-element E."""
    nodes, el = O.linear_to_quadratic(*O.cube_hexa(4, 4, 4))
    E0 = 1000.0 * torch.ones(len(el), dtype=DT)
    Eg = E0.clone().cuda().requires_grad_(True)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(Eg, 0.3))
    X = model.nodes
    model.constraints[X[:, 2] == 0.0] = True
    top = X[:, 2] == 1.0
    trac = torch.randn(9, 3, generator=torch.Generator().manual_seed(0), dtype=DT)
    load = model.integrate_surface_load(top, trac.cuda())
    load = load / load.sum(0).norm()
    model.forces = load
    u, f, _, _, _ = model.solve(differentiable_parameters=Eg, precision=precision, stol=1e-12)
    (g,) = torch.autograd.grad(torch.inner(load.ravel(), u.ravel()), Eg)
    Er = E0.clone().requires_grad_(True)
    o = O.OracleModel(nodes, el, O.Hexa2, "solid", O.iso3d_C(Er, torch.tensor(0.3)))
    lo, conn = o.integrate_surface_load(top.cpu(), trac)
    assert torch.allclose(load.cpu() * (lo.sum(0).norm()), lo, atol=1e-14)
    lo = lo / lo.sum(0).norm()
    uo, _, _, _ = o.solve(model.constraints.cpu(), lo, torch.zeros_like(lo), params=(Er,))
    (go,) = torch.autograd.grad(torch.inner(lo.ravel(), uo.ravel()), Er)
    assert rel(u, uo) < 1e-9
    assert rel(g, go) < 1e-8


def test_increments_scale_like_the_reference():
    """tests/test_gradients.py:27-76: incremental loading reaches the same
    state and gradient as a single step; intermediate states are returned."""
    nodes = torch.tensor([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, 1.0]])
    elements = torch.tensor([[0, 1, 4, 3], [1, 2, 5, 4]])

    def build():
        c = Planar(nodes.cuda(), elements.cuda(), IsotropicElasticityPlaneStress(E=1000.0, nu=0.3))
        c.forces[5, 1] = -1.0
        c.constraints[[0, 3], :] = True
        c.thickness.requires_grad = True
        return c

    c1 = build()
    u1, f1, _, _, _ = c1.solve(differentiable_parameters=c1.thickness)
    g1 = torch.autograd.grad(torch.inner(f1.ravel(), u1.ravel()), c1.thickness)[0]
    c2 = build()
    incs = torch.linspace(0.1, 1.0, 5)
    u2, f2, _, _, _ = c2.solve(increments=incs, return_intermediate=True,
                               differentiable_parameters=c2.thickness)
    assert u2.shape == (5, 6, 2)
    g2 = torch.autograd.grad(torch.inner(f2[-1].ravel(), u2[-1].ravel()), c2.thickness)[0]
    assert torch.allclose(g2, g1, atol=1e-9, rtol=1e-7)
    assert torch.allclose(u2[2], incs[2].item() * u1, atol=1e-12)


def test_negative_jacobian_raises():
    """base.py:311-312 / tests/test_base.py:77-84."""
    nodes, el = rect_quad(3, 3)
    el = el[:, [0, 3, 2, 1]]  # clockwise
    model = Planar(nodes.cuda(), el.cuda(), IsotropicElasticityPlaneStress(1000.0, 0.3))
    model.constraints[0] = True
    with pytest.raises(ValueError, match="Negative Jacobian"):
        model.solve()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_batched_load_cases_hexa2(precision):
    """Config 5 reduced to 3^3 Hexa2 with 5 traction load cases (padded to 8
    interleaved right-hand sides): every case against the oracle, and the
    gradient of the summed compliances against the sum of oracle gradients."""
    nodes, el = O.linear_to_quadratic(*O.cube_hexa(4, 4, 4))
    E0 = 1000.0 * torch.ones(len(el), dtype=DT)
    Eg = E0.clone().cuda().requires_grad_(True)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(Eg, 0.3))
    X = model.nodes
    model.constraints[X[:, 2] == 0.0] = True
    top = X[:, 2] == 1.0
    Er = E0.clone().requires_grad_(True)
    loads, oracle_u = [], []
    go = torch.zeros_like(E0)
    for k in range(5):
        o = O.OracleModel(nodes, el, O.Hexa2, "solid", O.iso3d_C(Er, torch.tensor(0.3)))
        trac = torch.randn(9, 3, generator=torch.Generator().manual_seed(k), dtype=DT)
        lo, _ = o.integrate_surface_load(top.cpu(), trac)
        lo = lo / lo.sum(0).norm()
        loads.append(lo)
        uo, _, _, _ = o.solve(model.constraints.cpu(), lo, torch.zeros_like(lo), params=(Er,))
        go += torch.autograd.grad(torch.inner(lo.ravel(), uo.ravel()), Er)[0]
        oracle_u.append(uo.detach())
    F = torch.stack(loads).cuda()
    u, f = model.solve_load_cases(F, differentiable_parameters=Eg, precision=precision, stol=1e-12)
    assert u.shape == (5, model.n_nod, 3) and f.shape == u.shape
    for k in range(5):
        assert rel(u[k], oracle_u[k]) < 1e-9
    C = (F * u).sum()
    (g,) = torch.autograd.grad(C, Eg)
    assert rel(g, go) < 1e-8


@pytest.mark.parametrize("precision", PRECISIONS)
def test_sixteen_load_cases_match_single_solves(precision):
    """16 right-hand sides (the batch limit) on a Hexa1 cube: the batched solve
    agrees with 16 separate FEM.solve() calls, and the force gradient of a
    weighted sum of displacements with the separate adjoints."""
    nodes, el = cube_hexa(6, 6, 6)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    X = model.nodes
    model.constraints[X[:, 0] == 0.0] = True
    gen = torch.Generator().manual_seed(3)
    F = torch.randn(16, model.n_nod, 3, generator=gen, dtype=DT).cuda()
    F.requires_grad_(True)
    W = torch.randn(16, model.n_nod, 3, generator=gen, dtype=DT).cuda()
    u, _ = model.solve_load_cases(F, differentiable_parameters=F, precision=precision, stol=1e-12)
    (gF,) = torch.autograd.grad((W * u).sum(), F)
    for k in (0, 7, 15):
        Fk = F[k].detach().clone().requires_grad_(True)
        model.forces = Fk
        uk, _, _, _, _ = model.solve(differentiable_parameters=Fk, precision="fp64", stol=1e-12)
        (gk,) = torch.autograd.grad((W[k] * uk).sum(), Fk)
        assert rel(u[k], uk) < 1e-9
        assert rel(gF[k], gk) < 1e-8


def test_wide_grid_solves_agree():
    """A 28^3 Hexa1 cube (24,389 nodes): the FP64 SpMV runs one node per team
    on more blocks than the partial-sum floor of 2048, and the mixed solver on
    a capped grid; both reach the tolerance and agree."""
    nodes, el = cube_hexa(29, 29, 29)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    X = model.nodes
    model.constraints[X[:, 2] == 0.0, :] = True
    model.constraints[X[:, 2] == 1.0, 2] = True
    model.displacements[X[:, 2] == 1.0, 2] = -0.01
    out = {}
    for precision in PRECISIONS:
        u, f, _, _, _ = model.solve(precision=precision, stol=1e-10)
        info = model.last_solve_info["forward"]["pcg"][-1]
        assert info["relres"][0] <= 1e-10
        out[precision] = u
    assert rel(out["mixed"], out["fp64"]) < 1e-8
