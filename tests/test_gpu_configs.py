"""The benchmark configurations at their full sizes on the GPU. An oracle solve
is out of reach there, so these tests check what holds at any size:
homogeneity of the solution in the moduli, the closed-form SIMP sensitivity,
central differences of a variational objective, mixed-precision against FP64
results, and batched against single solves. The same formulas are pinned
against the CPU oracle at reduced sizes in test_gpu_solve.py."""

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


def rel(a, b):
    a = a.detach().double()
    b = b.detach().double()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def test_config2_hexa1_64_mixed_matches_fp64_and_is_homogeneous():
    """Hexa1 64^3, z=0 clamped, u_z=-0.01 on z=1 (823,875 DOFs): the mixed
    result agrees with the FP64 one to 1e-7 in u and 1e-6 in d(sum u)/dE; with
    displacement loading u does not change when every E_e is scaled, so
    sum_e E_e d(sum u)/dE_e = 0."""
    nodes, el = cube_hexa(65, 65, 65)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    assert model.n_dofs == 823875
    X = model.nodes
    model.constraints[X[:, 2] == 0.0, :] = True
    model.constraints[X[:, 2] == 1.0, 2] = True
    model.displacements[X[:, 2] == 1.0, 2] = -0.01
    out = {}
    for precision, stol in (("fp64", 1e-11), ("mixed", 1e-8)):
        E = torch.full((model.n_elem,), 1000.0, dtype=DT, device="cuda", requires_grad=True)
        model.material = IsotropicElasticity3D(E, 0.3)
        u, _, _, _, _ = model.solve(differentiable_parameters=E, precision=precision, stol=stol)
        assert model.last_solve_info["forward"]["pcg"][-1]["relres"][0] <= stol
        (g,) = torch.autograd.grad(u.sum(), E)
        out[precision] = (u.detach(), g)
    assert rel(out["mixed"][0], out["fp64"][0]) < 1e-7
    assert rel(out["mixed"][1], out["fp64"][1]) < 1e-6
    Eg = 1000.0 * out["fp64"][1]
    assert abs(float(Eg.sum())) < 1e-7 * float(Eg.abs().sum())


def test_config3_simp48_closed_form_and_central_differences():
    """SIMP on Hexa1 48^3: rho ~ U(0.3, 1) (seed 0), E = 1000 rho^3, clamped at
    x=0, total -z load 1 split equally over the 2401 nodes of x=1. dC/drho by
    the adjoint against the closed form -3 rho^2 E0 u_e^T k0_e u_e
    (examples/optimization/solid/topology.ipynb), the degree -1 homogeneity of
    C in E, and central differences (h = 1e-6, FP64) on 32 elements."""
    n = 48
    nodes, el = cube_hexa(n + 1, n + 1, n + 1)
    rho0 = 0.3 + 0.7 * torch.rand(len(el), generator=torch.Generator().manual_seed(0), dtype=DT)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    X = model.nodes
    model.constraints[X[:, 0] == 0.0, :] = True
    face = X[:, 0] == 1.0
    assert int(face.sum()) == 2401
    model.forces[face, 2] = -1.0 / 2401
    F = model.forces.clone()

    def compliance(rho, **kw):
        model.material = IsotropicElasticity3D(1000.0 * rho**3, 0.3)
        u, f, _, _, _ = model.solve(precision="fp64", stol=1e-11, **kw)
        return torch.inner(F.ravel(), u.ravel()), u, f

    def stationary_compliance(rho):
        # 2 F.u - u.K u equals F.u at the solution, and its error is second
        # order in the solver's residual (F.u alone is first order), so its
        # differences resolve the h = 1e-6 perturbation
        C, u, f = compliance(rho)
        return float(2.0 * C - torch.inner(u.ravel(), f.ravel()))

    rho = rho0.clone().cuda().requires_grad_(True)
    C, u, _ = compliance(rho, differentiable_parameters=rho)
    (g,) = torch.autograd.grad(C, rho)
    C = float(C.detach())

    unit = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1.0, 0.3))
    k0 = unit.k0()
    ue = u.detach()[model.elements].reshape(len(el), -1)
    closed = -3.0 * rho0.cuda() ** 2 * 1000.0 * torch.einsum("ei,eij,ej->e", ue, k0, ue)
    assert rel(g, closed) < 1e-7
    # C(c E) = C(E) / c  =>  sum_e rho_e / 3 dC/drho_e = -C
    assert abs(float((rho0.cuda() * g).sum()) / 3.0 + C) < 1e-8 * abs(C)

    idx = torch.randperm(len(el), generator=torch.Generator().manual_seed(0))[:32]
    h = 1e-6
    for e in idx.tolist():
        cs = []
        for sgn in (1.0, -1.0):
            r = rho0.clone()
            r[e] += sgn * h
            cs.append(stationary_compliance(r.cuda()))
        fd = (cs[0] - cs[1]) / (2.0 * h)
        assert abs(fd - float(g[e])) < 1e-3 * abs(float(g[e])), (e, fd, float(g[e]))


def test_config4_thermal_tetra80_mixed_matches_fp64_and_is_homogeneous():
    """Tetra1 Kuhn split of 80^3 cells (531,441 DOFs, 3,072,000 tets), interior
    nodes jittered by U(-0.1h, 0.1h) (seed 0), kappa_e = exp(N(0, 0.5^2))
    (seed 1), T=0 at x=0 and T=1 at x=1. T and d(mean T)/d(kappa) of the mixed
    path against FP64; T does not change when every kappa_e is scaled, so
    sum_e kappa_e dJ/dkappa_e = 0; 0 <= T <= 1 up to the discretisation's
    overshoot."""
    nodes, el = cube_tetra6(80, jitter=0.1, seed=0)
    assert len(nodes) == 531441 and len(el) == 3072000
    k0 = torch.exp(0.5 * torch.randn(len(el), generator=torch.Generator().manual_seed(1), dtype=DT))
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(1.0))
    X = model.nodes
    model.constraints[X[:, 0] == 0.0] = True
    model.constraints[X[:, 0] == 1.0] = True
    model.temperatures[X[:, 0] == 1.0] = 1.0
    out = {}
    for precision, stol in (("fp64", 1e-11), ("mixed", 1e-8)):
        kap = k0.clone().cuda().requires_grad_(True)
        model.material = IsotropicConductivity3D(kap)
        T, _, _, _, _ = model.solve(differentiable_parameters=kap, precision=precision, stol=stol)
        (g,) = torch.autograd.grad(T.mean(), kap)
        out[precision] = (T.detach(), g)
    T, g = out["fp64"]
    assert float(T.min()) > -0.01 and float(T.max()) < 1.01
    assert rel(out["mixed"][0], T) < 1e-7
    assert rel(out["mixed"][1], g) < 1e-6
    kg = k0.cuda() * g
    assert abs(float(kg.sum())) < 1e-7 * float(kg.abs().sum())


def test_config5_hexa2_32_sixteen_load_cases():
    """Hexa2 32^3 (140,481 nodes, 421,443 DOFs), E=1000, nu=0.3, clamped at z=0;
    16 load cases, each a traction ~N(0,1) per face of z=1 (seed k) integrated
    to consistent nodal loads and scaled to unit total magnitude. One batched
    solve and one batched adjoint give u_k and d(sum_k C_k)/dE: every case
    against its single solve for three cases, the gradient against the
    closed form -sum_k u_ke^T k_e(E=1) u_ke, and the degree -1 homogeneity
    sum_e E_e d(sum C)/dE_e = -sum C."""
    nodes, el = linear_to_quadratic(*cube_hexa(33, 33, 33))
    assert len(nodes) == 140481
    E = torch.full((len(el),), 1000.0, dtype=DT, device="cuda", requires_grad=True)
    model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E, 0.3))
    assert model.n_dofs == 421443
    X = model.nodes
    model.constraints[X[:, 2] == 0.0, :] = True
    top = X[:, 2] == 1.0
    loads = []
    for k in range(16):
        trac = torch.randn(32 * 32, 3, generator=torch.Generator().manual_seed(k), dtype=DT)
        f = model.integrate_surface_load(top, trac.cuda())
        loads.append(f / f.sum(0).norm())
    Fb = torch.stack(loads)
    u, _ = model.solve_load_cases(Fb, differentiable_parameters=E, precision="mixed", stol=1e-8)
    Ck = (Fb * u).sum(dim=(1, 2))
    (g,) = torch.autograd.grad(Ck.sum(), E)
    u = u.detach()

    for k in (0, 7, 15):
        model.forces = Fb[k]
        uk, _, _, _, _ = model.solve(precision="fp64", stol=1e-11)
        assert rel(u[k], uk) < 1e-7

    unit = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1.0, 0.3))
    ke = unit.k0()
    closed = torch.zeros(len(el), dtype=DT, device="cuda")
    for k in range(16):
        ue = u[k][model.elements].reshape(len(el), -1)
        closed -= torch.einsum("ei,eij,ej->e", ue, ke, ue)
    assert rel(g, closed) < 1e-6
    total = float(Ck.detach().sum())
    assert abs(float((1000.0 * g).sum()) + total) < 1e-6 * abs(total)
