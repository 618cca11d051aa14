"""The two-level preconditioner (preconditioner="two_level": block-Jacobi plus
an additive coarse correction on node aggregates with their rigid-body
modes, in the place of the reference's AMG preconditioner,
sparse.py:422-442). Results against the oracle's direct solves for every
element type and both precisions, gradients against the SIMP closed form,
batched load cases, aggregates whose DOFs are all constrained, the solver
node order of Hexa2 meshes, Newton increments and implicit time steps,
the dense inverse of the coarse operator and the iteration counts it saves.
Void density designs are covered with the other preconditioners in
test_gpu_contrast.py."""

import pytest
import torch

import oracle as O
from torchfem import Planar, Solid, SolidHeat
from torchfem.materials import (IsotropicConductivity3D, IsotropicElasticity3D,
                                IsotropicElasticityPlaneStress)
from torchfem.mesh import cube_hexa, cube_tetra6, rect_quad

pytestmark = pytest.mark.gpu
DT = torch.float64
PRECISIONS = ["fp64", "mixed"]


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def _hexa(n, quadratic=False, seed=0):
    nodes, el = O.cube_hexa(n + 1, n + 1, n + 1)
    if quadratic:
        nodes, el = O.linear_to_quadratic(nodes, el)
    inner = ((nodes > 0) & (nodes < 1)).all(dim=1)
    nodes = nodes + 0.02 * torch.sin(5.0 * nodes[:, [1, 2, 0]]) * inner[:, None]
    g = torch.Generator().manual_seed(seed)
    E = 500.0 + 1000.0 * torch.rand(len(el), generator=g, dtype=DT)
    return nodes, el, E


def _solid_case(n, quadratic=False):
    nodes, el, E = _hexa(n, quadratic)
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E.cuda(), 0.3))
    X = m.nodes
    m.constraints[X[:, 0] == 0.0] = True
    face = X[:, 0] == 1.0
    m.forces[face, 2] = -1.0 / int(face.sum())
    m.forces[face, 1] = 0.3 / int(face.sum())
    o = O.OracleModel(nodes, el, O.Hexa2 if quadratic else O.Hexa1, "solid",
                      O.iso3d_C(E, torch.tensor(0.3)))
    return m, o


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("quadratic", [False, True])
def test_solid_matches_oracle(precision, quadratic):
    """Distorted Hexa1 (10^3) and Hexa2 (5^3) cubes with random moduli: u
    and reactions against the oracle."""
    m, o = _solid_case(5 if quadratic else 10, quadratic)
    u, f, _, _, _ = m.solve(precision=precision, preconditioner="two_level", stol=1e-11)
    uo, fo, _, _ = o.solve(m.constraints.cpu(), m.forces.cpu(), m.displacements.cpu())
    assert rel(u, uo) < 1e-9
    assert rel(f, fo) < 1e-8


def test_solver_node_order():
    """A 20^3 Hexa2 cube, whose mesh order does not fit 16-bit neighbour
    offsets, is solved in the locality node order: the coarse space is built
    on that pattern; against block-Jacobi in the same order."""
    nodes, el, E = _hexa(20, quadratic=True)
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E.cuda(), 0.3))
    assert m._solver_view() is not None
    m.constraints[m.nodes[:, 0] == 0.0] = True
    face = m.nodes[:, 0] == 1.0
    m.forces[face, 2] = -1.0 / int(face.sum())
    u, _, _, _, _ = m.solve(precision="mixed", preconditioner="two_level", stol=1e-10)
    ub, _, _, _, _ = m.solve(precision="fp64", preconditioner="block_jacobi", stol=1e-12)
    assert rel(u, ub) < 1e-8


@pytest.mark.parametrize("precision", PRECISIONS)
def test_heat_and_planar_match_oracle(precision):
    """A scalar field (Tetra1 conduction: one constant mode per aggregate)
    and a plane-stress Quad1 cantilever (two translations and a rotation)."""
    nodes, el = O.cube_tetra6(7, jitter=0.1, seed=2)
    g = torch.Generator().manual_seed(3)
    k = torch.exp(0.4 * torch.randn(len(el), generator=g, dtype=DT))
    h = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(k.cuda()))
    h.constraints[h.nodes[:, 0] == 0.0] = True
    h.heat_flux = torch.full_like(h.heat_flux, 1e-2)
    T, _, _, _, _ = h.solve(precision=precision, preconditioner="two_level", stol=1e-11)
    oh = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(k))
    To, _, _, _ = oh.solve(h.constraints.cpu(), h.heat_flux.cpu(), h.temperatures.cpu())
    assert rel(T, To) < 1e-9

    nodes, el = O.rect_quad(41, 21, 2.0, 1.0)
    P = Planar(nodes.cuda(), el.cuda(), IsotropicElasticityPlaneStress(1000.0, 0.3))
    P.constraints[P.nodes[:, 0] == 0.0] = True
    tip = P.nodes[:, 0] == 2.0
    P.forces[tip, 1] = -1.0 / int(tip.sum())
    u, _, _, _, _ = P.solve(precision=precision, preconditioner="two_level", stol=1e-11)
    op = O.OracleModel(nodes, el, O.Quad1, "planar",
                       O.plane_stress_C(torch.full((len(el),), 1000.0, dtype=DT),
                                        torch.tensor(0.3)),
                       thickness=torch.ones(len(el), dtype=DT))
    uo, _, _, _ = op.solve(P.constraints.cpu(), P.forces.cpu(), P.displacements.cpu())
    assert rel(u, uo) < 1e-9


@pytest.mark.parametrize("precision", PRECISIONS)
def test_simp_gradient_matches_closed_form(precision):
    """dC/drho through the adjoint solved with the two-level preconditioner,
    against -3 rho^2 u_e^T k_e u_e from the oracle's solve (8^3)."""
    n = 8
    nodes, el = O.cube_hexa(n + 1, n + 1, n + 1)
    rho0 = 0.3 + 0.7 * torch.rand(len(el), generator=torch.Generator().manual_seed(0), dtype=DT)
    rho = rho0.clone().cuda().requires_grad_(True)
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0 * rho**3, 0.3))
    m.constraints[m.nodes[:, 0] == 0.0] = True
    right = m.nodes[:, 0] == 1.0
    m.forces[right, 2] = -1.0 / int(right.sum())
    u, _, _, _, _ = m.solve(differentiable_parameters=rho, precision=precision,
                            preconditioner="two_level", stol=1e-12)
    (g,) = torch.autograd.grad(torch.inner(m.forces.ravel(), u.ravel()), rho)
    o = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(1000.0 * rho0**3, torch.tensor(0.3)))
    uo, _, _, _ = o.solve(m.constraints.cpu(), m.forces.cpu(), torch.zeros(len(nodes), 3, dtype=DT))
    unit = O.OracleModel(nodes, el, O.Hexa1, "solid",
                         O.iso3d_C(torch.ones(len(el), dtype=DT), torch.tensor(0.3)),
                         pattern=o.pattern)
    closed = O.simp_compliance_sensitivity(unit, uo, rho0, 1000.0)
    assert rel(u, uo) < 1e-9
    assert rel(g, closed) < 1e-8


@pytest.mark.parametrize("precision", PRECISIONS)
def test_load_cases_and_constrained_aggregates(precision):
    """Four load cases in one batched solve and adjoint, on a cube clamped
    over its lower third (whole aggregates have every DOF constrained, so
    their coarse modes vanish): against one two-level solve per case and
    against block-Jacobi."""
    nodes, el = cube_hexa(13, 13, 13)
    E = torch.full((len(el),), 1000.0, dtype=DT, device="cuda", requires_grad=True)
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E, 0.3))
    X = m.nodes
    m.constraints[X[:, 2] <= 1.0 / 3.0 + 1e-12] = True
    top = X[:, 2] == 1.0
    g = torch.Generator().manual_seed(7)
    F = torch.zeros(4, m.n_nod, 3, dtype=DT)
    F[:, top.cpu()] = torch.randn(4, int(top.sum()), 3, generator=g, dtype=DT)
    F = F.cuda()
    kw = dict(differentiable_parameters=E, precision=precision, stol=1e-11)
    U = m.solve_load_cases(F, preconditioner="two_level", **kw)[0]
    (gE,) = torch.autograd.grad(U.sum(), E)
    m.material = IsotropicElasticity3D(E, 0.3)  # a fresh graph from E
    Ub = m.solve_load_cases(F, preconditioner="block_jacobi", **kw)[0]
    (gEb,) = torch.autograd.grad(Ub.sum(), E)
    assert rel(U, Ub) < 1e-9
    assert rel(gE, gEb) < 1e-8
    m.material = IsotropicElasticity3D(E.detach(), 0.3)
    for c in (0, 3):
        m.forces = F[c]
        u, _, _, _, _ = m.solve(precision=precision, preconditioner="two_level", stol=1e-11)
        assert rel(U[c], u) < 1e-9


@pytest.mark.parametrize("n", [1, 7, 64, 100, 257])
def test_spd_inverse(n):
    """tfm_spd_inverse (block Gauss-Jordan, 64-column pivot blocks) against
    torch.linalg.inv on SPD matrices with condition number 1e6, sizes below,
    at and across block boundaries; an indefinite matrix reports the first
    pivot that is not positive."""
    from torchfem._coarse import spd_inverse
    g = torch.Generator().manual_seed(n)
    Q, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=DT))
    A = ((Q * torch.logspace(-6, 0, n, dtype=DT)) @ Q.T).cuda()
    A = 0.5 * (A + A.T)
    X = spd_inverse(A.clone())
    ref = torch.linalg.inv(A)
    assert X is not None
    # the inverse's error is bounded by cond(A) * eps relative to its norm
    assert rel(X, ref) < 1e-8
    assert float((X @ A - torch.eye(n, dtype=DT, device="cuda")).abs().max()) < 1e-8
    if n > 1:
        B = torch.eye(n, dtype=DT, device="cuda")
        B[n - 1, n - 1] = -1.0
        assert spd_inverse(B) is None


def test_newton_increments_and_time_stepping():
    """Inside Newton increments (a Neo-Hookean cantilever with nlgeom, 2,499
    DOFs in four aggregates: every tangent gets its own coarse operator and
    inverse) and inside implicit time stepping (a 33 x 33 heated plate,
    (M + dt/2 K) solves): the results equal block-Jacobi's."""
    from test_gpu_hyper import LBD, MU, psi
    from torchfem import PlanarHeat
    from torchfem.materials import Hyperelastic3D, IsotropicConductivity2D
    nodes, el = cube_hexa(17, 7, 7)
    nodes = nodes * torch.tensor([2.0, 0.5, 0.5], dtype=DT)
    params = torch.tensor([MU, LBD], dtype=DT)
    out = {}
    for pc in ("block_jacobi", "two_level"):
        beam = Solid(nodes.cuda(), el.cuda(), Hyperelastic3D(psi, params))
        X = beam.nodes
        beam.constraints[X[:, 0] == 0.0] = True
        tip = X[:, 0] == 2.0
        beam.forces[tip, 2] = -0.36 / int(tip.sum())
        beam.forces[tip, 1] = 0.108 / int(tip.sum())
        u, f, _, _, _ = beam.solve(increments=torch.tensor([0.0, 0.5, 1.0]), nlgeom=True,
                                   rtol=1e-12, atol=1e-9, stol=1e-12, max_iter=40,
                                   precision="fp64", preconditioner=pc)
        out[pc] = (u, f)
    assert float(out["two_level"][0].abs().max()) > 1e-3
    assert rel(out["two_level"][0], out["block_jacobi"][0]) < 1e-9
    assert rel(out["two_level"][1], out["block_jacobi"][1]) < 1e-8

    nodes, el = rect_quad(33, 33, 1.0, 1.0)
    T = {}
    for pc in ("block_jacobi", "two_level"):
        plate = PlanarHeat(nodes.cuda(), el.cuda(), IsotropicConductivity2D(kappa=400.0, rho=1.0e5))
        west = torch.isclose(plate.nodes[:, 0], plate.nodes[:, 0].min())
        east = torch.isclose(plate.nodes[:, 0], plate.nodes[:, 0].max())
        plate.constraints[west | east] = True
        plate.temperatures[west, 0] = 5.0
        plate.temperatures[east, 0] = 20.0
        T[pc] = plate.time_integration(torch.tensor([0.0, 0.7, 2.0]), delta_t=0.25, rtol=1e-12,
                                       atol=1e-12, stol=1e-12, precision="mixed",
                                       preconditioner=pc)[0]
    assert float((T["two_level"][-1] - 5.0).abs().max()) > 1.0
    assert rel(T["two_level"], T["block_jacobi"]) < 1e-9


@pytest.mark.parametrize("precision", PRECISIONS)
def test_every_element_case(precision):
    """Every element type and material layout of the kernel tests (plane
    stress with thickness, heat in 2-D and 3-D, Tetra1 / Hexa1 / Hexa2
    elasticity, isotropic and general tangents), clamped at the lowest x and
    loaded at the highest: the two-level solve against the oracle's direct
    solve."""
    from test_gpu_kernels import CASES, _case
    for name, general in CASES:
        m, o = _case(name, general)
        X = m.nodes
        m.constraints[X[:, 0] == X[:, 0].min()] = True
        far = X[:, 0] == X[:, 0].max()
        heat = name.endswith("heat")
        if heat:
            m.heat_flux[far] = 1.0 / int(far.sum())
            load, pres = m.heat_flux, m.temperatures
        else:
            m.forces[far, -1] = -1.0 / int(far.sum())
            m.forces[far, 0] = 0.4 / int(far.sum())
            load, pres = m.forces, m.displacements
        u, f, _, _, _ = m.solve(precision=precision, preconditioner="two_level", stol=1e-11)
        uo, fo, _, _ = o.solve(m.constraints.cpu(), load.cpu(), pres.cpu())
        assert rel(u, uo) < 1e-9, (name, general)
        assert rel(f, fo) < 1e-8, (name, general)


@pytest.mark.parametrize("kind", ["solid", "planar", "heat"])
def test_coarse_operator_on_a_free_body(kind):
    """k_coarse_operator (A_c = P^T K P) on an unconstrained body, whose K
    has the rigid motions (the constants for heat) as its null space: A_c is
    symmetric, annihilates the coefficients of every global rigid motion
    (translation t + rotation w about the origin is t + w x c_I plus the
    rotation w in the modes of aggregate I, whose centroid is c_I), and has
    exactly that many zero eigenvalues (6, 3 or 1). A rotation mode with a
    wrong sign or axis, or a dropped block, fails one of these."""
    import ctypes as C

    from torchfem import _lib as L
    from torchfem._coarse import CoarseSpace
    from torchfem.base import _System
    from torchfem.materials import IsotropicConductivity3D
    if kind == "solid":
        nodes, el, E = _hexa(7)
        m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E.cuda(), 0.3))
        nz = 6
    elif kind == "planar":
        nodes, el = O.rect_quad(25, 9, 3.0, 1.0)
        m = Planar(nodes.cuda(), el.cuda(), IsotropicElasticityPlaneStress(1000.0, 0.3))
        nz = 3
    else:
        nodes, el = O.cube_tetra6(6, jitter=0.1, seed=2)
        m = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(1.0))
        nz = 1
    kd, p0, p1, sc = m._material_params()
    S = _System(m, kd, p0, p1, sc, "fp64", renumber=False)
    pat = S._pat
    cs = CoarseSpace(pat, size=24 * cs_modes(kind))
    assert cs.n_agg >= 4
    Ac = torch.empty(cs.n_c, cs.n_c, dtype=DT, device="cuda")
    L.check(L.lib().tfm_coarse_operator(pat.stream(), cs.h, C.byref(S.sys), L.ptr(Ac)))
    Ac = Ac.cpu()
    scale = float(Ac.abs().max())
    assert float((Ac - Ac.T).abs().max()) < 1e-12 * scale
    cen = cs._arrays["cen"].cpu()
    g = torch.Generator().manual_seed(5)
    if kind == "heat":
        c = torch.ones(cs.n_c, dtype=DT)
    elif kind == "solid":
        t, w = torch.randn(3, generator=g, dtype=DT), torch.randn(3, generator=g, dtype=DT)
        c = torch.cat([t + torch.linalg.cross(w.expand_as(cen), cen), w.expand_as(cen)], 1).reshape(-1)
    else:
        t, w = torch.randn(2, generator=g, dtype=DT), float(torch.randn(1, generator=g, dtype=DT))
        rot = torch.stack([-w * cen[:, 1], w * cen[:, 0]], 1)
        c = torch.cat([t + rot, torch.full((cs.n_agg, 1), w, dtype=DT)], 1).reshape(-1)
    assert float((Ac @ c).norm()) < 1e-10 * scale * float(c.norm())
    ev = torch.linalg.eigvalsh(Ac)
    assert bool((ev[:nz].abs() < 1e-9 * scale).all())
    assert float(ev[nz]) > 1e-6 * scale


def cs_modes(kind):
    return {"solid": 6, "planar": 3, "heat": 1}[kind]


def test_coarse_inverse_fallback(monkeypatch):
    """When the Gauss-Jordan inverse reports a pivot that is not positive,
    the coarse inverse comes from the eigendecomposition instead (forced
    here on the host by a stand-in that reports failure): the solve still
    matches the oracle."""
    import torchfem._coarse as coarse
    monkeypatch.setattr(coarse, "spd_inverse", lambda A, stream=None: None)
    m, o = _solid_case(6)
    u, f, _, _, _ = m.solve(precision="fp64", preconditioner="two_level", stol=1e-11)
    uo, fo, _, _ = o.solve(m.constraints.cpu(), m.forces.cpu(), m.displacements.cpu())
    assert rel(u, uo) < 1e-9
    assert rel(f, fo) < 1e-8


def test_coarse_space_failure_falls_back(monkeypatch):
    """A coarse space that cannot be built (forced on the host by a stand-in
    that raises) leaves the solve to block-Jacobi, marked in the solve info,
    with the oracle's result."""
    import torchfem.base as base

    def refuse(pat, size=None):
        raise RuntimeError("no coarse space")
    monkeypatch.setattr(base, "CoarseSpace", refuse)
    m, o = _solid_case(6)
    u, f, _, _, _ = m.solve(precision="mixed", preconditioner="two_level", stol=1e-11)
    assert m.last_solve_info["forward"]["pcg"][0].get("two_level_fallback")
    uo, fo, _, _ = o.solve(m.constraints.cpu(), m.forces.cpu(), m.displacements.cpu())
    assert rel(u, uo) < 1e-9


def test_iterations_drop():
    """24^3 Hexa1 compression (the bench's load on a smaller cube): the
    coarse correction at least halves the PCG iterations of block-Jacobi,
    forward and adjoint."""
    nodes, el = cube_hexa(25, 25, 25)
    E = torch.full((len(el),), 1000.0, dtype=DT, device="cuda", requires_grad=True)
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E, 0.3))
    X = m.nodes
    m.constraints[X[:, 2] == 0.0, :] = True
    m.constraints[X[:, 2] == 1.0, 2] = True
    m.displacements[X[:, 2] == 1.0, 2] = -0.01
    its = {}
    for pc in ("block_jacobi", "two_level"):
        m.material = IsotropicElasticity3D(E, 0.3)
        u, _, _, _, _ = m.solve(differentiable_parameters=E, precision="fp64", preconditioner=pc,
                                stol=1e-8)
        torch.autograd.grad(u.sum(), E)
        info = m.last_solve_info
        its[pc] = (info["forward"]["pcg"][0]["iterations"][0], info["adjoint"]["iterations"][0])
    assert its["two_level"][0] * 2 <= its["block_jacobi"][0], its
    assert its["two_level"][1] * 2 <= its["block_jacobi"][1], its
