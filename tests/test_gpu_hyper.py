"""Finite-strain (hyperelastic) path on the GPU: element integration against
the oracle's restatement of Mechanics.integrate_material, the HIP operators'
adjoints, full incremental Newton solves against the oracle's solve, and the
reference's known answer (tests/test_gradients.py:80-128: the reaction of a
Neo-Hookean cube stretched to three times its length and its derivatives
with respect to the Lame parameters) with gradients from the HIP adjoint.
"""

import math

import pytest
import torch

import oracle as O
from test_oracle import uniaxial_stretch_closed_form
from torchfem import Solid
from torchfem._finite import Kinematics, _ForceOp, _GradOp
from torchfem.materials import Hyperelastic3D, IsotropicElasticity3D

pytestmark = pytest.mark.gpu
DT = torch.float64
MU, LBD = 350.0, 800.0


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def psi(F, params):
    """Compressible Neo-Hookean energy, written as the reference's test writes it."""
    C = F.transpose(-1, -2) @ F
    logJ = 0.5 * torch.logdet(C)
    return params[0] / 2 * (torch.trace(C) - 3.0) - params[0] * logJ + params[1] / 2 * logJ**2


def _mesh(name):
    if name == "Tetra1":
        return O.cube_tetra6(3, jitter=0.15, seed=5)
    nodes, el = O.cube_hexa(3, 4, 3)
    if name == "Hexa2":
        nodes, el = O.linear_to_quadratic(nodes, el)
    inner = ((nodes > 0) & (nodes < 1)).all(dim=1)
    return nodes + 0.04 * torch.sin(4.0 * nodes[:, [1, 2, 0]]) * inner[:, None], el


def _params(n_elem):
    g = torch.Generator().manual_seed(11)
    return torch.stack([MU * (1 + 0.3 * torch.rand(n_elem, generator=g, dtype=DT)),
                        LBD * (1 + 0.3 * torch.rand(n_elem, generator=g, dtype=DT))], dim=1)


@pytest.mark.parametrize("name", ["Hexa1", "Tetra1", "Hexa2"])
@pytest.mark.parametrize("nlgeom", [False, True])
def test_integrate_material_matches_oracle(name, nlgeom):
    """k, f, F and the flux of one finite-strain evaluation from a previous
    deformation gradient, per element parameters, on a distorted mesh."""
    nodes, el = _mesh(name)
    params = _params(len(el))
    fem = Solid(nodes.cuda(), el.cuda(), Hyperelastic3D(psi, params))
    m = O.OracleModel(nodes, el, O.ELEMENTS[name], "solid",
                      O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    g = torch.Generator().manual_seed(3)
    du = 0.05 * torch.randn(len(nodes) * 3, generator=g, dtype=DT)
    ngp = fem.n_int
    grad_prev = torch.eye(3, dtype=DT) + 0.05 * torch.randn(ngp, len(el), 3, 3, generator=g,
                                                              dtype=DT)
    k, f, F, flux, _ = fem.integrate_material(
        None, grad_prev.cuda(), None, torch.zeros(ngp, len(el), 0, device="cuda"), du.cuda(),
        None, 0, nlgeom)
    k_o, f_o, F_o, P_o = O.oracle_integrate_hyper(m, O.neo_hookean_psi, params, grad_prev, du)
    assert rel(F, F_o) < 1e-14
    assert rel(f, f_o) < 1e-12
    assert rel(k, k_o) < 1e-12
    q_o = F_o @ P_o / torch.linalg.det(F_o)[..., None, None] if nlgeom else P_o
    assert rel(flux, q_o) < 1e-12


@pytest.mark.parametrize("name", ["Hexa1", "Tetra1", "Hexa2"])
def test_pointwise_tangent_reduces_to_linear_element_matrices(name):
    """With the same elastic C at every Gauss point the per-point tangent
    kernel gives the linear element matrices (k0)."""
    nodes, el = _mesh(name)
    fem = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    C = fem.material.C.to("cuda", DT)  # [n_elem, 3, 3, 3, 3]
    A = C.expand(fem.n_int, *C.shape).contiguous()
    k = Kinematics(fem).matrices(A)
    assert rel(k, fem.k0()) < 1e-13


@pytest.mark.parametrize("name", ["Hexa1", "Tetra1", "Hexa2"])
def test_operator_adjoints(name):
    """<grad u, Hbar> = <u, grad^T Hbar> and <F_int(P), ubar> = <P, F_int^T ubar>
    for the HIP operators and the adjoints their autograd functions apply."""
    nodes, el = _mesh(name)
    fem = Solid(nodes.cuda(), el.cuda(), Hyperelastic3D(psi, torch.tensor([MU, LBD])))
    kin = Kinematics(fem)
    g = torch.Generator(device="cuda").manual_seed(1)
    u = torch.randn(fem.n_dofs, generator=g, device="cuda", dtype=DT, requires_grad=True)
    Hb = torch.randn(*kin.shape, generator=g, device="cuda", dtype=DT)
    H = _GradOp.apply(kin, u)
    (gu,) = torch.autograd.grad(H, u, Hb)
    lhs, rhs = float((H * Hb).sum().detach()), float((u * gu).sum().detach())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    P = torch.randn(*kin.shape, generator=g, device="cuda", dtype=DT, requires_grad=True)
    ub = torch.randn(fem.n_dofs, generator=g, device="cuda", dtype=DT)
    Fi = _ForceOp.apply(kin, P)
    (gP,) = torch.autograd.grad(Fi, P, ub)
    lhs, rhs = float((Fi * ub).sum().detach()), float((P * gP).sum().detach())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def _stretch_box(n, U, params, etype="Hexa1"):
    if etype == "Tetra1":
        nodes, el = O.cube_tetra6(n)
    else:
        nodes, el = O.cube_hexa(n, n, n)
        if etype == "Hexa2":
            nodes, el = O.linear_to_quadratic(nodes, el)
    box = Solid(nodes.cuda(), el.cuda(), Hyperelastic3D(psi, params))
    X = box.nodes
    right = X[:, 0] == 1.0
    box.constraints[X[:, 0] == 0.0, 0] = True
    box.constraints[right, 0] = True
    box.constraints[X[:, 1] == 0.5, 1] = True
    box.constraints[X[:, 2] == 0.5, 2] = True
    box.displacements[right, 0] = U
    return box, right


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_uniaxial_stretch_reaction_and_gradient(precision):
    """The reference's known answer (tests/test_gradients.py:80-128), same
    mesh, constraints, increments and tolerance: the reaction force of a cube
    stretched to 3x its length and its derivatives with respect to (mu, lbd)."""
    E, nu, U = 1000.0, 0.3, 2.0
    lbd = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    params = torch.tensor([mu, lbd], dtype=DT, requires_grad=True)
    box, right = _stretch_box(3, U, params)
    lam_pts = torch.logspace(0, math.log10(1.0 + U), 8, dtype=DT)
    u, f, _, _, _ = box.solve(increments=(lam_pts - 1.0) / U, nlgeom=True,
                              differentiable_parameters=params, precision=precision)
    reaction = f[right, 0].sum()
    (grad,) = torch.autograd.grad(reaction, params)
    R, dR = uniaxial_stretch_closed_form(mu, lbd, 1.0 + U)
    assert abs(float(reaction) - R) < 1e-7 * R
    assert torch.allclose(grad, torch.tensor(dR, dtype=DT), rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("etype,n", [("Tetra1", 10), ("Hexa2", 5), ("Hexa1", 17)])
def test_uniaxial_stretch_is_homogeneous_at_size(etype, n):
    """Larger meshes of every element type: the stretched block deforms
    homogeneously (u = (lam - 1) x along the stretch, lateral stretch from
    the closed form), and the reaction is the closed-form nominal stress."""
    E, nu = 1000.0, 0.3
    lbd = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    U = 0.8
    box, right = _stretch_box(n, U, torch.tensor([mu, lbd], dtype=DT), etype)
    lam_pts = torch.logspace(0, math.log10(1.0 + U), 6, dtype=DT)
    u, f, flux, grad, _ = box.solve(increments=(lam_pts - 1.0) / U, nlgeom=True)
    lam = 1.0 + U
    R, _ = uniaxial_stretch_closed_form(mu, lbd, lam)
    Jv = 1.0
    for _ in range(100):
        Jv -= (mu * (Jv / lam - 1.0) + lbd * math.log(Jv)) / (mu / lam + lbd / Jv)
    lat = math.sqrt(Jv / lam)
    X = box.nodes
    expect = torch.stack([(lam - 1.0) * X[:, 0], (lat - 1.0) * (X[:, 1] - 0.5),
                          (lat - 1.0) * (X[:, 2] - 0.5)], dim=1)
    assert (u - expect).abs().max() < 1e-8
    assert abs(float(f[right, 0].sum()) - R) < 1e-7 * R
    Fexp = torch.diag(torch.tensor([lam, lat, lat], dtype=DT, device="cuda"))
    assert (grad - Fexp).abs().max() < 1e-8
    # Cauchy stress: sigma_xx = R lam / J (nominal to true), laterals free
    assert abs(float(flux[:, 0, 0].mean()) - R * lam / Jv) < 1e-6 * R
    assert flux[:, 1, 1].abs().max() < 1e-6 * R


def _cantilever(params):
    nodes, el = O.cube_hexa(7, 3, 3)
    nodes = nodes * torch.tensor([2.0, 0.5, 0.5], dtype=DT)
    beam = Solid(nodes.cuda(), el.cuda(), Hyperelastic3D(psi, params))
    X = beam.nodes
    beam.constraints[X[:, 0] == 0.0] = True
    tip = X[:, 0] == 2.0
    beam.forces[tip, 2] = -0.04
    beam.forces[tip, 1] = 0.012
    return beam, nodes, el


@pytest.mark.parametrize("nlgeom", [True, False])
def test_incremental_solve_matches_oracle(nlgeom):
    """A bent block (non-homogeneous, three increments): displacements,
    forces, stresses and deformation gradients against the oracle's
    incremental Newton solve, each converged far below the difference.
    Without nlgeom the tangent stays the one of the undeformed state
    (modified Newton), so that case takes a quarter of the load."""
    params = torch.tensor([MU, LBD], dtype=DT)
    beam, nodes, el = _cantilever(params)
    if not nlgeom:
        beam.forces = 0.25 * beam.forces
    incs = torch.tensor([0.0, 0.4, 0.7, 1.0])
    u, f, q, g, _ = beam.solve(increments=incs, nlgeom=nlgeom, rtol=1e-12, atol=1e-9,
                               stol=1e-12, max_iter=40, precision="fp64")
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    u_o, f_o, q_o, g_o = O.oracle_solve_hyper(
        m, O.neo_hookean_psi, params.expand(len(el), 2), beam.constraints.cpu(),
        beam.forces.cpu(), beam.displacements.cpu(), incs, nlgeom=nlgeom, rtol=1e-12, atol=1e-9,
        max_iter=40)
    assert float(u.abs().max()) > 1e-3  # a visibly nonlinear load
    assert rel(u, u_o) < 1e-8
    assert rel(f, f_o) < 1e-7
    assert rel(q, q_o) < 1e-7
    assert rel(g, g_o) < 1e-9


def test_gradients_match_finite_differences():
    """d/d(mu, lbd) and d/d(forces) of an objective of u and the reactions,
    through three increments, against central differences."""
    params = torch.tensor([MU, LBD], dtype=DT, requires_grad=True)
    beam, _, _ = _cantilever(params)
    incs = torch.tensor([0.0, 0.5, 1.0])
    w = torch.linspace(-1.0, 1.0, beam.n_dofs, device="cuda", dtype=DT).reshape(-1, 3)
    kw = dict(increments=incs, nlgeom=True, rtol=1e-11, atol=1e-13, stol=1e-12, max_iter=40,
              precision="fp64")

    def objective(p, forces, track):
        beam.material = Hyperelastic3D(psi, p).vectorize(beam.n_elem)
        beam.forces = forces
        dp = (p, forces) if track else None
        u, f, _, _, _ = beam.solve(differentiable_parameters=dp, **kw)
        return (w * u).sum() + 1e-3 * f[beam.nodes[:, 0] == 0.0].sum()

    forces0 = beam.forces.detach().clone()
    forces = forces0.clone().requires_grad_(True)
    J = objective(params, forces, True)
    gp, gf = torch.autograd.grad(J, (params, forces))
    with torch.no_grad():
        # steps of a few 1e-4 relative: central-difference error ~1e-7, solve
        # noise ~1e-9 relative at these tolerances
        for i, h in ((0, 0.2), (1, 0.2)):
            e = torch.zeros(2, dtype=DT)
            e[i] = h
            fd = (objective(params.detach() + e, forces0, False)
                  - objective(params.detach() - e, forces0, False)) / (2 * h)
            assert abs(float(gp[i]) - float(fd)) < 1e-5 * abs(float(fd))
        tip = torch.nonzero(beam.nodes[:, 0] == 2.0).ravel()
        for node, comp in ((int(tip[0]), 2), (int(tip[-1]), 1)):
            h = 2e-5
            dF = torch.zeros_like(forces0)
            dF[node, comp] = h
            fd = (objective(params.detach(), forces0 + dF, False)
                  - objective(params.detach(), forces0 - dF, False)) / (2 * h)
            assert abs(float(gf[node, comp]) - float(fd)) < 1e-5 * abs(float(fd))


def test_nlgeom_for_linear_material_is_rejected():
    nodes, el = O.cube_hexa(2, 2, 2)
    fem = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    with pytest.raises(NotImplementedError):
        fem.solve(nlgeom=True)


# ------------------------------------------------ planar hyperelasticity
from test_oracle import plane_strain_uniaxial_closed_form  # noqa: E402
from torchfem import Planar  # noqa: E402
from torchfem.materials import HyperelasticPlaneStrain, HyperelasticPlaneStress  # noqa: E402


def _quad_mesh():
    nodes, el = O.rect_quad(5, 4, 2.0, 1.0)
    inner = (nodes[:, 0] > 0) & (nodes[:, 0] < 2) & (nodes[:, 1] > 0) & (nodes[:, 1] < 1)
    return nodes + 0.05 * torch.sin(5.0 * nodes[:, [1, 0]]) * inner[:, None], el


@pytest.mark.parametrize("kind", ["strain", "stress"])
def test_planar_integrate_material_matches_oracle(kind):
    """k, f, F, stress and state of one planar finite-strain evaluation with
    per-element parameters and thickness, from a previous deformation
    gradient (plane stress: and a previous out-of-plane stretch)."""
    nodes, el = _quad_mesh()
    params = _params(len(el))
    t = 0.5 + torch.rand(len(el), generator=torch.Generator().manual_seed(2), dtype=DT)
    cls = HyperelasticPlaneStress if kind == "stress" else HyperelasticPlaneStrain
    kw = dict(tolerance=1e-13, max_iter=50) if kind == "stress" else {}
    fem = Planar(nodes.cuda(), el.cuda(), cls(psi, params, **kw), thickness=t.cuda())
    m = O.OracleModel(nodes, el, O.Quad1, "planar",
                      O.plane_stress_C(torch.full((len(el),), 1.0), torch.tensor(0.3)),
                      thickness=t)
    g = torch.Generator().manual_seed(3)
    ngp = fem.n_int
    du = 0.05 * torch.randn(len(nodes) * 2, generator=g, dtype=DT)
    grad_prev = torch.eye(2, dtype=DT) + 0.05 * torch.randn(ngp, len(el), 2, 2, generator=g,
                                                              dtype=DT)
    state = (0.05 * torch.randn(ngp, len(el), 1, generator=g, dtype=DT) if kind == "stress"
             else torch.zeros(ngp, len(el), 0, dtype=DT))
    k, f, F, flux, st = fem.integrate_material(
        None, grad_prev.cuda(), None, state.cuda(), du.cuda(), None, 0, False)
    step = (O.hyper_plane_stress(O.neo_hookean_psi, params, tol=1e-13) if kind == "stress"
            else O.hyper_plane_strain(O.neo_hookean_psi, params))
    k_o, f_o, F_o, P_o, st_o = O.oracle_integrate_incremental(
        m, step, grad_prev, torch.zeros(ngp, len(el), 2, 2, dtype=DT), state, du)
    assert rel(F, F_o) < 1e-14
    assert rel(flux, P_o) < 1e-11
    assert rel(f, f_o) < 1e-11
    assert rel(k, k_o) < 1e-10
    if kind == "stress":
        assert rel(st, st_o) < 1e-11


@pytest.mark.parametrize("kind", ["strain", "stress"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_planar_uniaxial_stretch_at_size(kind, precision):
    """A Quad1 plate of thickness 0.3 stretched to 1.8x: plane stress (40 x 40
    cells) gives the 3-D uniaxial closed form and its gradient w.r.t.
    (mu, lbd) (times the thickness), plane strain (10 x 10 cells) the closed
    form with F_zz = 1. The first Newton iterate of an increment puts the
    whole increment into the last column of cells; in plane strain, with no
    out-of-plane contraction, J there grows with the refinement until the
    Neo-Hookean tangent turns indefinite and Newton from that iterate (the
    reference's scheme, which the oracle's restatement shares) diverges
    beyond about 20 x 20 cells."""
    mu, lbd, t = 350.0, 800.0, 0.3
    n = 41 if kind == "stress" else 11
    nodes, el = O.rect_quad(n, n, 1.0, 1.0)
    params = torch.tensor([mu, lbd], dtype=DT, requires_grad=True)
    cls = HyperelasticPlaneStress if kind == "stress" else HyperelasticPlaneStrain
    kw = dict(tolerance=1e-12, max_iter=50) if kind == "stress" else {}
    plate = Planar(nodes.cuda(), el.cuda(), cls(psi, params, **kw),
                   thickness=torch.full((len(el),), t, dtype=DT).cuda())
    X = plate.nodes
    right = X[:, 0] == 1.0
    plate.constraints[X[:, 0] == 0.0, 0] = True
    plate.constraints[X[:, 1] == 0.0, 1] = True
    plate.constraints[right, 0] = True
    plate.displacements[right, 0] = 0.8
    u, f, _, _, _ = plate.solve(increments=torch.tensor([0.0, 0.25, 0.5, 0.75, 1.0]),
                                nlgeom=True, differentiable_parameters=params,
                                precision=precision)
    reaction = f[right, 0].sum()
    if kind == "stress":
        R, dR = uniaxial_stretch_closed_form(mu, lbd, 1.8)
        (grad,) = torch.autograd.grad(reaction, params)
        assert torch.allclose(grad, t * torch.tensor(dR, dtype=DT), rtol=1e-6)
    else:
        R, lat = plane_strain_uniaxial_closed_form(mu, lbd, 1.8)
        assert (u[:, 1] - (lat - 1.0) * X[:, 1]).abs().max() < 1e-9
    assert abs(float(reaction) - R * t) < 1e-7 * R * t
