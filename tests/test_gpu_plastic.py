"""Small-strain J2 plasticity on the incremental Newton path: element
integration against the oracle (stress, state, consistent-tangent element
matrices from a pre-stressed state), the uniaxial closed form at size for
every element type, a two-increment plastic bending path against the
oracle's incremental solve, and gradients through the increments against
central differences.
"""

import pytest
import torch

import oracle as O
from test_oracle import uniaxial_plastic_closed_form
from torchfem import Solid
from torchfem.materials import IsotropicPlasticity3D

pytestmark = pytest.mark.gpu
DT = torch.float64
E0, NU, S0, HH = 1000.0, 0.3, 2.0, 100.0


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def sf(q):
    return S0 + HH * q


def sfp(q):
    return HH + 0.0 * q


def _mesh(name):
    if name == "Tetra1":
        return O.cube_tetra6(3, jitter=0.15, seed=5)
    nodes, el = O.cube_hexa(3, 4, 3)
    if name == "Hexa2":
        nodes, el = O.linear_to_quadratic(nodes, el)
    inner = ((nodes > 0) & (nodes < 1)).all(dim=1)
    return nodes + 0.04 * torch.sin(4.0 * nodes[:, [1, 2, 0]]) * inner[:, None], el


@pytest.mark.parametrize("name", ["Hexa1", "Tetra1", "Hexa2"])
def test_integrate_material_matches_oracle(name):
    """k, f, stress and state of one J2 evaluation from a pre-stressed,
    pre-strained state on a distorted mesh: some points yield, some do not."""
    nodes, el = _mesh(name)
    fem = Solid(nodes.cuda(), el.cuda(), IsotropicPlasticity3D(E0, NU, sf, sfp, tolerance=1e-12))
    m = O.OracleModel(nodes, el, O.ELEMENTS[name], "solid",
                      O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    g = torch.Generator().manual_seed(4)
    ngp = fem.n_int
    du = 2e-4 * torch.randn(len(nodes) * 3, generator=g, dtype=DT)
    grad_prev = torch.eye(3, dtype=DT).expand(ngp, len(el), 3, 3).clone()
    S = 0.6 * torch.randn(ngp, len(el), 3, 3, generator=g, dtype=DT)
    S = 0.5 * (S + S.transpose(-1, -2))
    q = 0.01 * torch.rand(ngp, len(el), 1, generator=g, dtype=DT)
    k, f, F, flux, state = fem.integrate_material(None, grad_prev.cuda(), S.cuda(), q.cuda(),
                                                  du.cuda(), None, 0, False)
    step = O.j2_step(E0, NU, sf, sfp, tol=1e-12)
    k_o, f_o, F_o, P_o, st_o = O.oracle_integrate_incremental(m, step, grad_prev, S, q, du)
    yielded = (st_o > q).float().mean()
    assert 0.05 < float(yielded) < 0.95
    assert rel(F, F_o) < 1e-14
    assert rel(flux, P_o) < 1e-11
    assert rel(state, st_o) < 1e-11
    assert rel(f, f_o) < 1e-11
    assert rel(k, k_o) < 1e-11


def _tension_box(etype, n, strain):
    if etype == "Tetra1":
        nodes, el = O.cube_tetra6(n)
    else:
        nodes, el = O.cube_hexa(n, n, n)
        if etype == "Hexa2":
            nodes, el = O.linear_to_quadratic(nodes, el)
    box = Solid(nodes.cuda(), el.cuda(), IsotropicPlasticity3D(E0, NU, sf, sfp, tolerance=1e-12))
    X = box.nodes
    right = X[:, 0] == 1.0
    box.constraints[X[:, 0] == 0.0, 0] = True
    box.constraints[X[:, 1] == 0.0, 1] = True
    box.constraints[X[:, 2] == 0.0, 2] = True
    box.constraints[right, 0] = True
    box.displacements[right, 0] = strain
    return box, right


@pytest.mark.parametrize("etype,n", [("Hexa1", 13), ("Tetra1", 8), ("Hexa2", 5)])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_uniaxial_tension_closed_form_at_size(etype, n, precision):
    """Pulled to three times its yield strain in four increments, every
    element type reproduces the homogeneous elastoplastic state: reaction =
    the closed-form uniaxial stress, lateral contraction from the elastic
    and plastically incompressible parts, q = the plastic strain."""
    strain = 3.0 * S0 / E0
    box, right = _tension_box(etype, n, strain)
    u, f, flux, grad, state = box.solve(increments=torch.tensor([0.0, 0.25, 0.5, 0.75, 1.0]),
                                        rtol=1e-12, atol=1e-12, stol=1e-12,
                                        precision=precision)
    sig, lat, ep = uniaxial_plastic_closed_form(E0, NU, S0, HH, strain)
    assert abs(float(f[right, 0].sum()) - sig) < 1e-8 * sig
    assert (u[:, 1] - lat * box.nodes[:, 1]).abs().max() < 1e-11
    assert (flux[:, 0, 0] - sig).abs().max() < 1e-8 * sig
    assert (state[:, 0] - ep).abs().max() < 1e-8 * ep


def _beam(E=E0):
    """A 2 x 0.5 x 0.5 Hexa1 beam clamped at x = 0 and bent into the plastic
    range by a tip force (force control: a prescribed tip displacement makes
    the first Newton iterate of every increment concentrate the increment in
    the last element layer, which Newton from the previous increment, the
    reference's initial guess, does not recover from)."""
    nodes, el = O.cube_hexa(7, 3, 3)
    nodes = nodes * torch.tensor([2.0, 0.5, 0.5], dtype=DT)
    beam = Solid(nodes.cuda(), el.cuda(), IsotropicPlasticity3D(E, NU, sf, sfp, tolerance=1e-12))
    X = beam.nodes
    beam.constraints[X[:, 0] == 0.0] = True
    tip = X[:, 0] == 2.0
    beam.forces[tip, 2] = -0.08 / int(tip.sum())
    return beam, nodes, el, tip


def test_bending_path_matches_oracle():
    """The beam loaded in two increments, the second from the plastic state
    the first leaves: displacements, reactions, stresses and plastic strains
    after each increment against the oracle's incremental solve."""
    beam, nodes, el, tip = _beam()
    incs = torch.tensor([0.0, 0.5, 1.0])
    u, f, q, g, s = beam.solve(increments=incs, rtol=1e-12, atol=1e-11, stol=1e-12,
                               max_iter=40, precision="fp64", return_intermediate=True)
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    step = O.j2_step(E0, NU, sf, sfp, tol=1e-12)
    for n in (1, 2):
        u_o, f_o, q_o, g_o, s_o = O.oracle_solve_incremental(
            m, step, 1, beam.constraints.cpu(), beam.forces.cpu(), beam.displacements.cpu(),
            incs[: n + 1], nlgeom=False, rtol=1e-12, atol=1e-11, max_iter=40)
        assert rel(u[n], u_o) < 1e-8
        assert rel(f[n], f_o) < 1e-7
        assert rel(q[n], q_o) < 1e-7
        assert rel(g[n], g_o) < 1e-9
        assert rel(s[n], s_o) < 1e-7
    assert 0.2 < float((s[-1] > 0).double().mean()) < 0.9  # a partly plastic beam


def test_gradients_match_finite_differences():
    """d(tip deflection)/dE through a plastic loading path of two increments,
    against central differences."""
    incs = torch.tensor([0.0, 0.5, 1.0])
    kw = dict(increments=incs, rtol=1e-12, atol=1e-13, stol=1e-12, max_iter=40,
              precision="fp64")

    def deflection(E, track):
        beam, _, _, tip = _beam(E)
        u, f, _, _, state = beam.solve(differentiable_parameters=E if track else None, **kw)
        assert float(state.max()) > 0.0
        return u[tip, 2].mean()

    E = torch.tensor(E0, dtype=DT, requires_grad=True)
    (g,) = torch.autograd.grad(deflection(E, True), E)
    with torch.no_grad():
        h = 0.5
        fd = (deflection(torch.tensor(E0 + h, dtype=DT), False)
              - deflection(torch.tensor(E0 - h, dtype=DT), False)) / (2 * h)
    assert abs(float(g) - float(fd)) < 1e-4 * abs(float(fd))
