"""Distributed loads, shape-function evaluation and the remaining drop-in API
methods on the GPU, against the CPU oracle and the reference's own load
invariants (tests/test_loads.py: closed surfaces carry no net pressure force,
pressure acts along the outward normal, gravity equals weight, thickness
scales planar loads, interior faces are not loaded, heat loads stay scalar).
"""

import pytest
import torch

import oracle as O
from torchfem import Planar, PlanarHeat, Solid, SolidHeat
from torchfem.materials import (IsotropicConductivity2D, IsotropicConductivity3D,
                                IsotropicElasticity3D, IsotropicElasticityPlaneStress)

pytestmark = pytest.mark.gpu
DT = torch.float64


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def _mesh(name, distort=True):
    if name == "Quad1":
        nodes, el = O.rect_quad(5, 4, 2.0, 1.0)
        if distort:
            inner = (nodes[:, 0] > 0) & (nodes[:, 0] < 2) & (nodes[:, 1] > 0) & (nodes[:, 1] < 1)
            nodes = nodes + 0.05 * torch.sin(5.0 * nodes[:, [1, 0]]) * inner[:, None]
        return nodes, el
    if name == "Tetra1":
        return O.cube_tetra6(3, jitter=0.15 if distort else 0.0, seed=2)
    if name == "Hexa1":
        nodes, el = O.cube_hexa(4, 3, 5)
    else:
        nodes, el = O.linear_to_quadratic(*O.cube_hexa(3, 4, 3))
    if distort:
        inner = ((nodes > 0) & (nodes < 1)).all(dim=1)
        nodes = nodes + 0.04 * torch.sin(4.0 * nodes[:, [1, 2, 0]]) * inner[:, None]
    return nodes, el


def _model(name, nodes, el, heat=False):
    if name == "Quad1":
        t = 0.5 + torch.rand(len(el), generator=torch.Generator().manual_seed(4), dtype=DT)
        if heat:
            return PlanarHeat(nodes.cuda(), el.cuda(), IsotropicConductivity2D(1.0), thickness=t.cuda())
        return Planar(nodes.cuda(), el.cuda(), IsotropicElasticityPlaneStress(1000.0, 0.3),
                      thickness=t.cuda())
    if heat:
        return SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(1.0))
    return Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))


def _oracle(name, nodes, el):
    return O.OracleModel(nodes, el, O.ELEMENTS[name], "solid" if name != "Quad1" else "planar",
                         O.iso3d_C(torch.full((len(el),), 1.0), 0.3) if name != "Quad1"
                         else O.plane_stress_C(torch.full((len(el),), 1.0), torch.tensor(0.3)))


def _masks(nodes):
    x = nodes[:, 0]
    lo = nodes.min(0).values
    hi = nodes.max(0).values
    return {
        "face": x == hi[0],
        "all": torch.ones(len(nodes), dtype=torch.bool),
        "two_faces": (x == lo[0]) | (nodes[:, 1] == hi[1]),
        # a box cutting through the mesh: its interior facets must not be loaded
        "box": (nodes < lo + 0.6 * (hi - lo)).all(dim=1),
        "none": torch.zeros(len(nodes), dtype=torch.bool),
    }


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
@pytest.mark.parametrize("which", ["face", "all", "two_faces", "box", "none"])
def test_boundary_facets_match_the_oracle_in_order(name, which):
    """The facet list (node ids, winding and order) is identical to the
    reference's torch.unique-based selection (base.py:452-470)."""
    nodes, el = _mesh(name)
    mask = _masks(nodes)[which]
    m = _model(name, nodes, el)
    o = _oracle(name, nodes, el)
    got = m._boundary_facets(mask.cuda()).cpu()
    ref = o.boundary_facets(mask)
    assert got.shape == ref.shape
    assert torch.equal(got, ref)


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
@pytest.mark.parametrize("which", ["face", "all", "box"])
def test_traction_loads_match_the_oracle(name, which):
    """Uniform and per-facet tractions (base.py:472-494) to 1e-13."""
    nodes, el = _mesh(name)
    mask = _masks(nodes)[which]
    m = _model(name, nodes, el)
    o = _oracle(name, nodes, el)
    d = nodes.shape[1]
    integrate = m.integrate_line_load if name == "Quad1" else m.integrate_surface_load
    t = torch.tensor([0.3, -1.2, 0.7][:d], dtype=DT)
    F = integrate(mask.cuda(), t.cuda())
    Fo, conn = o.integrate_surface_load(mask, t)
    assert rel(F, Fo) < 1e-13
    per = torch.randn(len(conn), d, generator=torch.Generator().manual_seed(7), dtype=DT)
    F = integrate(mask.cuda(), per.cuda())
    Fo, _ = o.integrate_surface_load(mask, per)
    assert rel(F, Fo) < 1e-13


@pytest.mark.parametrize("name", ["Tetra1", "Hexa1", "Hexa2"])
def test_pressure_on_a_closed_surface_has_no_net_force(name):
    """tests/test_loads.py:121-128: sum of p n dA over a closed surface is 0."""
    nodes, el = _mesh(name, distort=(name != "Hexa2"))
    m = _model(name, nodes, el)
    F = m.integrate_surface_load(torch.ones(len(nodes), dtype=torch.bool, device="cuda"), 2.5)
    assert F.shape == (len(nodes), 3)
    assert float(F.sum(0).abs().max()) < 1e-12 * float(F.abs().sum())


@pytest.mark.parametrize("name", ["Tetra1", "Hexa1", "Hexa2"])
def test_pressure_acts_along_the_outward_normal(name):
    """tests/test_loads.py:130-133: pressure p on the face x = 1 of the unit
    cube gives the total force p * area along +x."""
    nodes, el = _mesh(name, distort=False)
    m = _model(name, nodes, el)
    face = m.nodes[:, 0] == 1.0
    F = m.integrate_surface_load(face, 3.0)
    assert torch.allclose(F.sum(0).cpu(), torch.tensor([3.0, 0.0, 0.0], dtype=DT), atol=1e-13)
    assert float(F[~face].abs().max()) == 0.0


def test_planar_edge_pressure_and_line_traction():
    """tests/test_loads.py:167-195: a pressure on the closed boundary of a
    planar body has no net force; on the edge x = 2 it acts along +x with the
    edge length; a uniform traction splits equally onto the two edge nodes."""
    nodes, el = O.rect_quad(5, 4, 2.0, 1.0)
    m = _model("Quad1", nodes, el)
    F = m.integrate_line_load(torch.ones(len(nodes), dtype=torch.bool, device="cuda"), 1.5)
    assert float(F.sum(0).abs().max()) < 1e-13
    edge = m.nodes[:, 0] == 2.0
    F = m.integrate_line_load(edge, 2.0)
    assert torch.allclose(F.sum(0).cpu(), torch.tensor([2.0, 0.0], dtype=DT), atol=1e-13)
    F = m.integrate_line_load(edge, torch.tensor([0.0, -1.0], dtype=DT))
    ends = edge & ((m.nodes[:, 1] == 0.0) | (m.nodes[:, 1] == 1.0))
    assert torch.allclose(F[ends, 1].cpu(), torch.full((2,), -1.0 / 6.0, dtype=DT), atol=1e-14)
    assert torch.allclose(F[edge & ~ends, 1].cpu(), torch.full((2,), -1.0 / 3.0, dtype=DT), atol=1e-14)


def test_heat_flux_stays_scalar():
    """tests/test_loads.py:155-164: a scalar flux on a heat model gives one
    value per node, the flux times the loaded area in total."""
    nodes, el = _mesh("Hexa1", distort=False)
    m = _model("Hexa1", nodes, el, heat=True)
    face = m.nodes[:, 2] == 1.0
    F = m.integrate_surface_load(face, 4.0)
    assert F.shape == (len(nodes), 1)
    assert abs(float(F.sum()) - 4.0) < 1e-13


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_body_load_equals_weight_and_matches_the_shape_integrals(name):
    """tests/test_loads.py:211-222: a uniform load per unit volume sums to
    load * volume (times thickness for planar models); every node carries
    sum_e int N dV (the oracle's shape functions); a per-element load equal
    to the uniform one gives the same result."""
    nodes, el = _mesh(name)
    m = _model(name, nodes, el)
    d = nodes.shape[1]
    g = torch.tensor([0.0, -9.81, 1.5][:d], dtype=DT)
    F = m.integrate_body_load(g.cuda())
    o = _oracle(name, nodes, el)
    N, _, detJ = o.eval_shape_functions(o.etype.ipoints)
    w = o.etype.iweights.to(DT)
    Ne = (detJ * w[:, None]).T @ N  # [E, npe]
    if name == "Quad1":
        Ne = Ne * m.thickness.cpu()[:, None]
    vol = torch.zeros(len(nodes), dtype=DT).index_add_(0, el.reshape(-1), Ne.reshape(-1))
    assert rel(F, vol[:, None] * g) < 1e-13
    assert torch.allclose(F.sum(0).cpu(), Ne.sum() * g, rtol=1e-13, atol=1e-13)
    Fe = m.integrate_body_load(g.cuda().expand(len(el), d).contiguous())
    assert rel(Fe, F) < 1e-15


def test_loads_reject_gradients_and_wrong_shapes():
    nodes, el = _mesh("Hexa1", distort=False)
    m = _model("Hexa1", nodes, el)
    face = m.nodes[:, 0] == 1.0
    with pytest.raises(NotImplementedError):
        m.integrate_surface_load(face, torch.tensor([1.0, 0.0, 0.0], device="cuda", requires_grad=True))
    with pytest.raises(ValueError):
        m.integrate_surface_load(face, torch.ones(3, 3, dtype=DT, device="cuda"))
    with pytest.raises(NotImplementedError):
        m.integrate_line_load(face, torch.ones(3, dtype=DT, device="cuda"))
    with pytest.raises(ValueError):
        m.integrate_surface_load(face.float(), 1.0)


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_eval_shape_functions_match_the_oracle(name):
    """N, B and detJ (base.py:293-314) at the Gauss points and at random
    local points, against the oracle's einsum form."""
    nodes, el = _mesh(name)
    m = _model(name, nodes, el)
    o = _oracle(name, nodes, el)
    for xi in (o.etype.ipoints, 0.3 * torch.rand(5, nodes.shape[1], generator=torch.Generator().manual_seed(1), dtype=DT),
               o.etype.ipoints[0]):
        N, B, detJ = m.eval_shape_functions(xi)
        No, Bo, Jo = o.eval_shape_functions(xi)
        assert B.shape == Bo.shape and detJ.shape == Jo.shape
        assert rel(N, No) < 1e-14
        assert rel(B, Bo) < 1e-13
        assert rel(detJ, Jo) < 1e-13
    # element volumes by the element's own rule (for the curved 20-node
    # elements the 2x2x2 rule is not exact, so compare with the oracle's)
    _, _, Jo = o.eval_shape_functions(o.etype.ipoints)
    vol_o = (Jo * o.etype.iweights.to(DT)[:, None]).sum(0)
    assert rel(m.integrate_field(), vol_o) < 1e-13
    if name != "Hexa2":
        assert abs(float(vol_o.sum()) - (2.0 if name == "Quad1" else 1.0)) < 1e-12


def test_negative_jacobian_in_shape_functions_raises():
    """base.py:311-312: mirrored elements have detJ < 0."""
    nodes, el = O.cube_hexa(3, 2, 2)
    m = Solid(nodes.cuda(), el[:, [1, 0, 3, 2, 5, 4, 7, 6]].contiguous().cuda(),
              IsotropicElasticity3D(1.0, 0.3))
    with pytest.raises(ValueError, match="Negative Jacobian"):
        m.eval_shape_functions(m.etype.ipoints)


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_rigid_modes_are_in_the_null_space_of_the_unconstrained_stiffness(name):
    """compute_B (base.py:316-344): translations and rotations e_k x X give
    K B = 0 before boundary conditions."""
    nodes, el = _mesh(name)
    m = _model(name, nodes, el)
    B = m.compute_B()
    d = m.n_dof_per_node
    assert B.shape == (m.n_dofs, 3 if d == 2 else 6)
    K = m.assemble_matrix(m.k0(), torch.zeros(0, dtype=torch.int64)).to_dense()
    KB = K @ B
    assert float(KB.abs().max()) < 1e-10 * float(K.abs().max()) * float(B.abs().max())
    # the columns are the translations and the rotations about the axes
    X = m.nodes
    if d == 3:
        assert torch.equal(B[0::3, 4], X[:, 2]) and torch.equal(B[2::3, 4], -X[:, 0])
    else:
        assert torch.equal(B[0::2, 2], X[:, 1]) and torch.equal(B[1::2, 2], -X[:, 0])


# ------------------------------------------------------------- API methods
@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2", "Tetra1-heat"])
def test_integrate_material_matches_the_oracle_and_accumulates(name):
    """integrate_material (base.py:247-278, 982-1092, 1177-1286) from the
    initial state against the oracle's integration of du, and two increments
    du1, du2 (the second from the first's gradient and flux) against one of
    du1 + du2: k, f, gradients and fluxes per Gauss point."""
    heat = name.endswith("heat")
    base = name.split("-")[0]
    nodes, el = _mesh(base)
    m = _model(base, nodes, el, heat=heat)
    if heat:
        o = O.OracleModel(nodes, el, O.ELEMENTS[base], "heat", O.iso_KAPPA(torch.ones(len(el), dtype=DT)))
    elif base == "Quad1":
        o = O.OracleModel(nodes, el, O.Quad1, "planar",
                          O.plane_stress_C(torch.full((len(el),), 1000.0, dtype=DT), torch.tensor(0.3)),
                          thickness=m.thickness.cpu())
    else:
        o = O.OracleModel(nodes, el, O.ELEMENTS[base], "solid", O.iso3d_C(torch.full((len(el),), 1000.0), 0.3))
    gen = torch.Generator().manual_seed(5)
    du1 = 1e-3 * torch.randn(m.n_dofs, generator=gen, dtype=DT)
    du2 = 1e-3 * torch.randn(m.n_dofs, generator=gen, dtype=DT)
    shape = (m.n_int, m.n_elem, *m.n_flux)
    grad0 = m.initial_grad.expand(shape).contiguous()
    flux0 = torch.zeros(shape, dtype=DT, device="cuda")
    state0 = torch.zeros(m.n_int, m.n_elem, 0, dtype=DT, device="cuda")
    u0 = torch.zeros(m.n_dofs, dtype=DT, device="cuda")
    k, f, grad, flux, _ = m.integrate_material(u0, grad0, flux0, state0, du1.cuda(), None, 0, False)
    ko, fo, go, qo = o.integrate(du1)
    assert rel(k, ko) < 1e-12
    assert rel(f, fo) < 1e-12
    assert rel(grad, go.reshape(shape)) < 1e-12
    assert rel(flux, qo.reshape(shape)) < 1e-12
    k2, f2, grad2, flux2, _ = m.integrate_material(du1.cuda(), grad, flux, state0, du2.cuda(), None, 1,
                                                   False, compute_stiffness=False)
    assert k2 is None
    _, fo, go, qo = o.integrate(du1 + du2, need_k=False)
    assert rel(f2, fo) < 1e-12
    assert rel(grad2, go.reshape(shape)) < 1e-12
    assert rel(flux2, qo.reshape(shape)) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_K_property_is_the_assembled_matrix(precision):
    """FEM.K after a solve equals assemble_matrix(k0, con) of the oracle,
    with Dirichlet rows and columns replaced by the identity (base.py:398-426),
    for an index and for a boolean constraint mask."""
    nodes, el = _mesh("Hexa1")
    m = _model("Hexa1", nodes, el)
    o = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1000.0), 0.3))
    m.constraints[m.nodes[:, 0] == 0.0] = True
    m.forces[m.nodes[:, 0] == 1.0, 2] = -0.1
    m.solve(precision=precision, stol=1e-10)
    K = m.K.to_dense().cpu()
    con = torch.nonzero(m.constraints.cpu().ravel()).ravel()
    Ko = torch.tensor(o.csr(o.assemble_matrix(o.k0(), con)).toarray())
    assert rel(K, Ko) < 1e-13
    Kb = m.assemble_matrix(m.k0(), m.constraints.ravel()).to_dense().cpu()
    Ki = m.assemble_matrix(m.k0(), con.cuda()).to_dense().cpu()
    assert torch.equal(Kb, Ki)
    assert rel(Ki, Ko) < 1e-13


def test_cached_solve_warm_starts_and_agrees():
    """use_cached_solve=True (CachedSolve, sparse.py:103-128): the second solve
    of the same problem starts from the first solution, needs (almost) no
    iterations and returns the same field; gradients agree as well."""
    nodes, el = _mesh("Hexa1")
    E = torch.full((len(el),), 1000.0, dtype=DT, device="cuda", requires_grad=True)
    m = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(E, 0.3))
    m.constraints[m.nodes[:, 0] == 0.0] = True
    m.forces[m.nodes[:, 0] == 1.0, 2] = -0.1
    u_ref, _, _, _, _ = m.solve(precision="fp64", stol=1e-10)
    outs = []
    for _ in range(2):
        # a fresh material per solve: its moduli hold the graph back to E
        m.material = IsotropicElasticity3D(E, 0.3)
        u, _, _, _, _ = m.solve(precision="fp64", stol=1e-10, use_cached_solve=True,
                                differentiable_parameters=E)
        (g,) = torch.autograd.grad(u.sum(), E)
        outs.append((u.detach(), g, m.last_solve_info))
    assert rel(outs[0][0], u_ref) < 1e-9 and rel(outs[1][0], u_ref) < 1e-9
    assert rel(outs[1][1], outs[0][1]) < 1e-8
    it_first = outs[0][2]["forward"]["pcg"][0]["iterations"][0]
    it_second = sum(p["iterations"][0] for p in outs[1][2]["forward"]["pcg"])
    assert it_second < it_first // 4
    assert outs[1][2]["adjoint"]["iterations"][0] < outs[0][2]["adjoint"]["iterations"][0] // 4


@pytest.mark.parametrize("name", ["Tetra1", "Hexa1", "Hexa2"])
def test_per_gauss_point_outputs(name):
    """solve(aggregate_integration_points=False) returns stresses and
    deformation gradients per Gauss point (base.py:844-867), equal to the
    oracle's integration of the solution."""
    nodes, el = _mesh(name)
    m = _model(name, nodes, el)
    o = O.OracleModel(nodes, el, O.ELEMENTS[name], "solid", O.iso3d_C(torch.full((len(el),), 1000.0), 0.3))
    m.constraints[m.nodes[:, 2] == 0.0] = True
    m.forces[m.nodes[:, 2] == 1.0, 0] = 0.05
    u, f, sig, F, _ = m.solve(precision="fp64", stol=1e-12, aggregate_integration_points=False)
    assert sig.shape == (m.n_int, m.n_elem, 3, 3) and F.shape == (m.n_int, m.n_elem, 3, 3)
    _, _, go, qo = o.integrate(u.detach().cpu().reshape(-1), need_k=False)
    assert rel(sig, qo) < 1e-12
    assert rel(F, go) < 1e-12
    u2, _, sig2, F2, _ = m.solve(precision="fp64", stol=1e-12)
    assert rel(sig2, sig.mean(0)) < 1e-12 and rel(F2, F.mean(0)) < 1e-12
