"""Pins the CPU oracle (oracle/fem_oracle.py) before it is trusted as the
checker of the HIP path. Each test fixes the oracle by something other than
its own formulas: the reference's own known-answer tests and golden value,
closed forms, exact symbolic integration, invariants and finite differences.
"""

import itertools
import math

import numpy as np
import pytest
import sympy
import torch

import oracle as O

DT = torch.float64


# ---------------------------------------------------------------- elements
# tests/test_elements.py:25-66 of the reference, applied to the restated tables
@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_quadrature_volume(name):
    el = O.ELEMENTS[name]
    vol = 0.0
    for w, q in zip(el.iweights, el.ipoints):
        J = el.B(q) @ el.iso_coords
        vol += float(w) * float(torch.linalg.det(J))
    assert math.isclose(vol, el.iso_volume, rel_tol=1e-12)


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_shape_gradient_is_derivative_of_N(name):
    el = O.ELEMENTS[name]
    for q in el.ipoints:
        q = q.clone().requires_grad_(True)
        for i in range(el.nodes):
            (g,) = torch.autograd.grad(el.N(q)[i], q)
            assert torch.allclose(g, el.B(q.detach())[:, i], atol=1e-13)


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_kronecker_property(name):
    el = O.ELEMENTS[name]
    assert torch.allclose(el.N(el.iso_coords), torch.eye(el.nodes, dtype=DT), atol=1e-14)


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_weights_sum_to_iso_volume(name):
    el = O.ELEMENTS[name]
    assert math.isclose(float(el.iweights.sum()), el.iso_volume, rel_tol=1e-12)


# ---------------------------------------------------------------- meshes
def test_cube_hexa_numbering():
    """mesh.py:8-46: node i*Ny*Nz + j*Nz + k at (x_i, y_j, z_k)."""
    nodes, el = O.cube_hexa(3, 4, 5)
    assert nodes.shape == (60, 3) and el.shape == (2 * 3 * 4, 8)
    assert torch.allclose(nodes[1 * 20 + 2 * 5 + 3], torch.tensor([0.5, 2 / 3, 0.75], dtype=DT), rtol=0, atol=1e-15)
    # first element: [n0, n1, n3, n2, n4, n5, n7, n6] with n1 = +x, n2 = +y, n4 = +z
    assert el[0].tolist() == [0, 20, 25, 5, 1, 21, 26, 6]


def test_cube_tetra6_is_conforming_and_positive():
    nodes, el = O.cube_tetra6(3, jitter=0.1, seed=0)
    m = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(torch.ones(len(el))))
    _, _, detJ = m.eval_shape_functions(O.Tetra1.ipoints)
    assert (detJ > 0).all()
    assert math.isclose(float((detJ / 6.0).sum()), 1.0, rel_tol=1e-12)  # total volume
    faces = el[:, O.Tetra1.facets].reshape(-1, 3).sort(dim=1).values
    _, counts = torch.unique(faces, dim=0, return_counts=True)
    assert counts.max() == 2  # every face is shared by at most two tets
    n_boundary = int((counts == 1).sum())
    assert n_boundary == 6 * 2 * 3 * 3  # 2 triangles per boundary quad


def test_linear_to_quadratic_counts():
    nodes, el = O.linear_to_quadratic(*O.cube_hexa(3, 3, 3))
    # 27 vertices + 3*3*3*... edges: 3 * (2 * 3 * 3) = 54 edges
    assert nodes.shape[0] == 27 + 54 and el.shape == (8, 20)
    mid = nodes[el[:, 8:]]
    iso = O.Hexa2.iso_coords[8:]
    corners = nodes[el[:, :8]]
    # every mid-side node is the midpoint of its two corners
    pairs = torch.tensor([[0, 1], [1, 2], [2, 3], [3, 0], [4, 5], [5, 6], [6, 7], [7, 4],
                          [0, 4], [1, 5], [2, 6], [3, 7]])
    assert torch.allclose(mid, 0.5 * (corners[:, pairs[:, 0]] + corners[:, pairs[:, 1]]))
    assert iso.shape == (12, 3)


# ---------------------------------------------------------------- pattern
@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_hexa_pattern_nnz_closed_form(n):
    nodes, el = O.cube_hexa(n, n, n)
    _, glob_idx, _, _ = O.build_pattern(el, len(nodes), 3)
    assert glob_idx.shape[1] == O.hexa_nnz_closed_form(n)


@pytest.mark.parametrize(
    "mesh,ndof",
    [
        (lambda: O.cube_hexa(3, 4, 3), 3),
        (lambda: O.rect_quad(4, 3), 2),
        (lambda: O.cube_tetra6(2), 1),
        (lambda: O.linear_to_quadratic(*O.cube_hexa(3, 3, 2)), 3),
    ],
)
def test_pattern_matches_brute_force(mesh, ndof):
    """glob_idx is the sorted set of element DOF couplings plus the diagonal,
    recomputed here from a dense incidence product."""
    nodes, el = mesh()
    n = len(nodes) * ndof
    idx, glob_idx, k_map, diag_map = O.build_pattern(el, len(nodes), ndof)
    A = torch.zeros(n, n, dtype=torch.bool)
    for row in idx:
        A[row[:, None], row[None, :]] = True
    A |= torch.eye(n, dtype=torch.bool)
    expect = A.nonzero().T  # row-major order
    assert torch.equal(glob_idx, expect)
    # k_map sends each element entry to the slot of its (row, col)
    rows = idx[:, :, None].expand(-1, -1, idx.shape[1]).reshape(-1)
    cols = idx[:, None, :].expand(-1, idx.shape[1], -1).reshape(-1)
    assert torch.equal(glob_idx[0, k_map.long()], rows)
    assert torch.equal(glob_idx[1, k_map.long()], cols)
    assert torch.equal(glob_idx[0, diag_map.long()], torch.arange(n))
    assert torch.equal(glob_idx[1, diag_map.long()], torch.arange(n))


def test_pattern_orphan_node_keeps_only_its_diagonal():
    nodes, el = O.rect_quad(3, 3)
    nodes = torch.cat([nodes, torch.tensor([[5.0, 5.0]], dtype=DT)])
    _, glob_idx, _, _ = O.build_pattern(el, len(nodes), 2)
    last = glob_idx[:, glob_idx[0] >= 18]
    assert last.tolist() == [[18, 19], [18, 19]]


# ---------------------------------------------------------------- materials
def test_iso_C_lame_and_symmetries():
    """tests/test_materials.py:89-131: C symmetric, Lame constants."""
    E, nu = 1000.0, 0.3
    C = O.iso3d_C(torch.tensor(E), torch.tensor(nu))
    lbd = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    assert math.isclose(float(C[0, 0, 1, 1]), lbd)
    assert math.isclose(float(C[0, 1, 0, 1]), G)
    assert math.isclose(float(C[0, 0, 0, 0]), lbd + 2 * G)
    assert torch.allclose(C, C.permute(2, 3, 0, 1))
    assert torch.allclose(C, C.permute(1, 0, 2, 3))


def test_plane_stress_C_closed_form():
    E, nu = 210.0, 0.25
    C = O.plane_stress_C(torch.tensor(E), torch.tensor(nu))
    sig_xx = C[0, 0, 0, 0] * 1.0  # uniaxial strain eps_xx = 1 with sigma_zz = 0
    assert math.isclose(float(sig_xx), E / (1 - nu**2))
    assert math.isclose(float(C[0, 1, 0, 1]), E / (2 * (1 + nu)))


# ---------------------------------------------------------------- element matrices
def _unit_cube_hexa_ke_symbolic(E, nu):
    """Exact Ke of a unit-cube Hexa1 by symbolic integration of B^T C B,
    with the trilinear functions built from the corner coordinates only."""
    x, y, z = sympy.symbols("x y z")
    corners = O.Hexa1.iso_coords.tolist()
    N = []
    for c in corners:
        f = 1
        for v, s in zip((x, y, z), c):
            f *= v if s > 0 else (1 - v)
        N.append(f)
    lam = sympy.Rational(E) * sympy.Rational(nu) / ((1 + sympy.Rational(nu)) * (1 - 2 * sympy.Rational(nu)))
    G = sympy.Rational(E) / (2 * (1 + sympy.Rational(nu)))
    grads = [[sympy.diff(Ni, v) for v in (x, y, z)] for Ni in N]
    K = sympy.zeros(24, 24)
    for p, q in itertools.product(range(8), range(8)):
        bp, bq = grads[p], grads[q]
        for i, k in itertools.product(range(3), range(3)):
            expr = lam * bp[i] * bq[k] + G * bp[k] * bq[i] + (G * sum(bp[j] * bq[j] for j in range(3)) if i == k else 0)
            K[3 * p + i, 3 * q + k] = sympy.integrate(sympy.expand(expr), (x, 0, 1), (y, 0, 1), (z, 0, 1))
    return torch.tensor([[float(K[i, j]) for j in range(24)] for i in range(24)], dtype=DT)


def test_unit_cube_hexa1_ke_matches_exact_integration():
    nodes, el = O.cube_hexa(2, 2, 2)
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.tensor([1000.0]), 0.3))
    k = m.k0()[0]
    ref = _unit_cube_hexa_ke_symbolic(1000, sympy.Rational(3, 10))
    assert torch.allclose(k, ref, rtol=1e-12, atol=1e-10)


def _affine_mesh(name):
    g = torch.Generator().manual_seed(3)
    A = torch.eye(3, dtype=DT) + 0.2 * torch.rand(3, 3, generator=g, dtype=DT)
    if name == "Hexa1":
        nodes, el = O.cube_hexa(3, 2, 2)
    elif name == "Hexa2":
        nodes, el = O.linear_to_quadratic(*O.cube_hexa(2, 3, 2))
    else:
        nodes, el = O.cube_tetra6(2)
    return nodes @ A.T, el


@pytest.mark.parametrize("name", ["Hexa1", "Hexa2", "Tetra1"])
def test_ke_energy_of_linear_field_and_rigid_modes(name):
    """u = G x gives u^T Ke u = V eps:C:eps exactly (constant strain on an
    affine element); rigid motions give Ke u = 0; Ke is symmetric."""
    nodes, el = _affine_mesh(name)
    E = torch.linspace(500.0, 1500.0, len(el), dtype=DT)
    m = O.OracleModel(nodes, el, O.ELEMENTS[name], "solid", O.iso3d_C(E, 0.3))
    k = m.k0()
    _, _, detJ = m.eval_shape_functions(m.etype.ipoints)
    vol = torch.einsum("i,ie->e", m.etype.iweights.to(DT), detJ)
    Gm = torch.tensor([[1e-3, 2e-4, -1e-4], [3e-4, -1e-3, 5e-5], [1e-4, 2e-4, 7e-4]], dtype=DT)
    ue = (nodes[el] @ Gm.T).reshape(len(el), -1)
    eps = 0.5 * (Gm + Gm.T)
    lbd = E * 0.3 / (1.3 * 0.4)
    G = E / 2.6
    dens = lbd * torch.trace(eps) ** 2 + 2 * G * (eps * eps).sum()
    assert torch.allclose(torch.einsum("ei,eij,ej->e", ue, k, ue), vol * dens, rtol=1e-11)
    assert torch.allclose(k, k.transpose(1, 2), atol=1e-10)
    W = torch.tensor([[0.0, -0.3, 0.2], [0.3, 0.0, -0.1], [-0.2, 0.1, 0.0]], dtype=DT)
    urig = (nodes[el] @ W.T + torch.tensor([0.1, -0.2, 0.3], dtype=DT)).reshape(len(el), -1)
    assert torch.einsum("eij,ej->ei", k, urig).abs().max() < 1e-9


def test_heat_k0_symmetric_zero_row_sums_scales():
    """tests/test_base.py:108-131."""
    nodes, el = O.cube_tetra6(2, jitter=0.1)
    kap = torch.full((len(el),), 400.0, dtype=DT)
    m = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(kap))
    k = m.k0()
    assert torch.allclose(k, k.transpose(1, 2))
    assert torch.allclose(k.sum(-1), torch.zeros(len(el), 4, dtype=DT), atol=1e-9)
    m2 = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(2 * kap), pattern=m.pattern)
    assert torch.allclose(m2.k0(), 2 * k)


def test_heat_unit_tetrahedron_closed_form():
    """The unit tetrahedron (0,0,0), (1,0,0), (0,1,0), (0,0,1) has the constant
    shape gradients g0 = (-1,-1,-1), g1 = e1, g2 = e2, g3 = e3 and volume 1/6,
    so its conduction matrix is k_ab = (1/6) g_a . KAPPA g_b exactly. A wrong
    factor (volume, weight, KAPPA scaling or transposition) fails this."""
    nodes = torch.tensor([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]], dtype=DT)
    el = torch.tensor([[0, 1, 2, 3]])
    g = torch.tensor([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]], dtype=DT)
    for KAP in (2.5 * torch.eye(3, dtype=DT),
                torch.tensor([[2.0, 0.3, -0.2], [0.3, 1.5, 0.1], [-0.2, 0.1, 1.2]], dtype=DT)):
        m = O.OracleModel(nodes, el, O.Tetra1, "heat", KAP[None])
        k = m.k0()[0]
        assert torch.allclose(k, g @ KAP @ g.T / 6.0, atol=1e-14)
    # isotropic: (kappa / 6) [[3,-1,-1,-1],[-1,1,0,0],[-1,0,1,0],[-1,0,0,1]]
    m = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(torch.tensor([6.0], dtype=DT)))
    expect = torch.tensor([[3.0, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]], dtype=DT)
    assert torch.allclose(m.k0()[0], expect, atol=1e-14)


@pytest.mark.parametrize("name", ["Tetra1", "Hexa1", "Hexa2"])
def test_heat_patch_test_flux(name):
    """A linear temperature T = a.x + c prescribed on the boundary of a body of
    constant (anisotropic) conductivity is reproduced inside, and every
    Gauss point carries the flux KAPPA a (conductivity.py:90-92) with zero
    net heat at interior nodes. Pins the heat integration up to every factor
    that a symmetric, zero-row-sum matrix could hide."""
    if name == "Tetra1":
        nodes, el = O.cube_tetra6(3, jitter=0.15, seed=3)
    elif name == "Hexa1":
        nodes, el = O.cube_hexa(4, 4, 4)
    else:
        nodes, el = O.linear_to_quadratic(*O.cube_hexa(3, 3, 3))
    KAP = torch.tensor([[2.0, 0.3, -0.2], [0.3, 1.5, 0.1], [-0.2, 0.1, 1.2]], dtype=DT)
    m = O.OracleModel(nodes, el, O.ELEMENTS[name], "heat", KAP.expand(len(el), 3, 3))
    a = torch.tensor([0.7, -0.4, 0.25], dtype=DT)
    T_exact = (nodes @ a + 0.1)[:, None]
    bnd = _boundary(nodes)
    T, f, flux, _ = m.solve(bnd[:, None], torch.zeros_like(T_exact), T_exact)
    assert torch.allclose(T, T_exact, atol=1e-12)
    assert torch.allclose(flux, (KAP @ a).expand(len(el), 1, 3), atol=1e-12)
    assert f[~bnd].abs().max() < 1e-12
    # the net boundary heat balances: sum of reactions = 0
    assert abs(float(f.sum())) < 1e-12


# ---------------------------------------------------------------- solves
def _boundary(nodes):
    mask = torch.zeros(len(nodes), dtype=torch.bool)
    for d in range(nodes.shape[1]):
        c = nodes[:, d]
        mask |= torch.isclose(c, c.min()) | torch.isclose(c, c.max())
    return mask


@pytest.mark.parametrize("name", ["Quad1", "Tetra1", "Hexa1", "Hexa2"])
def test_patch_test(name):
    """tests/test_models.py:54-71: a linear field is reproduced exactly."""
    if name == "Quad1":
        nodes, el = O.rect_quad(4, 4)
        m = O.OracleModel(nodes, el, O.Quad1, "planar", O.plane_stress_C(torch.full((len(el),), 1000.0), 0.3),
                          thickness=torch.ones(len(el), dtype=DT))
        Gm = torch.tensor([[1.0e-3, 2.0e-4], [3.0e-4, -1.0e-3]], dtype=DT)
    else:
        if name == "Tetra1":
            nodes, el = O.cube_tetra6(2)
        elif name == "Hexa1":
            nodes, el = O.cube_hexa(3, 3, 3)
        else:
            nodes, el = O.linear_to_quadratic(*O.cube_hexa(3, 3, 3))
        m = O.OracleModel(nodes, el, O.ELEMENTS[name], "solid", O.iso3d_C(torch.full((len(el),), 1000.0), 0.3))
        Gm = torch.tensor([[1.0e-3, 2.0e-4, -1.0e-4], [3.0e-4, -1.0e-3, 5.0e-5], [1.0e-4, 2.0e-4, 7.0e-4]], dtype=DT)
    u_exact = nodes @ Gm.T
    con = _boundary(nodes)[:, None].repeat(1, m.ndof)
    u, f, sigma, _ = m.solve(con, torch.zeros_like(nodes), u_exact)
    assert torch.allclose(u, u_exact, atol=1e-12)
    assert torch.allclose(sigma, sigma[0].expand_as(sigma), atol=1e-12)


def test_solve_matches_dense_solution():
    """tests/test_sparse.py:49-56: the sparse solve agrees with a dense one."""
    nodes, el = O.cube_hexa(3, 3, 3)
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1000.0), 0.3))
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 0] == 0.0] = True
    k = m.k0()
    val = m.assemble_matrix(k, torch.nonzero(con.ravel()).ravel())
    K = torch.tensor(m.csr(val).toarray())
    b = torch.randn(m.n_dofs, generator=torch.Generator().manual_seed(0), dtype=DT)
    b[con.ravel()] = 0.0
    x_ref = torch.linalg.solve(K, b)
    x_cg = torch.as_tensor(m.solve_linear(m.csr(val), b.numpy(), method="cg", rtol=1e-14))
    x_lu = torch.as_tensor(m.solve_linear(m.csr(val), b.numpy(), method="spsolve"))
    assert torch.allclose(x_cg, x_ref, atol=1e-10)
    assert torch.allclose(x_lu, x_ref, atol=1e-12)


def _readme_cantilever(thickness):
    nodes = torch.tensor([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, 1.0]], dtype=DT)
    el = torch.tensor([[0, 1, 4, 3], [1, 2, 5, 4]])
    C = O.plane_stress_C(torch.full((2,), 1000.0, dtype=DT), torch.tensor(0.3, dtype=DT))
    m = O.OracleModel(nodes, el, O.Quad1, "planar", C, thickness=thickness)
    con = torch.zeros(6, 2, dtype=torch.bool)
    con[[0, 3], :] = True
    F = torch.zeros(6, 2, dtype=DT)
    F[5, 1] = -1.0
    return m, con, F


def test_readme_golden_thickness_gradient():
    """README.md:101-113: dC/dthickness = [-0.0208, -0.0053] (4 decimals)."""
    t = torch.ones(2, dtype=DT, requires_grad=True)
    m, con, F = _readme_cantilever(t)
    u, f, _, _ = m.solve(con, F, torch.zeros(6, 2, dtype=DT), params=(t,))
    compliance = torch.inner(f.ravel(), u.ravel())
    (g,) = torch.autograd.grad(compliance, t)
    assert torch.allclose(g, torch.tensor([-0.0208, -0.0053], dtype=DT), atol=5e-5)
    # and central finite differences of the same objective
    h = 1e-6
    for e in range(2):
        vals = []
        for s in (1, -1):
            tt = torch.ones(2, dtype=DT)
            tt[e] += s * h
            m2, _, _ = _readme_cantilever(tt)
            u2, f2, _, _ = m2.solve(con, F, torch.zeros(6, 2, dtype=DT))
            vals.append(float(torch.inner(f2.ravel(), u2.ravel())))
        assert math.isclose(float(g[e]), (vals[0] - vals[1]) / (2 * h), rel_tol=1e-6)


def _simp_cube(n, rho):
    nodes, el = O.cube_hexa(n + 1, n + 1, n + 1)
    C = O.iso3d_C(1000.0 * rho**3, torch.tensor(0.3, dtype=DT))
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", C)
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 0] == 0.0] = True
    F = torch.zeros(len(nodes), 3, dtype=DT)
    right = nodes[:, 0] == 1.0
    F[right, 2] = -1.0 / int(right.sum())
    return m, con, F


def test_simp_compliance_gradient_closed_form_and_fd():
    """Implicit-function gradient vs the self-adjoint closed form
    -3 rho^2 E0 u_e^T k0_e u_e (topology.ipynb cell 8) and central FD."""
    n = 3
    g = torch.Generator().manual_seed(0)
    rho = (0.3 + 0.7 * torch.rand(n**3, generator=g, dtype=DT)).requires_grad_(True)
    m, con, F = _simp_cube(n, rho)
    u, f, _, _ = m.solve(con, F, torch.zeros_like(F), params=(rho,))
    comp = torch.inner(F.ravel(), u.ravel())
    (grad,) = torch.autograd.grad(comp, rho)
    unit = O.OracleModel(m.nodes, m.elements, O.Hexa1, "solid",
                         O.iso3d_C(torch.ones(n**3, dtype=DT), torch.tensor(0.3, dtype=DT)),
                         pattern=m.pattern)
    closed = O.simp_compliance_sensitivity(unit, u.detach(), rho.detach(), 1000.0)
    assert torch.allclose(grad, closed, rtol=1e-8)
    h = 1e-6
    for e in [0, 7, 13, 26]:
        vals = []
        for s in (1, -1):
            r2 = rho.detach().clone()
            r2[e] += s * h
            m2, _, _ = _simp_cube(n, r2)
            u2, _, _, _ = m2.solve(con, F, torch.zeros_like(F))
            vals.append(float(torch.inner(F.ravel(), u2.ravel())))
        assert math.isclose(float(grad[e]), (vals[0] - vals[1]) / (2 * h), rel_tol=1e-5)


def test_surface_load_totals():
    """tests/test_loads.py:118-160: a uniform traction integrates to traction x area."""
    nodes, el = O.linear_to_quadratic(*O.cube_hexa(3, 3, 3))
    m = O.OracleModel(nodes, el, O.Hexa2, "solid", O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    top = nodes[:, 2] == 1.0
    load, conn = m.integrate_surface_load(top, torch.tensor([0.3, -0.2, 1.0], dtype=DT))
    assert conn.shape == (4, 8)
    assert torch.allclose(load.sum(0), torch.tensor([0.3, -0.2, 1.0], dtype=DT), atol=1e-13)
    assert (load[~top] == 0).all()


# ------------------------------------------------------------ hyperelasticity
def _nh_P_closed(F, lam, mu):
    Fit = torch.linalg.inv(F).transpose(-1, -2)
    lnJ = torch.log(torch.linalg.det(F))[..., None, None]
    return mu * (F - Fit) + lam * lnJ * Fit


def test_neo_hookean_stress_and_tangent():
    """hyper_step's automatic derivatives of the Neo-Hookean energy against
    P = mu (F - F^-T) + lam ln J F^-T, the tangent against central
    differences of that closed form, major symmetry, and the small-strain
    limit: at F = I the tangent is the isotropic C with G = mu."""
    g = torch.Generator().manual_seed(0)
    F = torch.eye(3, dtype=DT) + 0.2 * torch.randn(5, 3, 3, generator=g, dtype=DT)
    lam, mu = 800.0, 350.0
    params = torch.tensor([mu, lam], dtype=DT)
    P, A = O.hyper_step(O.neo_hookean_psi, F, params)
    assert torch.allclose(P, _nh_P_closed(F, lam, mu), rtol=1e-12, atol=1e-10)
    h = 1e-6
    for k in range(3):
        for L_ in range(3):
            dF = torch.zeros(3, 3, dtype=DT)
            dF[k, L_] = h
            fd = (_nh_P_closed(F + dF, lam, mu) - _nh_P_closed(F - dF, lam, mu)) / (2 * h)
            assert torch.allclose(A[..., k, L_], fd, rtol=1e-6, atol=1e-5)
    assert torch.allclose(A, A.permute(0, 3, 4, 1, 2), atol=1e-9)
    _, A0 = O.hyper_step(O.neo_hookean_psi, torch.eye(3, dtype=DT)[None], params)
    E = mu * (3 * lam + 2 * mu) / (lam + mu)
    nu = lam / (2 * (lam + mu))
    C = O.iso3d_C(torch.tensor([E], dtype=DT), torch.tensor(nu, dtype=DT))
    assert torch.allclose(A0, C, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("nlgeom", [False, True])
def test_hyperelastic_affine_deformation_is_exact(nlgeom):
    """An affine large deformation x = F0 X prescribed on the whole boundary
    of a distorted Hexa1 mesh is reproduced inside (the elements contain
    linear fields and a uniform P is in equilibrium), in two increments; the
    stress is P(F0) (nlgeom: Cauchy F0 P / det F0) in every element."""
    nodes, el = O.cube_hexa(4, 4, 4)
    inner = ((nodes > 0) & (nodes < 1)).all(dim=1)
    nodes = nodes + 0.05 * torch.sin(4.0 * nodes[:, [1, 2, 0]]) * inner[:, None]
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    F0 = torch.tensor([[1.25, 0.10, 0.0], [0.05, 0.90, 0.08], [0.0, -0.06, 1.10]], dtype=DT)
    u_exact = nodes @ (F0 - torch.eye(3, dtype=DT)).T
    bnd = _boundary(nodes)
    con = bnd[:, None].repeat(1, 3)
    params = torch.tensor([350.0, 800.0], dtype=DT).expand(len(el), 2)
    u, f, flux, grad = O.oracle_solve_hyper(m, O.neo_hookean_psi, params, con,
                                            torch.zeros_like(nodes), u_exact,
                                            torch.tensor([0.0, 0.5, 1.0]), nlgeom=nlgeom)
    assert torch.allclose(u, u_exact, atol=1e-12)
    assert torch.allclose(grad, F0.expand_as(grad), atol=1e-12)
    P0 = _nh_P_closed(F0, 800.0, 350.0)
    expect = F0 @ P0 / torch.linalg.det(F0) if nlgeom else P0
    assert torch.allclose(flux, expect.expand_as(flux), rtol=1e-11, atol=1e-9)
    assert f[~bnd].abs().max() < 1e-9


def test_hyperelastic_small_load_matches_linear_elasticity():
    """For a small load the Neo-Hookean solution approaches the linear
    elastic one with the same Lame constants (difference second order)."""
    nodes, el = O.cube_hexa(3, 3, 3)
    lam, mu = 800.0, 350.0
    E = mu * (3 * lam + 2 * mu) / (lam + mu)
    nu = lam / (2 * (lam + mu))
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), E), nu))
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 2] == 0.0] = True
    F = torch.zeros(len(nodes), 3, dtype=DT)
    F[nodes[:, 2] == 1.0, 0] = 1e-3
    params = torch.tensor([mu, lam], dtype=DT).expand(len(el), 2)
    u_lin, _, _, _ = m.solve(con, F, torch.zeros_like(F))
    errs = []
    for s in (1.0, 0.1):
        u_nl, _, _, _ = O.oracle_solve_hyper(m, O.neo_hookean_psi, params, con, s * F,
                                             torch.zeros_like(F), torch.tensor([0.0, 1.0]),
                                             nlgeom=True)
        errs.append(float((u_nl - s * u_lin).norm() / (s * u_lin).norm()))
    assert errs[0] < 1e-2
    assert errs[1] < 0.2 * errs[0]  # the nonlinear part shrinks with the load


def uniaxial_stretch_closed_form(mu, lbd, stretch):
    """tests/test_gradients.py:114-125: the lateral stretch of a uniaxially
    stretched Neo-Hookean block from mu (J / lam - 1) + lbd ln J = 0 (Newton),
    the nominal reaction R = mu (lam - 1/lam) + lbd ln J / lam and its
    derivatives with respect to (mu, lbd) by implicit differentiation."""
    lam = stretch
    J = 1.0
    for _ in range(100):
        J -= (mu * (J / lam - 1.0) + lbd * math.log(J)) / (mu / lam + lbd / J)
    den = mu / lam + lbd / J
    dJ_dmu = -(J / lam - 1.0) / den
    dJ_dlbd = -math.log(J) / den
    R = mu * (lam - 1.0 / lam) + lbd * math.log(J) / lam
    dR = (lam - 1.0 / lam) + lbd / (lam * J) * dJ_dmu, math.log(J) / lam + lbd / (lam * J) * dJ_dlbd
    return R, dR


def test_hyperelastic_uniaxial_stretch_reaction():
    """The reference's known answer (tests/test_gradients.py:80-128): a unit
    cube of Neo-Hookean material stretched to 3x its length (8 geometric
    increments, nlgeom) carries the analytical nominal reaction force."""
    E, nu, U = 1000.0, 0.3, 2.0
    lbd = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    nodes, el = O.cube_hexa(3, 3, 3)
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    right = nodes[:, 0] == 1.0
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 0] == 0.0, 0] = True
    con[right, 0] = True
    con[nodes[:, 1] == 0.5, 1] = True
    con[nodes[:, 2] == 0.5, 2] = True
    D = torch.zeros(len(nodes), 3, dtype=DT)
    D[right, 0] = U
    lam_pts = torch.logspace(0, math.log10(1.0 + U), 8, dtype=DT)
    params = torch.tensor([mu, lbd], dtype=DT).expand(len(el), 2)
    u, f, _, _ = O.oracle_solve_hyper(m, O.neo_hookean_psi, params, con, torch.zeros_like(D), D,
                                      (lam_pts - 1.0) / U, nlgeom=True)
    R, _ = uniaxial_stretch_closed_form(mu, lbd, 1.0 + U)
    assert abs(float(f[right, 0].sum()) - R) < 1e-8 * R


# --------------------------------------------------------------- plasticity
def _sf_lin(q):
    return 200.0 + 50.0 * q


def _sfp_lin(q):
    return 50.0 + 0.0 * q


def _sf_sat(q):  # saturating (nonlinear) hardening
    return 200.0 + 150.0 * (1.0 - torch.exp(-20.0 * q))


def _sfp_sat(q):
    return 3000.0 * torch.exp(-20.0 * q)


def test_j2_step_elastic_and_return_mapping():
    """The restated J2 step: a small strain stays elastic (stress C : eps,
    state unchanged, the reference's test_materials.py:229-246); a large
    deviatoric strain returns to the yield surface, |dev sigma| =
    sqrt(2/3) sigma_f(q), along the trial direction, with q > 0; the
    hydrostatic part stays elastic (3 K tr eps / 3)."""
    E, nu = 210e3, 0.3
    step = O.j2_step(E, nu, _sf_lin, _sfp_lin, tol=1e-10)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    H = torch.zeros(2, 3, 3, dtype=DT)
    H[0] = 1e-6 * torch.eye(3, dtype=DT)
    H[1] = torch.tensor([[0.01, 0.002, 0.0], [0.0, -0.004, 0.001], [0.0, 0.0, -0.003]], dtype=DT)
    P, q, A = step(H, torch.eye(3, dtype=DT).expand(2, 3, 3), torch.zeros(2, 3, 3, dtype=DT),
                   torch.zeros(2, 1, dtype=DT))
    eps = 0.5 * (H + H.transpose(-1, -2))
    C = O.iso3d_C(torch.tensor([E], dtype=DT), torch.tensor(nu, dtype=DT))[0]
    assert torch.allclose(P[0], torch.einsum("ijkl,kl->ij", C, eps[0]), rtol=1e-12)
    assert float(q[0]) == 0.0
    I = torch.eye(3, dtype=DT)
    dev = P[1] - torch.trace(P[1]) / 3 * I
    assert abs(float(dev.norm()) - math.sqrt(2 / 3) * float(_sf_lin(q[1]))) < 1e-8
    trial = torch.einsum("ijkl,kl->ij", C, eps[1])
    dev_t = trial - torch.trace(trial) / 3 * I
    assert torch.allclose(dev / dev.norm(), dev_t / dev_t.norm(), atol=1e-12)
    assert float(q[1]) > 0.0
    K = lam + 2 * G / 3
    assert abs(float(torch.trace(P[1])) - 3 * K * float(torch.trace(eps[1]))) < 1e-6


@pytest.mark.parametrize("hard", ["linear", "saturating"])
def test_j2_consistent_tangent_matches_differences(hard):
    """The restated consistent tangent against central differences of the
    step's stress in the strain, from a pre-stressed, pre-strained state, at
    an elastic and at a yielding point."""
    sf, sfp = (_sf_lin, _sfp_lin) if hard == "linear" else (_sf_sat, _sfp_sat)
    step = O.j2_step(1000.0, 0.3, sf, sfp, tol=1e-12, max_iter=50)
    base = torch.tensor([[0.1, 0.03, 0.0], [0.02, -0.05, 0.01], [0.0, 0.0, -0.04]], dtype=DT)
    S0 = torch.tensor([[50.0, 10.0, 0.0], [10.0, -20.0, 5.0], [0.0, 5.0, 0.0]], dtype=DT)
    q0 = torch.tensor([[0.01]], dtype=DT)
    for scale in (0.01, 5.0):  # elastic, plastic
        H = scale * base
        P, q, A = step(H[None], torch.eye(3, dtype=DT)[None], S0[None], q0)
        assert (float(q[0, 0]) > 0.01) == (scale > 1.0)
        h = 1e-7
        for k in range(3):
            for l in range(3):
                d = torch.zeros(3, 3, dtype=DT)
                d[k, l] = h
                Pp, _, _ = step((H + d)[None], torch.eye(3, dtype=DT)[None], S0[None], q0)
                Pm, _, _ = step((H - d)[None], torch.eye(3, dtype=DT)[None], S0[None], q0)
                fd = (Pp - Pm)[0] / (2 * h)
                # A is the derivative w.r.t. the symmetric strain: dP/dH_kl
                # = (A[..., k, l] + A[..., l, k]) / 2
                an = 0.5 * (A[0, :, :, k, l] + A[0, :, :, l, k])
                assert torch.allclose(fd, an, rtol=1e-5, atol=1e-6 * float(A.abs().max()))


def uniaxial_plastic_closed_form(E, nu, s0, Hh, strain):
    """Uniaxial stress in a J2 solid with sigma_f(q) = s0 + Hh q: sigma =
    E strain below s0 / E, else (s0 + Hh strain) / (1 + Hh / E); the
    lateral strain -nu sigma / E - eps_p / 2 (plastic incompressibility)."""
    if E * strain <= s0:
        sig, ep = E * strain, 0.0
    else:
        sig = (s0 + Hh * strain) / (1.0 + Hh / E)
        ep = (sig - s0) / Hh
    return sig, -nu * sig / E - 0.5 * ep, ep


def _plastic_box(n, s0, Hh, strain):
    nodes, el = O.cube_hexa(n, n, n)
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1.0), 0.3))
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 0] == 0.0, 0] = True
    con[nodes[:, 1] == 0.0, 1] = True
    con[nodes[:, 2] == 0.0, 2] = True
    right = nodes[:, 0] == 1.0
    con[right, 0] = True
    D = torch.zeros(len(nodes), 3, dtype=DT)
    D[right, 0] = strain
    return nodes, m, con, D, right


def test_j2_uniaxial_tension_closed_form():
    """A J2 cube with linear hardening pulled to three times its yield
    strain in four increments (roller supports on three faces): the
    reaction is the closed-form uniaxial stress, the lateral contraction its
    elastic plus plastically incompressible part, q the plastic strain."""
    E, nu, s0, Hh = 1000.0, 0.3, 2.0, 100.0
    strain = 3.0 * s0 / E
    nodes, m, con, D, right = _plastic_box(3, s0, Hh, strain)
    step = O.j2_step(E, nu, lambda q: s0 + Hh * q, lambda q: Hh + 0.0 * q, tol=1e-12)
    u, f, flux, grad, state = O.oracle_solve_incremental(
        m, step, 1, con, torch.zeros_like(D), D, torch.tensor([0.0, 0.25, 0.5, 0.75, 1.0]),
        nlgeom=False, rtol=1e-12, atol=1e-12)
    sig, lat, ep = uniaxial_plastic_closed_form(E, nu, s0, Hh, strain)
    assert abs(float(f[right, 0].sum()) - sig) < 1e-9 * sig
    assert torch.allclose(u[:, 1], lat * nodes[:, 1], atol=1e-12)
    assert torch.allclose(flux[:, 0, 0], torch.full_like(flux[:, 0, 0], sig), rtol=1e-9)
    assert torch.allclose(state[:, 0], torch.full_like(state[:, 0], ep), rtol=1e-9)


# ------------------------------------------------------- time integration
def test_mass_matrix_partition_of_unity_and_symmetry():
    """oracle_mass: symmetric, positive, and (N_I summing to one) its entries
    sum to rho x volume per element and DOF component."""
    nodes, el = O.cube_hexa(3, 4, 3)
    inner = ((nodes > 0) & (nodes < 1)).all(dim=1)
    nodes = nodes + 0.04 * torch.sin(4.0 * nodes[:, [1, 2, 0]]) * inner[:, None]
    for kind, mat in (("heat", O.iso_KAPPA(torch.ones(len(el), dtype=DT), 3)),
                      ("solid", O.iso3d_C(torch.ones(len(el), dtype=DT), 0.3))):
        m = O.OracleModel(nodes, el, O.Hexa1, kind, mat)
        rho = 1.0 + torch.rand(len(el), generator=torch.Generator().manual_seed(0), dtype=DT)
        M = O.oracle_mass(m, rho)
        _, _, detJ = m.eval_shape_functions(O.Hexa1.ipoints)
        V = (detJ * O.Hexa1.iweights[:, None]).sum(0)
        assert torch.allclose(M, M.transpose(1, 2), atol=1e-15)
        assert torch.allclose(M.sum((1, 2)), m.ndof * rho * V, rtol=1e-13)
        assert bool((torch.linalg.eigvalsh(M) > 0).all())


def test_time_integration_decays_an_eigenmode_exactly():
    """oracle_time_integration on a generalized eigenvector K phi = lam M phi
    (zero boundary temperature): the first step, whose old internal vector
    is zero as in the reference, gives phi / (1 + dt lam / 2); every later
    (Crank-Nicolson) step multiplies by (1 - dt lam / 2) / (1 + dt lam / 2)."""
    import scipy.linalg as sla

    nodes, el = O.cube_hexa(5, 5, 5)
    kappa, rho = 2.0, 3.0
    m = O.OracleModel(nodes, el, O.Hexa1, "heat", O.iso_KAPPA(torch.full((len(el),), kappa,
                                                                        dtype=DT), 3))
    bnd = _boundary(nodes)
    free = torch.nonzero(~bnd).ravel()
    K = m.csr(m.assemble_matrix(m.k0(), torch.zeros(0, dtype=torch.int64))).toarray()
    M = m.csr(m.assemble_matrix(O.oracle_mass(m, torch.full((len(el),), rho, dtype=DT)),
                                torch.zeros(0, dtype=torch.int64))).toarray()
    f = free.numpy()
    lam, vec = sla.eigh(K[np.ix_(f, f)], M[np.ix_(f, f)])
    phi = torch.zeros(len(nodes), dtype=DT)
    phi[free] = torch.as_tensor(vec[:, 1])
    dt = 0.01
    T = O.oracle_time_integration(m, torch.full((len(el),), rho, dtype=DT), bnd[:, None],
                                  phi[:, None], torch.zeros(len(nodes), 1, dtype=DT),
                                  torch.tensor([0.0, 0.03, 0.06]), dt, rtol=1e-13, atol=1e-14)
    a = 0.5 * dt * lam[1]
    g = (1 - a) / (1 + a)
    for row, n in ((0, 0), (1, 3), (2, 6)):
        amp = 1.0 if n == 0 else g ** (n - 1) / (1 + a)
        assert torch.allclose(T[row], amp * phi, atol=1e-12 * float(phi.abs().max()))


# ------------------------------------------------ planar hyperelasticity
def plane_strain_uniaxial_closed_form(mu, lbd, lam):
    """Neo-Hookean plate in plane strain stretched by lam along x, free
    along y: F = diag(lam, s, 1) with P_yy = mu (s - 1/s) + lbd ln(lam s) / s
    = 0 (scalar Newton); the nominal stress P_xx = mu (lam - 1/lam)
    + lbd ln(lam s) / lam."""
    s = 1.0
    for _ in range(100):
        r = mu * (s - 1.0 / s) + lbd * math.log(lam * s) / s
        dr = mu * (1.0 + 1.0 / s**2) + lbd * (1.0 - math.log(lam * s)) / s**2
        s -= r / dr
    return mu * (lam - 1.0 / lam) + lbd * math.log(lam * s) / lam, s


def _stretch_plate(n=4, t=0.3):
    nodes, el = O.rect_quad(n, n, 1.0, 1.0)
    m = O.OracleModel(nodes, el, O.Quad1, "planar",
                      O.plane_stress_C(torch.full((len(el),), 1.0), torch.tensor(0.3)),
                      thickness=torch.full((len(el),), t, dtype=DT))
    con = torch.zeros(len(nodes), 2, dtype=torch.bool)
    con[nodes[:, 0] == 0.0, 0] = True
    con[nodes[:, 1] == 0.0, 1] = True
    right = nodes[:, 0] == 1.0
    con[right, 0] = True
    return nodes, el, m, con, right


@pytest.mark.parametrize("kind", ["stress", "strain"])
def test_planar_hyperelastic_uniaxial_stretch(kind):
    """A Neo-Hookean plate stretched to 1.8x in four increments: in plane
    stress the lateral and out-of-plane stretches are equal and the reaction
    is the 3-D uniaxial closed form; in plane strain the lateral stretch
    solves P_yy = 0 with F_zz = 1. Reactions per unit reference width times
    the thickness."""
    mu, lbd, t = 350.0, 800.0, 0.3
    nodes, el, m, con, right = _stretch_plate(t=t)
    D = torch.zeros(len(nodes), 2, dtype=DT)
    D[right, 0] = 0.8
    params = torch.tensor([mu, lbd], dtype=DT).expand(len(el), 2)
    step = (O.hyper_plane_stress(O.neo_hookean_psi, params) if kind == "stress"
            else O.hyper_plane_strain(O.neo_hookean_psi, params))
    u, f, _, _, state = O.oracle_solve_incremental(
        m, step, 1 if kind == "stress" else 0, con, torch.zeros_like(D), D,
        torch.tensor([0.0, 0.25, 0.5, 0.75, 1.0]), nlgeom=True)
    if kind == "stress":
        R, _ = uniaxial_stretch_closed_form(mu, lbd, 1.8)
        lat = None
    else:
        R, lat = plane_strain_uniaxial_closed_form(mu, lbd, 1.8)
    assert abs(float(f[right, 0].sum()) - R * t) < 1e-9 * R * t
    if lat is not None:
        assert torch.allclose(u[:, 1], (lat - 1.0) * nodes[:, 1], atol=1e-11)
    else:
        # out-of-plane stretch equals the lateral one (isotropic material)
        s_lat = 1.0 + float(u[nodes[:, 1] == 1.0, 1].mean())
        assert abs(float(state.mean()) + 1.0 - s_lat) < 1e-10
