"""Implicit time integration on the GPU (FEM.time_integration with the
HIP mass matrices and the (M + dt/2 K) PCG): the reference's own tests of
tests/test_time_integration.py on the same heated plate, parity with the
oracle's restatement of the time loop, and the exact decay of a
generalized eigenmode.
"""

import numpy as np
import pytest
import scipy.linalg as sla
import torch

import oracle as O
from torchfem import PlanarHeat, SolidHeat
from torchfem.materials import IsotropicConductivity2D, IsotropicConductivity3D
from torchfem.mesh import rect_quad

pytestmark = pytest.mark.gpu
DT = torch.float64


def _plate() -> PlanarHeat:
    """tests/test_time_integration.py:9-17."""
    model = PlanarHeat(*(t.cuda() for t in rect_quad(5, 5, 1.0, 1.0)),
                       IsotropicConductivity2D(kappa=400.0, rho=1.0e5))
    west = torch.isclose(model.nodes[:, 0], model.nodes[:, 0].min())
    east = torch.isclose(model.nodes[:, 0], model.nodes[:, 0].max())
    model.constraints[west | east] = True
    model.temperatures[west, 0] = 5.0
    model.temperatures[east, 0] = 20.0
    return model


def test_one_result_per_output_time():
    model = _plate()
    t_output = torch.tensor([0.0, 6.0, 12.0, 18.0, 24.0])
    out = model.time_integration(t_output, delta_t=1.0)
    for result in out:
        assert result.shape[0] == len(t_output)


def test_single_output_time_keeps_time_axis():
    model = _plate()
    temp, _, flux, grad, _ = model.time_integration(torch.tensor([10.0]), delta_t=1.0)
    assert temp.shape[0] == 1 and flux.shape[0] == 1 and grad.shape[0] == 1


def test_sparse_output_matches_dense_output():
    model = _plate()
    dense_times = torch.arange(0.0, 10.0 + 0.1, 0.1)
    dense, *_ = model.time_integration(dense_times, 0.1)
    sparse_times = torch.tensor([0.0, 2.5, 5.0, 10.0])
    sparse, *_ = model.time_integration(sparse_times, 0.1)
    rows = torch.searchsorted(dense_times.contiguous(), sparse_times)
    assert torch.allclose(sparse, dense[rows], atol=1e-9)


def test_delta_t_controls_accuracy():
    model = _plate()
    t_output = torch.tensor([0.0, 1.0])
    reference, *_ = model.time_integration(t_output, delta_t=0.005)
    coarse, *_ = model.time_integration(t_output, delta_t=99.0)
    fine, *_ = model.time_integration(t_output, delta_t=0.05)
    assert (fine[-1] - reference[-1]).abs().max() < (coarse[-1] - reference[-1]).abs().max()


def test_starts_from_equilibrium_when_output_starts_late():
    model = _plate()
    full, *_ = model.time_integration(torch.arange(0.0, 10.0 + 0.5, 0.5), 0.5)
    late, *_ = model.time_integration(torch.tensor([5.0, 10.0]), 0.5)
    assert torch.allclose(late, full[torch.tensor([10, 20])], atol=1e-9)


@pytest.mark.parametrize("t_output", [torch.tensor([]), torch.tensor([-1.0, 1.0]),
                                      torch.tensor([1.0, 0.5]), torch.tensor([1.0, 1.0])])
def test_rejects_invalid_output_times(t_output):
    with pytest.raises(ValueError):
        _plate().time_integration(t_output, delta_t=0.5)


def test_gradients_are_refused():
    model = _plate()
    kappa = torch.tensor(400.0, dtype=DT, requires_grad=True)
    with pytest.raises(NotImplementedError):
        model.time_integration(torch.tensor([0.0, 1.0]), 0.5, differentiable_parameters=kappa)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_plate_matches_oracle(precision):
    """Temperatures, and the internal vector's output, of the heated plate
    against the oracle's restatement of the time loop."""
    model = _plate()
    t_output = torch.tensor([0.0, 0.7, 2.0, 5.0])
    T, f, flux, grad, _ = model.time_integration(t_output, delta_t=0.25, rtol=1e-12, atol=1e-12,
                                                 stol=1e-12, precision=precision)
    nodes, el = rect_quad(5, 5, 1.0, 1.0)
    m = O.OracleModel(nodes, el, O.Quad1, "heat", O.iso_KAPPA(torch.full((len(el),), 400.0), 2),
                      thickness=torch.ones(len(el), dtype=DT))
    T_o = O.oracle_time_integration(m, torch.full((len(el),), 1.0e5, dtype=DT),
                                    model.constraints.cpu(), model.temperatures.cpu(),
                                    model.heat_flux.cpu(), t_output, 0.25, rtol=1e-12, atol=1e-12)
    assert float((T.cpu().reshape(len(t_output), -1) - T_o).abs().max()) < 1e-9 * 20.0
    # the flux is KAPPA grad T (conductivity.py sign convention) per element
    assert torch.allclose(flux[-1], 400.0 * grad[-1], rtol=1e-12, atol=1e-9)


def test_mass_matrix_matches_oracle():
    nodes, el = O.cube_tetra6(3, jitter=0.15, seed=2)
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(1.0, rho=2.5))
    m = O.OracleModel(nodes, el, O.Tetra1, "heat", O.iso_KAPPA(torch.ones(len(el)), 3))
    M = model.integrate_mass().cpu()
    M_o = O.oracle_mass(m, torch.full((len(el),), 2.5, dtype=DT))
    assert float((M - M_o).abs().max()) < 1e-14 * float(M_o.abs().max())


def test_eigenmode_decays_exactly():
    """A generalized eigenvector of (K, M) with zero boundary temperature:
    the first step gives phi / (1 + dt lam / 2) (its old internal vector is
    zero, as in the reference), later steps multiply by
    (1 - dt lam / 2) / (1 + dt lam / 2); 9^3-cell Hexa1 cube."""
    nodes, el = O.cube_hexa(10, 10, 10)
    kappa, rho = 2.0, 3.0
    model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(kappa, rho=rho))
    bnd = ((nodes == 0) | (nodes == 1)).any(dim=1)
    m = O.OracleModel(nodes, el, O.Hexa1, "heat",
                      O.iso_KAPPA(torch.full((len(el),), kappa, dtype=DT), 3))
    none = torch.zeros(0, dtype=torch.int64)
    K = m.csr(m.assemble_matrix(m.k0(), none)).toarray()
    M = m.csr(m.assemble_matrix(O.oracle_mass(m, torch.full((len(el),), rho, dtype=DT)),
                                none)).toarray()
    f = torch.nonzero(~bnd).ravel().numpy()
    lam, vec = sla.eigh(K[np.ix_(f, f)], M[np.ix_(f, f)], subset_by_index=[0, 2])
    phi = torch.zeros(len(nodes), dtype=DT)
    phi[torch.as_tensor(f)] = torch.as_tensor(vec[:, 2])
    model.constraints[bnd.cuda()] = True
    model.temperatures = phi[:, None].cuda()
    dt = 0.002
    T, *_ = model.time_integration(torch.tensor([0.0, 0.01, 0.02]), dt, rtol=1e-13, atol=1e-14,
                                   stol=1e-13, precision="fp64")
    a = 0.5 * dt * lam[2]
    g = (1 - a) / (1 + a)
    for row, n in ((1, 5), (2, 10)):
        amp = g ** (n - 1) / (1 + a)
        assert float((T[row, :, 0].cpu() - amp * phi).abs().max()) < 1e-10 * float(phi.abs().max())
