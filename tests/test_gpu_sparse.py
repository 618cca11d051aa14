"""torchfem.sparse on the GPU backend, following the reference's
tests/test_sparse.py:23-245 (dense parity of solves and adjoint gradients)."""

import inspect

import pytest
import torch

from torchfem.sparse import (CachedSolve, describe_method, differentiable_sparse_solve,
                             newton_solve, resolve_method, sparse_solve)

pytestmark = pytest.mark.gpu
N = 6


def _spd(n, seed):
    torch.manual_seed(seed)
    A = torch.randn(n, n, dtype=torch.float64)
    return A @ A.T + n * torch.eye(n, dtype=torch.float64)


def _nonsymmetric(n, seed):
    torch.manual_seed(seed)
    return torch.randn(n, n, dtype=torch.float64) + n * torch.eye(n, dtype=torch.float64)


def _to_sparse(dense, values=None):
    idx = torch.cartesian_prod(torch.arange(dense.shape[0]), torch.arange(dense.shape[1])).T
    vals = dense.flatten() if values is None else values
    return torch.sparse_coo_tensor(idx, vals, dense.shape).coalesce(), idx


@pytest.mark.parametrize("method", [None, "spsolve", "cg", "minres"])
def test_matches_dense_solution(method):
    A_dense = _spd(N, 0)
    b = torch.randn(N, dtype=torch.float64)
    A, _ = _to_sparse(A_dense)
    x, _ = sparse_solve(A.cuda(), b.cuda(), method=method)
    assert torch.allclose(x.cpu(), torch.linalg.solve(A_dense, b), atol=1e-8)


@pytest.mark.parametrize("method", ["cg", "minres"])
def test_iterative_method_returns_a_preconditioner(method):
    A, _ = _to_sparse(_spd(N, 0))
    _, M = sparse_solve(A.cuda(), torch.randn(N, dtype=torch.float64).cuda(), method=method)
    assert M is not None


def test_direct_method_returns_no_preconditioner():
    A, _ = _to_sparse(_spd(N, 0))
    _, M = sparse_solve(A.cuda(), torch.randn(N, dtype=torch.float64).cuda(), method="spsolve")
    assert M is None


def test_rejects_non_square_matrix():
    A = torch.sparse_coo_tensor(torch.tensor([[0, 1], [0, 1]]), torch.ones(2, dtype=torch.float64),
                                (2, 3)).coalesce()
    with pytest.raises(ValueError, match="square 2D matrix"):
        sparse_solve(A.cuda(), torch.ones(2, dtype=torch.float64).cuda())


def test_rejects_unknown_method():
    A, _ = _to_sparse(_spd(N, 0))
    with pytest.raises(ValueError, match="is not supported"):
        sparse_solve(A.cuda(), torch.randn(N, dtype=torch.float64).cuda(), method="not-a-solver")


def test_initial_guess_does_not_change_the_solution():
    A_dense = _spd(N, 0)
    b = torch.randn(N, dtype=torch.float64)
    A, _ = _to_sparse(A_dense)
    ref = torch.linalg.solve(A_dense, b)
    x, _ = sparse_solve(A.cuda(), b.cuda(), method="cg", x0=(ref + 0.1 * torch.randn(N, dtype=torch.float64)).cuda())
    assert torch.allclose(x.cpu(), ref, atol=1e-8)


def _sparse_grads(A_dense, b, w, method=None):
    values = A_dense.flatten().clone().cuda().requires_grad_(True)
    idx = torch.cartesian_prod(torch.arange(A_dense.shape[0]), torch.arange(A_dense.shape[1])).T.cuda()
    A = torch.sparse_coo_tensor(idx, values, A_dense.shape).coalesce()
    bg = b.clone().cuda().requires_grad_(True)
    x = differentiable_sparse_solve(A, bg, method=method)
    gv, gb = torch.autograd.grad((x * w.cuda()).sum(), [values, bg])
    gA = torch.sparse_coo_tensor(idx.cpu(), gv.cpu(), A_dense.shape).to_dense()
    return x.cpu(), gA, gb.cpu()


def _dense_grads(A_dense, b, w):
    A = A_dense.clone().requires_grad_(True)
    bg = b.clone().requires_grad_(True)
    x = torch.linalg.solve(A, bg)
    gA, gb = torch.autograd.grad((x * w).sum(), [A, bg])
    return x, gA, gb


@pytest.mark.parametrize("build,method", [(_spd, None), (_spd, "cg"), (_spd, "minres"),
                                          (_nonsymmetric, None), (_nonsymmetric, "spsolve")])
def test_adjoint_gradients_match_dense(build, method):
    A_dense, b, w = build(N, 0), torch.randn(N, dtype=torch.float64), torch.randn(N, dtype=torch.float64)
    x, gA, gb = _sparse_grads(A_dense, b, w, method)
    xd, gAd, gbd = _dense_grads(A_dense, b, w)
    tol = 1e-10 if method in (None, "spsolve") else 1e-8
    assert torch.allclose(x, xd, atol=tol)
    assert torch.allclose(gb, gbd, atol=tol)
    assert torch.allclose(gA, gAd, atol=tol)
    # dL/db solves the adjoint system A^T lam = dL/dx (pins the transpose)
    assert torch.allclose(gb, torch.linalg.solve(A_dense.T, w), atol=tol)


def test_warm_start_reaches_the_same_solution():
    A_dense = _spd(N, 0)
    b = torch.randn(N, dtype=torch.float64)
    A, _ = _to_sparse(A_dense)
    ref = torch.linalg.solve(A_dense, b)
    cache = CachedSolve()
    first = differentiable_sparse_solve(A.cuda(), b.cuda(), method="cg", cached_solve=cache,
                                        update_cache=True)
    assert cache.previous_x is not None
    second = differentiable_sparse_solve(A.cuda(), b.cuda(), method="cg", cached_solve=cache,
                                         update_cache=True)
    assert torch.allclose(first.cpu(), ref, atol=1e-8)
    assert torch.allclose(second.cpu(), ref, atol=1e-8)


def test_solvers_do_not_share_a_default_cache():
    for func in (differentiable_sparse_solve, newton_solve):
        assert inspect.signature(func).parameters["cached_solve"].default is None


def test_resolve_and_describe():
    assert resolve_method(10, "cuda", "cg") == "cg"
    assert resolve_method(9999, "cuda", None) == "spsolve"
    assert resolve_method(10000, "cuda", None) == "minres"
    assert "jacobi" in describe_method(20000, "cuda", None)


def _nonlinear_problem(n=120, seed=0):
    """R(u; p) = A u + p * u^3 - f with the 1-D Dirichlet Laplacian A (SPD)
    and a positive cubic term, so the tangent A + 3 p diag(u^2) stays SPD."""
    g = torch.Generator().manual_seed(seed)
    A = (2.0 * torch.eye(n) - torch.diag(torch.ones(n - 1), 1) - torch.diag(torch.ones(n - 1), -1))
    A = A.double().cuda()
    f = (1.0 + torch.rand(n, generator=g, dtype=torch.float64)).cuda()
    return A, f


def _residual_fn(A, f, p):
    def eval_residual(du, i, u_prev, grad_prev, flux_prev, state_prev):
        R = A @ du + p * du**3 - f
        K = A + torch.diag(3.0 * p.detach() * du.detach() ** 2)
        return R, K.to_sparse_coo().coalesce()
    return eval_residual


def test_newton_adjoint_gradient_matches_implicit_function_and_differences():
    """sparse.py:517-795: Newton to rtol 1e-12, then d(sum u)/dp through the
    adjoint against -1^T K^-1 u^3 (implicit function) and central differences."""
    A, f = _nonlinear_problem()
    p = torch.tensor(0.5, dtype=torch.float64, device="cuda", requires_grad=True)
    u0 = torch.zeros(A.shape[0], dtype=torch.float64, device="cuda")
    u = newton_solve(_residual_fn(A, f, p), u0, None, 30, 1e-12, 1e-14, 1e-12, None,
                     None, None, None, False, None, None, None, None, p)
    R = A @ u.detach() + 0.5 * u.detach() ** 3 - f
    assert float(R.norm()) < 1e-10 * float(f.norm())
    (g,) = torch.autograd.grad(u.sum(), p)
    K = A + torch.diag(1.5 * u.detach() ** 2)
    implicit = -torch.linalg.solve(K, u.detach() ** 3).sum()
    assert abs(float(g) - float(implicit)) < 1e-9 * abs(float(implicit))
    h = 1e-6
    vals = []
    for s in (1.0, -1.0):
        ps = torch.tensor(0.5 + s * h, dtype=torch.float64, device="cuda")
        us = newton_solve(_residual_fn(A, f, ps), u0, None, 30, 1e-13, 1e-14, 1e-12, None,
                          None, None, None, False, None, None, None, None, ps)
        vals.append(float(us.sum()))
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(fd - float(g)) < 1e-6 * abs(float(g))


def test_newton_raises_without_convergence():
    """sparse.py:622-623: RuntimeError when max_iter runs out."""
    A, f = _nonlinear_problem()
    p = torch.tensor(50.0, dtype=torch.float64, device="cuda")
    u0 = torch.zeros(A.shape[0], dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="did not converge"):
        newton_solve(_residual_fn(A, f, p), u0, None, 2, 1e-12, 1e-14, 1e-12, None,
                     None, None, None, False, None, None, None, None, p)


# ------------------------------------------------ systems of 10^4 unknowns
def _fem_matrix(n=15):
    """The oracle's assembled Hexa1 stiffness of an n^3 cube clamped at z=0
    (Dirichlet rows/columns replaced by the identity): 3 (n+1)^3 unknowns,
    symmetric positive definite."""
    import oracle as O

    nodes, el = O.cube_hexa(n + 1, n + 1, n + 1)
    o = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(torch.full((len(el),), 1000.0), 0.3))
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 2] == 0.0] = True
    cidx = torch.nonzero(con.ravel()).ravel()
    return o.csr(o.assemble_matrix(o.k0(), cidx)).tocsr(), con.ravel()


def _coo(K):
    import numpy as np

    Kc = K.tocoo()
    idx = torch.as_tensor(np.vstack([Kc.row, Kc.col]).astype("int64"))
    return torch.sparse_coo_tensor(idx, torch.as_tensor(Kc.data), K.shape).coalesce().cuda()


def _scipy(K, b):
    import scipy.sparse.linalg as spla

    return torch.as_tensor(spla.spsolve(K.tocsc(), b.numpy()))


@pytest.mark.parametrize("method", [None, "cg", "minres", "spsolve"])
def test_fem_matrix_of_12k_unknowns_matches_scipy(method):
    """An assembled stiffness of 12,288 unknowns (above the 10k switch, so
    method None is MINRES) against SciPy's direct solve."""
    K, con = _fem_matrix()
    assert K.shape[0] == 12288
    b = torch.randn(K.shape[0], generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    b[con] = 0.0
    x, M = sparse_solve(_coo(K), b.cuda(), stol=1e-12, method=method)
    ref = _scipy(K, b)
    assert float((x.cpu() - ref).norm() / ref.norm()) < 1e-7
    assert (M is None) == (method == "spsolve")


def _nonsym_matrix(K, seed=0):
    """K plus a skew part on K's own pattern (a convection-like term): no
    longer symmetric, still invertible (its symmetric part is K)."""
    import scipy.sparse as sp

    Kc = K.tocoo()
    g = torch.Generator().manual_seed(seed)
    up = Kc.row < Kc.col
    r, c = Kc.row[up], Kc.col[up]
    v = (0.2 * torch.rand(len(r), generator=g, dtype=torch.float64) * abs(K.diagonal()).mean()).numpy()
    S = sp.coo_matrix((v, (r, c)), shape=K.shape)
    return (K + S - S.T).tocsr()


@pytest.mark.parametrize("method", [None, "spsolve"])
def test_non_symmetric_system_and_its_transposed_adjoint(method):
    """A non-symmetric matrix of 12k unknowns: the solve (GMRES) and the
    adjoint A^T lam = w of differentiable_sparse_solve against SciPy; method
    None sends it to GMRES instead of MINRES, and 'minres' refuses it."""
    K, con = _fem_matrix()
    A = _nonsym_matrix(K)
    Ac = _coo(A)
    vals = Ac._values().clone().requires_grad_(True)
    Ag = torch.sparse_coo_tensor(Ac._indices(), vals, A.shape).coalesce()
    b = torch.randn(A.shape[0], generator=torch.Generator().manual_seed(4), dtype=torch.float64)
    w = torch.randn(A.shape[0], generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    bg = b.cuda().requires_grad_(True)
    x = differentiable_sparse_solve(Ag, bg, stol=1e-12, method=method)
    assert float((x.detach().cpu() - _scipy(A, b)).norm() / _scipy(A, b).norm()) < 1e-8
    gv, gb = torch.autograd.grad((x * w.cuda()).sum(), [vals, bg])
    lam = _scipy(A.T.tocsr(), w)
    assert float((gb.cpu() - lam).norm() / lam.norm()) < 1e-8
    with pytest.raises(ValueError, match="symmetric"):
        sparse_solve(Ac, b.cuda(), method="minres")


def test_indefinite_matrix_minres_converges_and_cg_fails_fast():
    """K - sigma I with sigma inside the spectrum is symmetric indefinite:
    MINRES solves it, CG stops at the first p.Ap <= 0 with an error instead
    of running to its iteration limit."""
    import scipy.sparse as sp

    K, _ = _fem_matrix(n=10)
    sigma = 0.05 * float(abs(K.diagonal()).mean())
    A = (K - sigma * sp.eye(K.shape[0])).tocsr()
    b = torch.randn(A.shape[0], generator=torch.Generator().manual_seed(6), dtype=torch.float64)
    x, _ = sparse_solve(_coo(A), b.cuda(), stol=1e-10, method="minres")
    ref = _scipy(A, b)
    assert float((x.cpu() - ref).norm() / ref.norm()) < 1e-6
    with pytest.raises(RuntimeError, match="positive definite"):
        sparse_solve(_coo(A), b.cuda(), stol=1e-10, method="cg")
