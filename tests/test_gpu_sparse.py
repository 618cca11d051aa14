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


@pytest.mark.parametrize("build,method", [(_spd, None), (_spd, "cg"), (_nonsymmetric, None)])
def test_adjoint_gradients_match_dense(build, method):
    A_dense, b, w = build(N, 0), torch.randn(N, dtype=torch.float64), torch.randn(N, dtype=torch.float64)
    x, gA, gb = _sparse_grads(A_dense, b, w, method)
    xd, gAd, gbd = _dense_grads(A_dense, b, w)
    tol = 1e-10 if method is None else 1e-8
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
    assert resolve_method(10000, "cuda", None) == "cg"
    assert "jacobi" in describe_method(20000, "cuda", None)
