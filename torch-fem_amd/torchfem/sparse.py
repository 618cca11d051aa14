"""Sparse solves with implicit autograd (src/torchfem/sparse.py), on the GPU.

Same entry points and contracts as the reference: `CachedSolve`
(sparse.py:103-128), the `Solve` autograd Function (131-243) whose backward
solves A^T lam = grad_x and returns dL/dA = -lam[row] x[col] on A's own
pattern, `differentiable_sparse_solve` (246-268), `sparse_solve` (271-347),
`resolve_method`/`describe_method` (74-100) and the generic Newton driver
`NewtonRaphsonAdjoint`/`newton_solve` (517-795).

The backend is libtfm_hip.so, with Jacobi preconditioning throughout, as the
reference's GPU path (sparse.py:406-421):

* "cg": preconditioned CG for symmetric positive definite matrices; it stops
  with an error as soon as an iteration meets p.Ap <= 0;
* "minres": preconditioned MINRES for symmetric (possibly indefinite)
  matrices; a non-symmetric matrix is refused;
* "spsolve" (the default below 10,000 unknowns, sparse.py:74-84): the
  reference factorises; here restarted GMRES runs to a relative residual of
  1e-12, or to the attainable accuracy (at most 1e-10) when FP64 rounding
  stalls it first, as it would a direct solve;
* the default above 10,000 unknowns is "minres", as in the reference; a
  non-symmetric matrix goes to GMRES instead.

There is no CPU path: CPU tensors are moved to the GPU and the result is
returned on the device of b.
"""

from __future__ import annotations

import ctypes as C
import time
import weakref
from collections.abc import Callable
from typing import Any

import torch
from torch import Tensor
from torch.autograd import Function

from . import _lib as L

available_backends = ["hip"]
_DEFAULT_PRECISION = "mixed"
_DIRECT_RTOL = 1e-12   # relative residual "spsolve" aims at
_DIRECT_ACCEPT = 1e-10  # ... and accepts when rounding stalls GMRES above it
_GMRES_RESTART = 300   # basis vectors per GMRES cycle (fewer for small systems)


def default_precision() -> str:
    """Precision of FEM solves when `solve(precision=None)`: "mixed" or "fp64"."""
    return _DEFAULT_PRECISION


def set_default_precision(precision: str) -> None:
    global _DEFAULT_PRECISION
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'")
    _DEFAULT_PRECISION = precision


# FEM solves with preconditioner=None: "auto" takes the two-level
# preconditioner from AUTO_TWO_LEVEL_DOFS unknowns and block-Jacobi below.
# Its coarse set-up (aggregates once per mesh; coarse operator and dense
# inverse, 4 ms on the 824k-DOF cube, per assembled matrix) then pays for
# itself several times over: 2-6x shorter solves on the benchmark meshes of
# 320k-824k DOFs.
_DEFAULT_PRECONDITIONER = "auto"
AUTO_TWO_LEVEL_DOFS = 100_000
PRECONDITIONERS = ("auto", "two_level", "block_jacobi", "jacobi")


def default_preconditioner() -> str:
    """Preconditioner of FEM solves when `solve(preconditioner=None)`."""
    return _DEFAULT_PRECONDITIONER


def set_default_preconditioner(name: str) -> None:
    global _DEFAULT_PRECONDITIONER
    if name not in PRECONDITIONERS:
        raise ValueError(f"preconditioner must be one of {PRECONDITIONERS}")
    _DEFAULT_PRECONDITIONER = name


def resolve_method(n_dofs: int, device: str, method: str | None) -> str:
    """Backend `sparse_solve` uses for a system of this size (sparse.py:74-84):
    "spsolve" below 10,000 DOFs, "minres" above (no AmgX in this build)."""
    if method is not None:
        return method
    return "spsolve" if n_dofs < 10000 else "minres"


def describe_method(n_dofs: int, device: str, method: str | None) -> str:
    resolved = resolve_method(n_dofs, device, method)
    kind = "gmres to 1e-12" if resolved == "spsolve" else "iterative | jacobi"
    return f"{resolved} | {kind} | hip | cuda"


class CachedSolve:
    """Previous solution and adjoint used as warm starts (sparse.py:103-128)."""

    def __init__(self, previous_x: Tensor | None = None, previous_grad: Tensor | None = None):
        self.previous_x = previous_x
        self.previous_grad = previous_grad

    def update_grad(self, grad: Tensor | None) -> None:
        self.previous_grad = grad.detach().clone() if grad is not None else None

    def update_x(self, x: Tensor | None) -> None:
        self.previous_x = x.detach().clone() if x is not None else None


class JacobiPreconditioner:
    """Marker returned by iterative solves, passed on to the adjoint solve
    (the reference returns its preconditioner object, sparse.py:224-243)."""

    kind = "jacobi"


class PCGSolver:
    """Owner of a libtfm_hip PCG workspace (like AmgXSolver, amgx.py:211-395):
    created for n unknowns and up to `nrhs` right-hand sides, released by
    `close()`."""

    def __init__(self, n: int, nrhs: int = 1, device=None, max_iter: int | None = None):
        self.device = L.require_gpu(device)
        self.n = int(n)
        self.nrhs = int(nrhs)
        self.max_iter = int(max_iter or max(100, min(10 * self.n, 20000)))
        self._h = C.c_void_p()
        self._release = None
        self._create()

    def _create(self) -> None:
        with torch.cuda.device(self.device):
            L.check(L.lib().tfm_pcg_create(C.byref(self._h), self.n, self.nrhs, self.max_iter))
        # no __del__ (it may run while the interpreter tears the library down,
        # the reason amgx.py:13-20 gives); a finalizer releases an unclosed
        # workspace at collection or at exit, while the modules are alive
        self._release = weakref.finalize(self, L.lib().tfm_pcg_destroy, C.c_void_p(self._h.value))

    def close(self) -> None:
        """Releases the workspace (idempotent)."""
        if self._release is not None:
            self._release()
            self._release = None
        self._h = C.c_void_p()

    def _ensure(self, nrhs: int) -> None:
        if nrhs > self.nrhs:
            self.close()
            self.nrhs = nrhs
            self._create()

    def timing(self, stride: int) -> tuple[float, int]:
        """Event-timed inner SpMV launches since the last call (ms, count);
        resets the totals, then times the SpMV of every `stride`-th
        iteration from now on (0 switches timing off)."""
        ms = L.F64()
        cnt = L.I64()
        L.check(L.lib().tfm_pcg_timing(self._h, int(stride), C.byref(ms), C.byref(cnt)))
        return float(ms.value), int(cnt.value)

    def solve_system(self, sys: L.System, b: Tensor, x: Tensor, precision="mixed",
                     preconditioner="block_jacobi", rtol=1e-10, inner_rtol=0.1,
                     final_rtol=None, coarse=None) -> dict:
        """K_bc x = b for the rows of b [nrhs, n] (x in/out), on an assembled system.
        Mixed precision: the last FP32 segment aims at final_rtol (default
        0.7 rtol) so that the FP64 residual ends below rtol with a margin."""
        nrhs = b.shape[0]
        self._ensure(nrhs)
        # the coarse space of preconditioner="two_level" (a tfm_coarse handle)
        L.check(L.lib().tfm_pcg_set_coarse(self._h, coarse))
        its = (L.I32 * nrhs)()
        rel = (L.F64 * nrhs)()
        outer = L.I32()
        prec = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
        pc = {"block_jacobi": L.PC_BLOCK_JACOBI, "two_level": L.PC_TWO_LEVEL}.get(
            preconditioner, L.PC_JACOBI)
        t0 = time.perf_counter()
        rc = L.lib().tfm_pcg_solve(self._h, L.stream_of(self.device), C.byref(sys), prec, pc,
                                   L.ptr(b), L.ptr(x), nrhs, float(rtol), float(inner_rtol),
                                   float(0.7 * rtol if final_rtol is None else final_rtol),
                                   its, rel, C.byref(outer))
        fb = L.I32()
        L.check(L.lib().tfm_pcg_fp64_fallback(self._h, C.byref(fb)))
        info = dict(iterations=list(its), relres=list(rel), outer=int(outer.value),
                    time=time.perf_counter() - t0, precision=precision, preconditioner=preconditioner,
                    fp64_fallback=int(fb.value))
        if rc == L.TFM_ERR_NOCONV:
            msg = L.lib().tfm_last_error().decode(errors="replace")
            raise RuntimeError(f"PCG did not converge ({msg}): {info}")
        L.check(rc)
        return info

    def solve_csr(self, rowptr: Tensor, col: Tensor, val: Tensor, b: Tensor, x: Tensor,
                  rtol: float) -> dict:
        its = L.I32()
        rel = L.F64()
        rc = L.lib().tfm_csr_pcg_solve(self._h, L.stream_of(self.device), self.n, L.ptr(rowptr),
                                       L.ptr(col), L.ptr(val), L.ptr(b), L.ptr(x), float(rtol),
                                       C.byref(its), C.byref(rel))
        info = dict(iterations=int(its.value), relres=float(rel.value))
        if rc == L.TFM_ERR_NOCONV:
            msg = L.lib().tfm_last_error().decode(errors="replace")
            raise RuntimeError(f"CG failed after {info['iterations']} iterations: {msg}")
        L.check(rc)
        return info


class KrylovSolver:
    """Owner of a libtfm_hip MINRES / GMRES workspace (tfm_krylov_*), released
    by `close()` like the reference's AmgXSolver (amgx.py:385-395)."""

    def __init__(self, n: int, restart: int, device):
        self.device = device
        self.n = int(n)
        h = L.P()
        L.check(L.lib().tfm_krylov_create(C.byref(h), self.n, int(restart)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().tfm_krylov_destroy(self._h)
            self._h = None

    def solve(self, kind: str, rowptr, col, val, b, x, rtol: float, max_iter: int,
              accept: float | None = None) -> dict:
        """Runs MINRES or GMRES; a stop short of rtol is accepted when the true
        relative residual is at most `accept` (the attainable accuracy)."""
        its = L.I32()
        rel = L.F64()
        fn = L.lib().tfm_csr_minres_solve if kind == "minres" else L.lib().tfm_csr_gmres_solve
        rc = fn(self._h, L.stream_of(self.device), self.n, L.ptr(rowptr), L.ptr(col), L.ptr(val),
                L.ptr(b), L.ptr(x), float(rtol), int(max_iter), C.byref(its), C.byref(rel))
        info = dict(iterations=int(its.value), relres=float(rel.value))
        if rc == L.TFM_ERR_NOCONV:
            if accept is not None and info["relres"] <= accept:
                return info
            msg = L.lib().tfm_last_error().decode(errors="replace")
            raise RuntimeError(f"{kind} failed to converge (relative residual "
                               f"{info['relres']:.3e}): {msg}")
        L.check(rc)
        return info


def is_symmetric(A: Tensor, rtol: float = 1e-12) -> bool:
    """Whether a sparse matrix equals its transpose up to rtol * max|A|."""
    if A.layout == torch.sparse_csr:
        A = A.to_sparse_coo()
    A = A.coalesce()
    vmax = float(A._values().abs().max()) if A._nnz() else 0.0
    D = (A - A.t()).coalesce()
    return D._nnz() == 0 or float(D._values().abs().max()) <= rtol * vmax


def _csr_arrays(A: Tensor, device) -> tuple[Tensor, Tensor, Tensor]:
    """int64 row pointer, int32 columns and FP64 values of a sparse matrix."""
    A = A.to(device)
    n = A.shape[0]
    if A.layout == torch.sparse_csr:
        rowptr = A.crow_indices().to(torch.int64)
        col = A.col_indices().to(torch.int32)
        val = A.values()
    else:
        if not A.is_coalesced():
            A = A.coalesce()
        idx = A._indices()
        rows = idx[0]
        col = idx[1].to(torch.int32)
        val = A._values()
        rowptr = torch.zeros(n + 1, dtype=torch.int64, device=device)
        rowptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
    return rowptr.contiguous(), col.contiguous(), val.to(torch.float64).contiguous()


def sparse_solve(
    A: Tensor,
    b: Tensor,
    B: Tensor | None = None,
    stol: float = 1e-10,
    device: str | None = None,
    method: str | None = None,
    M: Any = None,
    x0: Tensor | None = None,
) -> tuple[Tensor, Any]:
    """Solve A x = b (sparse.py:271-347). Returns (x, preconditioner or None)."""
    if A.ndim != 2 or (A.shape[0] != A.shape[1]):
        raise ValueError("A should be a square 2D matrix.")
    if method is not None and method not in ["spsolve", "minres", "cg", "pardiso", "amgx"]:
        raise ValueError(
            f"Method {method} is not supported. "
            "Choose from 'spsolve', 'minres', 'cg', 'pardiso', or 'amgx'."
        )
    dev = L.require_gpu(device if device is not None else (b.device if b.is_cuda else None))
    out_device = b.device
    n = A.shape[0]
    auto = method is None
    method = resolve_method(n, "cuda", method)
    if method in ("pardiso", "amgx"):
        raise RuntimeError(f"{method} is not available in this build (no CPU or NVIDIA backend).")
    bb = b.detach().to(device=dev, dtype=torch.float64).contiguous()
    Ad = A.detach().to(dev)
    rowptr, col, val = _csr_arrays(Ad, dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev) if x0 is None else (
        x0.detach().to(device=dev, dtype=torch.float64).clone().contiguous())
    if not bool(bb.any()):  # b = 0: x = 0 whatever the initial guess (as scipy's solvers)
        x = torch.zeros(n, dtype=torch.float64, device=dev)
        return x.to(device=out_device, dtype=b.dtype), (None if method == "spsolve" else (
            M if M is not None else JacobiPreconditioner()))
    if method == "minres" and not is_symmetric(Ad):
        if not auto:
            raise ValueError("minres needs a symmetric matrix; use method='spsolve' (GMRES)")
        method = "gmres"  # the default for a large non-symmetric system
    if method == "cg":
        solver = PCGSolver(n, 1, dev)
        try:
            solver.solve_csr(rowptr, col, val, bb, x, stol)
        finally:
            solver.close()
        return x.to(device=out_device, dtype=b.dtype), (M if M is not None else JacobiPreconditioner())
    rtol = _DIRECT_RTOL if method == "spsolve" else stol
    restart = min(n, _GMRES_RESTART)
    ks = KrylovSolver(n, restart, dev)
    try:
        # "spsolve" returns the attainable accuracy when FP64 rounding stops
        # GMRES short of 1e-12 (a direct solve has the same floor)
        ks.solve("minres" if method == "minres" else "gmres", rowptr, col, val, bb, x, rtol,
                 max(1000, 20 * n), accept=_DIRECT_ACCEPT if method == "spsolve" else None)
    finally:
        ks.close()
    if method == "spsolve":
        return x.to(device=out_device, dtype=b.dtype), None
    return x.to(device=out_device, dtype=b.dtype), (M if M is not None else JacobiPreconditioner())


class Solve(Function):
    """x = A^{-1} b with the adjoint backward of sparse.py:131-243."""

    @staticmethod
    def forward(A, b, B=None, stol=1e-10, device=None, method=None, M=None, cached_solve=None,
                update_cache=False):
        x0 = None
        if cached_solve is not None and cached_solve.previous_x is not None:
            x0 = cached_solve.previous_x
        x, M = sparse_solve(A, b, B, stol, device, method, M, x0)
        if update_cache and cached_solve is not None:
            cached_solve.update_x(x)
        return x, M

    @staticmethod
    def setup_context(ctx, inputs, output):
        A, b, B, stol, device, method, M, cached_solve, update_cache = inputs
        x, M_computed = output
        ctx.save_for_backward(A, x)
        ctx.B, ctx.stol, ctx.device, ctx.method = B, stol, device, method
        ctx.M, ctx.cached_solve, ctx.update_cache = M_computed, cached_solve, update_cache

    @staticmethod
    def backward(ctx, *grad_outputs):
        grad_x = grad_outputs[0]
        A, x = ctx.saved_tensors
        x0 = None
        if ctx.cached_solve is not None and ctx.cached_solve.previous_grad is not None:
            x0 = ctx.cached_solve.previous_grad
        At = A.t().coalesce() if A.layout == torch.sparse_coo else A.t()
        gradb, _ = sparse_solve(At, grad_x, ctx.B, ctx.stol, ctx.device, ctx.method, ctx.M, x0=x0)
        Ac = A if A.layout != torch.sparse_coo or A.is_coalesced() else A.coalesce()
        indices = Ac._indices()
        val = -gradb[indices[0]] * x[indices[1]]
        gradA = torch.sparse_coo_tensor(indices, val, A.shape, is_coalesced=True)
        if ctx.update_cache and ctx.cached_solve is not None:
            ctx.cached_solve.update_grad(gradb.detach().clone())
        return gradA, gradb, None, None, None, None, None, None, None


def differentiable_sparse_solve(A, b, B=None, stol=1e-10, device=None, method=None, M=None,
                                cached_solve: CachedSolve | None = None,
                                update_cache: bool = False) -> Tensor:
    """Solve A x = b with the custom adjoint backward (sparse.py:246-268)."""
    result, _ = Solve.apply(A, b, B, stol, device, method, M, cached_solve, update_cache)
    if result is None:
        raise RuntimeError("Solve.apply returned None, expected a Tensor.")
    return result


class NewtonRaphsonAdjoint(Function):
    """Newton iterations on a residual callback with an implicit adjoint
    backward (sparse.py:517-724), on the sparse solves of this module."""

    @staticmethod
    def forward(ctx, eval_residual: Callable, du, B, max_iter, rtol, atol, stol, report,
                method=None, device=None, cached_solve=None, update_cache=False, u_prev=None,
                grad_prev=None, flux_prev=None, state_prev=None, *parameters):
        M = None
        converged_iter = max_iter - 1
        for i in range(max_iter):
            residual, K = eval_residual(du, i, u_prev, grad_prev, flux_prev, state_prev)
            res_norm = torch.linalg.norm(residual)
            if i == 0:
                res_norm0 = res_norm
            if report is not None:
                report.iteration(i, res_norm)
            if res_norm < rtol * res_norm0 or res_norm < atol:
                converged_iter = i
                break
            if not torch.isfinite(res_norm):
                break
            x0 = None
            if i == 0 and cached_solve is not None and cached_solve.previous_x is not None:
                x0 = cached_solve.previous_x
            du_i, M = sparse_solve(K, residual, B, stol, device, method, M, x0=x0)
            if i == 0 and update_cache and cached_solve is not None:
                cached_solve.update_x(du_i)
            du = du - du_i
        if not (res_norm < rtol * res_norm0 or res_norm < atol):
            raise RuntimeError("Newton-Raphson iteration did not converge.")
        ctx.save_for_backward(K, du, u_prev, grad_prev, flux_prev, state_prev, *parameters)
        ctx.B, ctx.M, ctx.stol, ctx.device, ctx.method = B, M, stol, device, method
        ctx.eval_residual, ctx.cached_solve, ctx.update_cache = eval_residual, cached_solve, update_cache
        ctx.converged_iter = converged_iter
        return du

    @staticmethod
    def backward(ctx, *grad_outputs):
        grad_du = grad_outputs[0]
        K, du, u_prev, grad_prev, flux_prev, state_prev, *parameters = ctx.saved_tensors
        x0 = None
        if ctx.cached_solve is not None and ctx.cached_solve.previous_grad is not None:
            x0 = ctx.cached_solve.previous_grad
        Kt = K.t().coalesce() if K.layout == torch.sparse_coo else K.t()
        lambda_, _ = sparse_solve(Kt, grad_du, ctx.B, ctx.stol, ctx.device, ctx.method, ctx.M,
                                  x0=x0)
        if ctx.update_cache and ctx.cached_solve is not None:
            ctx.cached_solve.update_grad(lambda_)
        du_local = du.detach().requires_grad_(True)
        prev_local = tuple(
            p.detach().requires_grad_(True) if p is not None else None
            for p in (u_prev, grad_prev, flux_prev, state_prev)
        )
        with torch.enable_grad():
            residual, _ = ctx.eval_residual(du_local, ctx.converged_iter, *prev_local)
        inputs = [t for t in (du_local, *prev_local, *parameters) if t is not None]
        grads = torch.autograd.grad(residual, inputs, grad_outputs=-lambda_, allow_unused=True,
                                    retain_graph=True)
        it = iter(grads)
        full = [next(it) if t is not None else None for t in (du_local, *prev_local, *parameters)]
        return (None,) * 12 + tuple(full[1:5]) + tuple(full[5:])


def newton_solve(eval_residual, du, B, max_iter, rtol, atol, stol, report, method=None,
                 device=None, cached_solve: CachedSolve | None = None, update_cache=False,
                 u_prev=None, grad_prev=None, flux_prev=None, state_prev=None, *parameters):
    """Adjoint-safe Newton solve of a residual callback (sparse.py:727-795)."""
    du = NewtonRaphsonAdjoint.apply(eval_residual, du, B, max_iter, rtol, atol, stol, report,
                                    method, device, cached_solve, update_cache, u_prev,
                                    grad_prev, flux_prev, state_prev, *parameters)
    if du is None:
        raise RuntimeError("Solve.apply returned None, expected a Tensor.")
    return du
