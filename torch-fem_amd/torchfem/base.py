"""FEM core: the assemble-and-solve path on the GPU (src/torchfem/base.py).

`FEM.solve()` keeps the reference's signature and outputs (base.py:597-923)
for linear, stateless materials. The work runs in libtfm_hip.so:

* pattern: host-built node adjacency (FEM.__init__, base.py:72-122);
* integration + assembly: one fused kernel writes the CSR values with the
  Dirichlet epilogue (integrate_material + assemble_matrix, base.py:398-426,
  982-1092, 1177-1286), deterministic, no float atomics;
* solve: Jacobi / block-Jacobi PCG, all-FP64 or FP32-inner with FP64
  iterative refinement on the matrix-free FP64 residual (sparse_solve,
  sparse.py:271-444);
* gradients: a hand-written backward (NewtonRaphsonAdjoint.backward,
  sparse.py:641-724): the adjoint reuses the PCG and a fused kernel forms
  -lam_e^T (dK_e/dp_e) u_e for every element with K_e recomputed.

For a linear material the Newton loop of the reference converges after one
linear solve per load state; the implementation keeps that loop (residual
check with rtol/atol, further corrections if needed) on the total
displacement, and scales the unit-load solution to the requested increments,
which is exact for a linear model.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from abc import ABC, abstractmethod
from collections.abc import Iterable
from functools import cached_property

import torch
from torch import Tensor

from . import _lib as L
from ._pattern import DevicePattern, SolverView, locality_order
from .elements import Element
from .materials import Material
from ._coarse import CoarseSpace

# FP32 residual reduction between two reliable updates of a mixed solve, per
# preconditioner (timed on the Hexa1 64^3 cube: block-Jacobi 0.1 against 0.03
# and 0.3; two-level 0.03 against 0.01, 0.1 and 0.3)
INNER_RTOL = 0.1
INNER_RTOL_TWO_LEVEL = 0.03
from ._finite import Kinematics, solve_finite
from .sparse import (AUTO_TWO_LEVEL_DOFS, PRECONDITIONERS, CachedSolve, PCGSolver,
                     default_precision, default_preconditioner)

_PREC = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}
_PC = {"jacobi": L.PC_JACOBI, "block_jacobi": L.PC_BLOCK_JACOBI, "two_level": L.PC_TWO_LEVEL}


def _n_out(kind: int, dim: int) -> int:
    return {L.MAT_ISO: 3, L.MAT_GENERAL: dim**4 + 1, L.MAT_HEAT_ISO: 2,
            L.MAT_HEAT_GENERAL: dim**2 + 1}[kind]


class _System:
    """Assembled operator of one solve: CSR values (FP64 and/or FP32), nodal
    diagonal blocks, the Dirichlet mask and the material it was built from."""

    def __init__(self, fem: "FEM", kind: int, p0: Tensor, p1: Tensor | None,
                 scale: Tensor | None, precision: str, nrhs: int = 1, renumber: bool = True):
        # the solver's node order: the mesh's own, or a locality order when the
        # mesh numbering scatters neighbours (FEM._solver_view); `perm` maps a
        # solver DOF to the model DOF, and FEM._pcg translates vectors
        view = fem._solver_view() if renumber else None
        pat = view.pattern if view is not None else fem._pattern
        self.perm = view.perm_dof if view is not None else None
        self._pat = pat
        dev = pat.device
        self.fem = fem
        self.kind = kind
        self.p0 = p0.detach().contiguous()
        self.p1 = p1.detach().contiguous() if p1 is not None else None
        self.scale = scale.detach().contiguous() if scale is not None else None
        self.mat = L.Material(kind=kind, p0=self.p0.data_ptr(),
                              p1=self.p1.data_ptr() if self.p1 is not None else None,
                              scale=self.scale.data_ptr() if self.scale is not None else None)
        self.con = fem._con_mask()  # model order
        con_s = self.con[self.perm].contiguous() if self.perm is not None else self.con
        self._con_s = con_s
        self.precision = precision
        # mixed precision keeps the FP64 values too: its refinement residuals
        # b - K x are then one FP64 SpMV instead of a matrix-free pass
        self.val64 = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
        self.val32 = torch.empty(pat.nnz, dtype=torch.float32, device=dev) if precision == "mixed" else None
        nd = pat.ndof
        self.diag = torch.empty(pat.n_nod * nd * nd, dtype=torch.float64, device=dev)
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        # FP32 values in difference form for one or two right-hand sides (the
        # form the solver's FP32 products take for that many), plain for more
        form = L.FP32_DIFF if nrhs <= 2 else L.FP32_PLAIN
        L.check(L.lib().tfm_assemble(pat.stream(), C.byref(pat.mesh), C.byref(self.mat),
                                     L.ptr(con_s), L.ptr(self.val64), L.ptr(self.val32),
                                     L.ptr(self.diag), L.ptr(bad), form))
        fem._check_jacobian(bad)
        self.sys = L.System(mesh=C.pointer(pat.mesh), mat=C.pointer(self.mat),
                            con=con_s.data_ptr(),
                            val64=L.ptr(self.val64), val32=L.ptr(self.val32),
                            diag_blocks=self.diag.data_ptr(), val32_form=form)

    def force(self, u: Tensor) -> Tensor:
        """K u with K unconstrained (matrix-free, FP64)."""
        return self.fem._forces(self.mat, u)[0]

    def sensitivity(self, u: Tensor, lam: Tensor, sign: float) -> Tensor:
        pat = self.fem._pattern
        u = u.reshape(-1, pat.n_dofs).contiguous()
        lam = lam.reshape(-1, pat.n_dofs).contiguous()
        nr = u.shape[0]
        out = torch.empty(nr, pat.n_elem, _n_out(self.kind, pat.dim), dtype=torch.float64,
                          device=pat.device)
        L.check(L.lib().tfm_sensitivity(pat.stream(), C.byref(pat.mesh), C.byref(self.mat),
                                        L.ptr(u), L.ptr(lam), nr, sign, L.ptr(out)))
        return out

    def param_grads(self, out: Tensor, needs: tuple[bool, bool, bool]):
        """Split sensitivity outputs [E, n_out] into grads of (p0, p1, scale)."""
        d = self.fem._pattern.dim
        g0 = g1 = gs = None
        if self.kind == L.MAT_ISO:
            g0, g1, gs = out[:, 0], out[:, 1], out[:, 2]
        elif self.kind == L.MAT_HEAT_ISO:
            g0, gs = out[:, 0], out[:, 1]
        elif self.kind == L.MAT_GENERAL:
            g0, gs = out[:, : d**4], out[:, d**4]
        else:
            g0, gs = out[:, : d**2], out[:, d**2]
        return (g0 if needs[0] else None, g1 if needs[1] else None, gs if needs[2] else None)


class _StaticSolve(torch.autograd.Function):
    """u = K(p)^{-1} [F_ext on free DOFs; U_D on constrained DOFs], with the
    implicit-function backward of NewtonRaphsonAdjoint (sparse.py:641-724):
    K^T lam = grad_u (free part), dL/dp = -lam^T d(K u)/dp, dL/dF = lam,
    dL/dU_D = grad_u - (K lam) on constrained DOFs.

    F_ext may hold several load cases ([nrhs, n_dofs]); they are solved as one
    batched multi-RHS PCG, and so are their adjoints."""

    @staticmethod
    def forward(ctx, fem, cfg, F_ext, U_D, p0, p1, scale):
        n = fem.n_dofs
        F2 = F_ext.reshape(-1, n)
        nr = F2.shape[0]
        U2 = U_D.reshape(-1, n).expand(nr, n)
        system = _System(fem, cfg["kind"], p0, p1, scale, cfg["precision"], nrhs=nr)
        con = system.con.bool()
        u0 = torch.where(con, U2, torch.zeros_like(U2))
        d = fem._newton(system, cfg, F2, u0, con)
        u = u0 + d
        ctx.fem, ctx.cfg, ctx.system = fem, cfg, system
        ctx.shapes = (F_ext.shape, U_D.shape)
        ctx.save_for_backward(u)
        fem._last_system = system
        return u.reshape(F_ext.shape)

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        fem, cfg, system = ctx.fem, ctx.cfg, ctx.system
        n = fem.n_dofs
        con = system.con.bool()
        g2 = g.reshape(-1, n).contiguous()
        g_free = torch.where(con, torch.zeros_like(g2), g2)
        t0 = time.perf_counter()
        warm = cfg.get("x0_adjoint")
        if warm is not None and warm.numel() != g_free.numel():
            warm = None
        lam = fem._pcg(system, g_free, cfg, warm=warm, adjoint=True)
        if cfg.get("cached") is not None and cfg.get("update_cache"):
            cfg["cached"].update_grad(lam.reshape(g.shape))
        fem.last_solve_info["adjoint"] = dict(fem._last_pcg, time=time.perf_counter() - t0)
        needs = ctx.needs_input_grad
        F_shape, D_shape = ctx.shapes
        gF = lam.reshape(F_shape) if needs[2] else None
        gD = None
        if needs[3]:
            Klam = torch.stack([system.force(row) for row in lam])
            gD = torch.where(con, g2 - Klam, torch.zeros_like(g2))
            gD = gD.sum(0).reshape(D_shape) if gD.shape[0] != 1 and len(D_shape) == 1 else gD.reshape(D_shape)
        g0 = g1 = gs = None
        if any(needs[4:7]):
            out = system.sensitivity(u, lam, -1.0).sum(0)
            g0, g1, gs = system.param_grads(out, needs[4:7])
        return None, None, gF, gD, g0, g1, gs


class _Integrate(torch.autograd.Function):
    """Converged-state evaluation (base.py:844-867): internal forces f = K u
    (reactions included), Gauss-point mean gradient and flux, for one state
    [n_dofs] or several [nrhs, n_dofs]."""

    @staticmethod
    def forward(ctx, fem, kind, u, p0, p1, scale):
        h = _MatHolder(kind, p0, p1, scale)
        rows = u.reshape(-1, fem.n_dofs)
        outs = [fem._forces(h.mat, row, means=True) for row in rows]
        F = torch.stack([o[0] for o in outs])
        gm = torch.stack([o[2] for o in outs])
        qm = torch.stack([o[3] for o in outs])
        ctx.fem, ctx.holder, ctx.single = fem, h, u.dim() == 1
        ctx.save_for_backward(u)
        if u.dim() == 1:
            return F[0], gm[0], qm[0]
        return F, gm, qm

    @staticmethod
    def backward(ctx, gF, ggrad, gflux):
        (u,) = ctx.saved_tensors
        fem, h = ctx.fem, ctx.holder
        needs = ctx.needs_input_grad
        if (ggrad is not None and bool(ggrad.abs().max() > 0)) or (
            gflux is not None and bool(gflux.abs().max() > 0)
        ):
            raise NotImplementedError(
                "gradients through the stress/flux or deformation-gradient outputs are "
                "not part of this build; differentiate u or f"
            )
        gu = g0 = g1 = gs = None
        if gF is None:
            return None, None, gu, g0, g1, gs
        gF2 = gF.reshape(-1, fem.n_dofs).contiguous()
        if needs[2]:
            gu = torch.stack([fem._forces(h.mat, row)[0] for row in gF2]).reshape(u.shape)
        if any(needs[3:6]):
            sysv = _SensView(fem, h)
            out = sysv.sensitivity(u.reshape(-1, fem.n_dofs), gF2, 1.0).sum(0)
            g0, g1, gs = sysv.param_grads(out, needs[3:6])
        return None, None, gu, g0, g1, gs


class _MatHolder:
    def __init__(self, kind, p0, p1, scale):
        self.kind = kind
        self.p0 = p0.detach().contiguous()
        self.p1 = p1.detach().contiguous() if p1 is not None else None
        self.scale = scale.detach().contiguous() if scale is not None else None
        self.mat = L.Material(kind=kind, p0=self.p0.data_ptr(),
                              p1=self.p1.data_ptr() if self.p1 is not None else None,
                              scale=self.scale.data_ptr() if self.scale is not None else None)


class _SensView(_System):
    """Sensitivity helpers for a material without an assembled matrix."""

    def __init__(self, fem, holder):  # noqa: D401 - no assembly here
        self.fem = fem
        self.kind = holder.kind
        self.mat = holder.mat


class FEM(ABC):
    """Finite-element model on one GPU (base.py:23-923)."""

    def __init__(self, nodes: Tensor, elements: Tensor, material: Material | None):
        device = L.require_gpu(nodes.device if nodes.is_cuda else None)
        self.device = device
        self.nodes = nodes.to(device=device, dtype=torch.float64)
        self.elements = elements.to(device=device, dtype=torch.int64)
        self.n_dofs = self.n_dof_per_node * nodes.shape[0]
        self.n_nod = nodes.shape[0]
        self.n_dim = nodes.shape[1]
        self.n_elem = len(self.elements)
        self.n_int = len(self.etype.iweights)

        dt = torch.float64
        self._neumann = torch.zeros(self.n_nod, self.n_dof_per_node, dtype=dt, device=device)
        self._dirichlet = torch.zeros(self.n_nod, self.n_dof_per_node, dtype=dt, device=device)
        self._constraints = torch.zeros(self.n_nod, self.n_dof_per_node, dtype=torch.bool,
                                        device=device)
        self._external_gradient = torch.zeros(self.n_elem, *self.n_flux, dtype=dt, device=device)

        # local-to-global DOF map, node-major / component-minor (base.py:73-76, 119)
        idx = (self.n_dof_per_node * self.elements).unsqueeze(-1) + torch.arange(
            self.n_dof_per_node, device=device
        )
        self.idx = idx.reshape(self.n_elem, -1).to(torch.int32)

        t0 = time.perf_counter()
        self._pattern = DevicePattern(self.nodes, self.elements, self.n_dof_per_node,
                                      self.etype.__name__, device)
        self.setup_time = time.perf_counter() - t0

        if material is None or material.is_vectorized:
            self.material = material
        else:
            self.material = material.vectorize(self.n_elem)
        self.cached_solve = CachedSolve()
        self._pcg_handle: PCGSolver | None = None
        self._sv_checked = False
        self._sv = None
        self._last_system: _System | None = None
        self._K_cache = None
        self.last_solve_info: dict = {}

    # ------------------------------------------------------------- abstract
    @property
    def n_state(self) -> int:
        return 0 if self.material is None else self.material.n_state

    @property
    def volume_scale(self) -> Tensor:
        return torch.ones(self.n_elem, dtype=torch.float64, device=self.device)

    @property
    @abstractmethod
    def n_flux(self) -> list[int]: ...

    @property
    @abstractmethod
    def n_dof_per_node(self) -> int: ...

    @property
    @abstractmethod
    def etype(self) -> type[Element]: ...

    @property
    @abstractmethod
    def initial_grad(self) -> Tensor: ...

    def _scale_tensor(self) -> Tensor | None:
        """Per-element volume scale passed to the kernels (thickness), or None."""
        return None

    # ----------------------------------------------------- reference indices
    @cached_property
    def glob_idx(self) -> Tensor:
        """Row/column indices [2, nnz] int64 of the pattern (base.py:112-117)."""
        e = self._pattern.expand(rows=True, cols=True)
        return torch.stack([e["rows"], e["cols"]])

    @cached_property
    def k_map(self) -> Tensor:
        """Element entry -> value slot, int32 [n_elem*(npe*ndof)^2] (base.py:94-104)."""
        return self._pattern.expand(rows=False, cols=False, k_map=True)["k_map"]

    @cached_property
    def diag_map(self) -> Tensor:
        """Diagonal slots, int32 [n_dofs] (base.py:106-108)."""
        return self._pattern.expand(rows=False, cols=False, diag_map=True)["diag_map"]

    # ------------------------------------------------------ boundary setters
    @property
    def constraints(self) -> Tensor:
        return self._constraints

    @constraints.setter
    def constraints(self, value: Tensor):
        if not value.shape == (self.n_nod, self.n_dof_per_node):
            raise ValueError("Constraints must have the same shape as nodes.")
        if value.dtype != torch.bool:
            raise TypeError("Constraints must be a boolean tensor.")
        self._constraints = value.to(self.device)

    def _con_mask(self) -> Tensor:
        return self._constraints.reshape(-1).to(torch.uint8).contiguous()

    # ---------------------------------------------------- shape functions
    def eval_shape_functions(self, xi: Tensor) -> tuple[Tensor, Tensor, Tensor]:
        """N, B = J^{-1} dN/dxi and detJ at local points xi [.., dim]
        (base.py:293-314): B [.., n_elem, dim, npe] and detJ [.., n_elem] come
        from one HIP kernel over (point, element) pairs; raises on detJ <= 0."""
        xi = torch.as_tensor(xi, dtype=torch.float64)
        lead = xi.shape[:-1]
        pts = xi.reshape(-1, xi.shape[-1]).to(self.device).contiguous()
        if pts.shape[1] != self.n_dim:
            raise ValueError(f"local points need {self.n_dim} coordinates")
        pat = self._pattern
        nq = pts.shape[0]
        B = torch.empty(nq, self.n_elem, self.n_dim, pat.npe, dtype=torch.float64,
                        device=self.device)
        detJ = torch.empty(nq, self.n_elem, dtype=torch.float64, device=self.device)
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        L.check(L.lib().tfm_shape_gradients(pat.stream(), C.byref(pat.mesh), L.ptr(pts), nq,
                                            L.ptr(B), L.ptr(detJ), L.ptr(bad)))
        self._check_jacobian(bad)
        N = self.etype.N(xi).to(self.device)
        return N, B.reshape(*lead, self.n_elem, self.n_dim, pat.npe), detJ.reshape(*lead, self.n_elem)

    def compute_B(self) -> Tensor:
        """Rigid-body modes [n_dofs, n_modes] (base.py:316-344): the unit
        translations, then the rotations e_k x X about the coordinate axes
        (in 2-D the single in-plane rotation (y, -x)); one column of ones for
        a scalar field."""
        d = self.n_dof_per_node
        X = self.nodes
        if d == 1:
            return torch.ones((self.n_nod, 1), dtype=torch.float64, device=self.device)
        eye = torch.eye(d, dtype=torch.float64, device=self.device)
        trans = eye.expand(self.n_nod, d, d)  # [node, component, mode]
        if d == 3:
            rot = torch.stack([torch.linalg.cross(eye[k].expand_as(X), X, dim=-1)
                               for k in range(3)], dim=-1)
        else:
            rot = torch.stack([X[:, 1], -X[:, 0]], dim=-1)[:, :, None]
        return torch.cat([trans, rot], dim=-1).reshape(self.n_nod * d, -1)

    def integrate_shape_functions(self) -> Tensor:
        """Integrals of the shape functions over each element [n_elem, npe]:
        (w_g detJ_ge) contracted with N_gI over the Gauss points."""
        N, _, detJ = self.eval_shape_functions(self.etype.ipoints)
        w = self.etype.iweights.to(device=self.device, dtype=torch.float64)
        return (detJ * w[:, None]).T @ N

    def integrate_field(self, field: Tensor | None = None) -> Tensor:
        """Element integrals of a nodal field (its volume for None)."""
        w = self.integrate_shape_functions()
        if field is None:
            return w.sum(dim=1)
        return (w * field[self.elements]).sum(dim=1)

    # --------------------------------------------------------- kernels
    def _check_jacobian(self, bad: Tensor) -> None:
        if int(bad.item()) != 0:
            raise ValueError("Negative Jacobian. Check element numbering.")

    def _material_params(self):
        if self.material is None:
            raise ValueError("this model has no material")
        kind, p0, p1 = self.material.kernel_params(self.n_elem, self.device)
        return kind, p0, p1, self._scale_tensor()

    def _forces(self, mat: L.Material, u: Tensor, means=False, per_gp=False,
                flux_prev: Tensor | None = None):
        """Internal forces K u (FP64, matrix-free) and optional gradients/fluxes."""
        pat = self._pattern
        dev = pat.device
        u = u.reshape(-1).to(torch.float64).contiguous()
        F = torch.empty(pat.n_dofs, dtype=torch.float64, device=dev)
        work = torch.empty(pat.n_elem * pat.npe * pat.ndof, dtype=torch.float64, device=dev)
        nq = pat.ndof * pat.dim
        gm = torch.empty(pat.n_elem, nq, dtype=torch.float64, device=dev) if means else None
        qm = torch.empty(pat.n_elem, nq, dtype=torch.float64, device=dev) if means else None
        gg = torch.empty(self.n_int, pat.n_elem, nq, dtype=torch.float64, device=dev) if per_gp else None
        qg = torch.empty(self.n_int, pat.n_elem, nq, dtype=torch.float64, device=dev) if per_gp else None
        qp = flux_prev.reshape(self.n_int, pat.n_elem, nq).to(torch.float64).contiguous() if flux_prev is not None else None
        L.check(L.lib().tfm_internal_force(pat.stream(), C.byref(pat.mesh), C.byref(mat),
                                           L.ptr(u), L.ptr(qp), L.ptr(F), L.ptr(gm), L.ptr(qm),
                                           L.ptr(gg), L.ptr(qg), L.ptr(work)))
        return F, work, gm, qm, gg, qg

    def _solver(self) -> PCGSolver:
        if self._pcg_handle is None:
            self._pcg_handle = PCGSolver(self.n_dofs, nrhs=1, device=self.device)
        return self._pcg_handle

    def _pcg(self, system: _System, b: Tensor, cfg: dict, warm: Tensor | None = None,
             adjoint: bool = False) -> Tensor:
        """K_bc x = b for rows of b ([nrhs, n]) with the configured PCG. A mixed
        forward solve ends its last FP32 segment at 0.25 stol, since the error
        of u is reported against the FP64 result; an adjoint solve, whose error
        enters the gradient, at 0.7 stol."""
        b = b.to(torch.float64).contiguous()
        x = torch.zeros_like(b) if warm is None else warm.reshape(b.shape).to(torch.float64).clone()
        perm = getattr(system, "perm", None)
        if perm is not None:  # to the solver's node order
            b, x = b[:, perm].contiguous(), x[:, perm].contiguous()
        solver = self._solver()
        # once the FP32 iterations of a mixed solve broke down on this system
        # (its FP32 operator is not positive definite), later solves with the
        # same matrix (the adjoint, further Newton steps) run in FP64 directly
        precision = "fp64" if getattr(system, "fp32_broke", False) else cfg["precision"]
        kw = dict(precision=precision, rtol=cfg["stol"],
                  final_rtol=(0.7 if adjoint else 0.25) * cfg["stol"])
        ir = cfg["inner_rtol"]
        use_two = cfg["pc"] == "two_level" and not getattr(system, "coarse_failed", False)
        if use_two:
            try:
                coarse = self._coarse_for(system)
                info = solver.solve_system(system.sys, b, x, preconditioner="two_level",
                                           coarse=coarse, inner_rtol=INNER_RTOL_TWO_LEVEL
                                           if ir is None else ir, **kw)
            except RuntimeError:
                # the coarse space could not be built, or the coarse correction
                # did not reach the tolerance (seen on density designs with a
                # 1e9 stiffness contrast): block-Jacobi continues from the
                # iterate, here and for later solves with this matrix
                system.coarse_failed = True
                use_two = False
        if not use_two:
            pc = "block_jacobi" if cfg["pc"] == "two_level" else cfg["pc"]
            info = solver.solve_system(system.sys, b, x, preconditioner=pc,
                                       inner_rtol=INNER_RTOL if ir is None else ir, **kw)
            if cfg["pc"] == "two_level":
                info["two_level_fallback"] = True
        if info.get("fp64_fallback", 0) > 0:
            system.fp32_broke = True
        self._last_pcg = info
        if perm is not None:
            out = torch.empty_like(x)
            out[:, perm] = x
            return out
        return x

    def _resolve_pc(self, preconditioner: str | None) -> str:
        """The preconditioner a solve runs with: None takes
        default_preconditioner(); "auto" the two-level preconditioner from
        AUTO_TWO_LEVEL_DOFS unknowns, block-Jacobi below."""
        name = preconditioner or default_preconditioner()
        if name not in PRECONDITIONERS:
            raise ValueError(f"preconditioner must be one of {sorted(PRECONDITIONERS)}")
        if name == "auto":
            name = "two_level" if self.n_dofs >= AUTO_TWO_LEVEL_DOFS else "block_jacobi"
        return name

    def _coarse_for(self, system):
        """The two-level preconditioner's coarse space for the system's
        pattern (built once per pattern), holding the inverse of the system's
        coarse operator (formed once per system); returns its handle."""
        pat = getattr(system, "_pat", None) or self._pattern
        cache = self.__dict__.setdefault("_coarse_spaces", {})
        cs = cache.get(id(pat))
        if cs is None or cs.pat is not pat:
            cs = CoarseSpace(pat)
            cache[id(pat)] = cs
        inv = getattr(system, "_coarse_inv", None)
        if inv is None:
            inv = cs.inverse(system)
            system._coarse_inv = inv
        L.check(L.lib().tfm_coarse_set_inverse(cs.h, L.ptr(inv[0]), L.ptr(inv[1])))
        return cs.h

    def _solver_view(self):
        """The node order the PCG runs in. A mesh whose numbering puts
        neighbours far apart (16-bit neighbour offsets do not fit, e.g. the
        mid-side nodes that linear_to_quadratic appends after the corners)
        scatters the SpMV's gathers over HBM; the solver then works on a copy
        of the pattern in a locality order (`locality_order`: nodes sorted by
        element-size bins of their coordinates) when that order brings every
        offset within 16 bits. Elements keep their order, so material arrays
        and element outputs are unchanged; only nodal vectors are permuted,
        in FEM._pcg. None when the mesh's own order is kept."""
        if self._sv_checked:
            return self._sv
        self._sv_checked = True
        pat = self._pattern
        if pat.adj16 is not None or self.n_nod < 2 or self.n_elem == 0:
            return None
        perm = locality_order(self.nodes, self.n_elem)
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(self.n_nod, device=self.device)
        spat = DevicePattern(self.nodes[perm], inv[self.elements], self.n_dof_per_node,
                             self.etype.__name__, self.device)
        if spat.adj16 is None:
            return None
        nd = self.n_dof_per_node
        perm_dof = (perm[:, None] * nd + torch.arange(nd, device=self.device)).reshape(-1)
        self._sv = SolverView(spat, perm_dof.contiguous())
        return self._sv

    def _newton(self, system: _System, cfg: dict, F_ext: Tensor, u0: Tensor, con: Tensor) -> Tensor:
        """Newton loop of NewtonRaphsonAdjoint.forward (sparse.py:541-639) on the
        corrections d (zero on constrained DOFs) of a linear model, for one or
        several load cases (rows of F_ext and u0) at once."""
        t0 = time.perf_counter()
        zero = torch.zeros_like(F_ext)

        def residual(d):
            Ku = torch.stack([system.force(row) for row in u0 + d])
            return torch.where(con, zero, F_ext - Ku)

        d = torch.zeros_like(F_ext)
        r = residual(d)
        res0 = torch.linalg.norm(r, dim=1)
        iters = []
        for i in range(cfg["max_iter"]):
            res = torch.linalg.norm(r, dim=1) if i else res0
            if not bool(torch.isfinite(res).all()):
                break
            ok = (res < cfg["rtol"] * res0) | (res < cfg["atol"])
            if bool(ok.all()):
                self.last_solve_info["forward"] = dict(
                    newton_iterations=i, pcg=iters, res0=res0.tolist(), res=res.tolist(),
                    time=time.perf_counter() - t0)
                return d
            rhs = torch.where(ok[:, None], zero, r)  # converged cases get a zero rhs
            warm = cfg.get("x0") if i == 0 else None
            if warm is not None and warm.numel() != rhs.numel():
                warm = None
            dd = self._pcg(system, rhs, cfg, warm=warm)
            iters.append(self._last_pcg)
            if i == 0 and cfg.get("cached") is not None and cfg.get("update_cache"):
                cfg["cached"].update_x(dd.reshape(-1) if dd.shape[0] == 1 else dd)
            d = d + dd
            r = residual(d)
        raise RuntimeError("Newton-Raphson iteration did not converge.")

    # ------------------------------------------------------- API methods
    def k0(self) -> Tensor:
        """Element matrices of the reference state (base.py:228-245)."""
        kind, p0, p1, scale = self._material_params()
        h = _MatHolder(kind, p0, p1, scale)
        pat = self._pattern
        W = pat.npe * pat.ndof
        k = torch.empty(pat.n_elem, W, W, dtype=torch.float64, device=self.device)
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        L.check(L.lib().tfm_element_matrices(pat.stream(), C.byref(pat.mesh), C.byref(h.mat),
                                             L.ptr(k), L.ptr(bad)))
        self._check_jacobian(bad)
        return k

    def integrate_material(self, u_prev, grad_prev, flux_prev, state_prev, du, de0, iter, nlgeom,
                           compute_stiffness: bool = True):
        """Element stiffness, forces, gradients and fluxes for the increment du
        from a previous state (base.py:247-278, 982-1092, 1177-1286). A
        hyperelastic material integrates F = grad_prev + grad(du) with its
        stress and tangent per Gauss point; the flux is then the Cauchy stress
        F P / J with nlgeom and P without it."""
        if getattr(self.material, "incremental", False):
            from ._finite import FiniteStrainRun

            run = FiniteStrainRun.__new__(FiniteStrainRun)
            run.mat, run.kin = self.material, Kinematics(self)
            dev, dt = self.device, torch.float64
            du = torch.as_tensor(du).reshape(-1).to(dev, dt)
            grad_prev = grad_prev.to(dev, dt)
            flux_prev = (torch.zeros_like(grad_prev) if flux_prev is None
                         else flux_prev.to(dev, dt))
            state_prev = state_prev.to(dev, dt)
            H = run.kin.grad(du)
            P, state_new, A = run.update(H, grad_prev, flux_prev, state_prev, iter,
                                         tangent=compute_stiffness)
            k = run.kin.matrices(A) if compute_stiffness else None
            F = grad_prev + H
            f = run.kin.element_forces(P)
            flux = F @ P / torch.linalg.det(F)[..., None, None] if nlgeom else P
            return k, f, F, flux, state_new
        if nlgeom:
            raise NotImplementedError("nlgeom is supported for hyperelastic materials only")
        if de0 is not None and torch.as_tensor(de0).abs().max() > 0:
            raise NotImplementedError("external gradient increments are not part of this build")
        kind, p0, p1, scale = self._material_params()
        h = _MatHolder(kind, p0, p1, scale)
        _, fe, _, _, gg, qg = self._forces(h.mat, du, per_gp=True, flux_prev=flux_prev)
        nd, dim = self.n_dof_per_node, self.n_dim
        shape = (self.n_int, self.n_elem, *self.n_flux)
        grad_new = grad_prev + gg.reshape(shape)
        flux_new = qg.reshape(shape)
        f = fe.reshape(self.n_elem, -1)
        k = self.k0() if compute_stiffness else None
        return k, f, grad_new, flux_new, state_prev.clone()

    def assemble_matrix(self, k: Tensor, con: Tensor) -> Tensor:
        """Global sparse matrix from element matrices with Dirichlet rows and
        columns replaced by the identity (base.py:398-426)."""
        pat = self._pattern
        k = k.to(device=self.device, dtype=torch.float64).contiguous()
        con = torch.as_tensor(con, device=self.device)
        if con.dtype == torch.bool:
            # a boolean mask over the DOFs, as is_constrained[con] = True accepts
            if con.numel() != self.n_dofs:
                raise ValueError("a boolean constraint mask needs one entry per DOF")
            mask = con.reshape(-1).to(torch.uint8).contiguous()
        else:
            mask = torch.zeros(self.n_dofs, dtype=torch.uint8, device=self.device)
            mask[con.long()] = 1
        val = torch.empty(pat.nnz, dtype=torch.float64, device=self.device)
        diag = torch.empty(pat.n_nod * pat.ndof**2, dtype=torch.float64, device=self.device)
        L.check(L.lib().tfm_assemble_elements(pat.stream(), C.byref(pat.mesh), L.ptr(k),
                                              L.ptr(mask), L.ptr(val), None, L.ptr(diag),
                                              L.FP32_PLAIN))
        self.is_constrained = mask.bool()
        with torch.sparse.check_sparse_tensor_invariants(False):
            return torch.sparse_coo_tensor(self.glob_idx, val, size=(self.n_dofs, self.n_dofs),
                                           is_coalesced=True)

    def assemble_rhs(self, f: Tensor) -> Tensor:
        """Global vector from element vectors [n_elem, npe*ndof] (base.py:428-445),
        gathered in ascending element order."""
        pat = self._pattern
        if f.requires_grad:
            # keep autograd: the gather is linear, its adjoint is an index select
            return _Gather.apply(self, f)
        fe = f.detach().to(device=self.device, dtype=torch.float64).contiguous()
        F = torch.empty(self.n_dofs, dtype=torch.float64, device=self.device)
        L.check(L.lib().tfm_gather_nodes(pat.stream(), C.byref(pat.mesh), L.ptr(fe), L.ptr(F)))
        return F

    @property
    def K(self) -> Tensor:
        """The last assembled global stiffness as sparse COO (base.py:699, 738)."""
        if self._K_cache is not None:
            return self._K_cache
        s = self._last_system
        if s is None:
            return torch.empty(0, device=self.device)
        if s.val64 is not None and s.perm is None:
            val = s.val64
        else:  # values in the solver's node order: assemble in the model's
            s64 = _System(self, s.kind, s.p0, s.p1, s.scale, "fp64", renumber=False)
            val = s64.val64
        with torch.sparse.check_sparse_tensor_invariants(False):
            self._K_cache = torch.sparse_coo_tensor(self.glob_idx, val,
                                                    size=(self.n_dofs, self.n_dofs),
                                                    is_coalesced=True)
        return self._K_cache

    @K.setter
    def K(self, value):
        self._K_cache = None if (value is None or value.numel() == 0) else value

    # -------------------------------------------------------------- loads
    def _facet_set(self, mask: Tensor) -> dict:
        """Boundary facets of a nodal selection, found on the device
        (tfm_boundary_facets): their list position per (element, facet) slot,
        the list of slots in the reference's order, and their count."""
        sel = torch.as_tensor(mask, device=self.device).reshape(-1)
        if sel.dtype != torch.bool or sel.numel() != self.n_nod:
            raise ValueError("The mask must be a boolean tensor with one entry per node.")
        sel = sel.to(torch.uint8).contiguous()
        pat = self._pattern
        nf, nfn = L.I32(), L.I32()
        L.check(L.lib().tfm_facet_count(C.byref(pat.mesh), C.byref(nf), C.byref(nfn)))
        slots = self.n_elem * nf.value
        i32 = dict(dtype=torch.int32, device=self.device)
        flag = torch.empty(max(slots, 1), dtype=torch.int8, device=self.device)
        pos = torch.empty(max(slots, 1), **i32)
        off = torch.empty(self.n_nod + 1, **i32)
        bsum = torch.empty((self.n_nod + 1023) // 1024 + 1, **i32)
        lst = torch.empty(max(slots, 1), **i32)
        n = L.I64()
        L.check(L.lib().tfm_boundary_facets(pat.stream(), C.byref(pat.mesh), L.ptr(sel),
                                            L.ptr(flag), L.ptr(pos), L.ptr(off), L.ptr(bsum),
                                            L.ptr(lst), slots, C.byref(n)))
        return dict(mask=sel, pos=pos, list=lst, n=n.value, nfn=nfn.value)

    def _boundary_facets(self, mask: Tensor) -> Tensor:
        """Node ids [n_facets, nodes per facet] of the boundary facets whose
        nodes all lie in `mask` (base.py:452-470), in the reference's order."""
        fs = self._facet_set(mask)
        out = torch.empty(fs["n"], fs["nfn"], dtype=torch.int64, device=self.device)
        pat = self._pattern
        L.check(L.lib().tfm_facet_nodes(pat.stream(), C.byref(pat.mesh), L.ptr(fs["list"]),
                                        fs["n"], L.ptr(out)))
        return out

    @staticmethod
    def _load_tensor(load, device) -> Tensor:
        load = torch.as_tensor(load)
        if load.requires_grad:
            raise NotImplementedError(
                "gradients with respect to distributed loads are not part of this build")
        return load.to(device=device, dtype=torch.float64)

    def _facet_load(self, mask: Tensor, load) -> Tensor:
        """Consistent nodal loads of a load per unit facet measure on the
        boundary facets of `mask` (base.py:472-494), by tfm_facet_load."""
        load = self._load_tensor(load, self.device)
        fs = self._facet_set(mask)
        if load.dim() == 0 and self.n_dof_per_node > 1:
            kind, k, buf = 0, self.n_dim, load.reshape(1)  # pressure along the outward normal
        elif load.dim() == 2:
            if load.shape[0] != fs["n"]:
                raise ValueError(f"a per-facet load needs one row per facet ({fs['n']})")
            kind, k, buf = 2, load.shape[1], load
        else:
            buf = load.reshape(-1)
            kind, k = 1, buf.numel()
        buf = buf.contiguous()
        fc = torch.empty(max(fs["n"] * fs["nfn"] * k, 1), dtype=torch.float64, device=self.device)
        F = torch.empty(self.n_nod, k, dtype=torch.float64, device=self.device)
        pat = self._pattern
        L.check(L.lib().tfm_facet_load(pat.stream(), C.byref(pat.mesh), L.ptr(fs["mask"]),
                                       L.ptr(fs["pos"]), L.ptr(fs["list"]), fs["n"], L.ptr(buf),
                                       kind, k, L.ptr(fc), L.ptr(F)))
        return F

    def integrate_body_load(self, load) -> Tensor:
        """Consistent nodal loads from a load per unit volume, [k] for every
        element or [n_elem, k] (base.py:496-513), by tfm_body_load."""
        load = self._load_tensor(load, self.device)
        per_elem = load.dim() == 2
        if per_elem and load.shape[0] != self.n_elem:
            raise ValueError("a per-element load needs one row per element")
        buf = (load if per_elem else load.reshape(-1)).contiguous()
        k = buf.shape[-1]
        scale = self.volume_scale.to(torch.float64).contiguous()
        pat = self._pattern
        fe = torch.empty(max(self.n_elem * pat.npe * k, 1), dtype=torch.float64, device=self.device)
        F = torch.empty(self.n_nod, k, dtype=torch.float64, device=self.device)
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        L.check(L.lib().tfm_body_load(pat.stream(), C.byref(pat.mesh), L.ptr(scale), L.ptr(buf),
                                      int(per_elem), k, L.ptr(fe), L.ptr(F), L.ptr(bad)))
        self._check_jacobian(bad)
        return F

    def integrate_surface_load(self, mask: Tensor, load) -> Tensor:
        """Consistent nodal loads from a load per unit area (base.py:515-539):
        a float is a pressure along the outward normal, [k] or [n_facet, k] a
        traction in global coordinates."""
        if self.etype.iso_dim != 3:
            raise NotImplementedError(
                f"{type(self).__name__} has no surfaces to load. Use "
                "integrate_line_load(...) or integrate_body_load(...) instead."
            )
        return self._facet_load(mask, load)

    def integrate_line_load(self, mask: Tensor, load) -> Tensor:
        """Consistent nodal loads from a load per unit length (base.py:541-568)."""
        if self.etype.iso_dim != 2:
            raise NotImplementedError(
                f"{type(self).__name__} has no edges to load. Use "
                "integrate_surface_load(...) or integrate_body_load(...) instead."
            )
        load = torch.as_tensor(load)
        if load.dim() == 0 and self.n_dim == 3:
            raise ValueError("A line in 3D has no unique normal, so a scalar load is ambiguous. "
                             "Pass a load vector instead.")
        return self._facet_load(mask, load)

    # --------------------------------------------------------------- solve
    def solve(
        self,
        increments: Tensor | None = None,
        max_iter: int = 10,
        rtol: float = 1e-8,
        atol: float = 1e-6,
        stol: float = 1e-10,
        cutback_factor: float = 0.5,
        growth_factor: float = 1.1,
        max_cutbacks: int = 10,
        verbose: bool = False,
        method: str | None = None,
        device: str | None = None,
        return_intermediate: bool = False,
        aggregate_integration_points: bool = True,
        use_cached_solve: bool = False,
        nlgeom: bool = False,
        alpha: float = 0.0,
        differentiable_parameters: Tensor | Iterable[Tensor] | None = None,
        precision: str | None = None,
        preconditioner: str | None = None,
        inner_rtol: float | None = None,
    ):
        """Quasi-static solve by load increments (base.py:597-923).

        Extra keywords of this build:
        * `precision`: "mixed" (FP32-inner PCG with FP64 refinement, the
          default) or "fp64" (all-FP64 PCG).
        * `preconditioner`: "block_jacobi", "jacobi", "two_level"
          (block-Jacobi with an additive coarse correction on node
          aggregates and their rigid-body modes, in the place of the
          reference's AMG) or "auto" (two-level from 100,000 unknowns,
          block-Jacobi below). None takes default_preconditioner(), "auto"
          unless set_default_preconditioner changed it.
        * `inner_rtol` (mixed precision): the factor by which the FP32
          recursive residual drops between two reliable updates. None takes
          0.1 with block-Jacobi or Jacobi and 0.03 with the two-level
          preconditioner, whose solves take few iterations, so that the fixed
          cost of each refinement weighs more.
        """
        finite = getattr(self.material, "incremental", False)
        if nlgeom and not finite:
            raise NotImplementedError(
                "nlgeom is supported for hyperelastic and plastic materials in this build"
            )
        if alpha > 0.0:
            raise NotImplementedError("viscous stabilization is not part of this build")
        if self.n_state != 0 and not finite:
            raise NotImplementedError("this stateful material is not part of this build")
        if self._external_gradient.abs().max() > 0:
            raise NotImplementedError("external strains are not part of this build")
        if method not in (None, "cg", "minres"):
            if method in ("spsolve", "pardiso", "amgx"):
                raise RuntimeError(
                    f"method {method!r} is not available in this build: FEM solves run on the "
                    "hand-written PCG (method None, 'cg' or 'minres')"
                )
            raise ValueError(
                f"Method {method} is not supported. "
                "Choose from 'spsolve', 'minres', 'cg', 'pardiso', or 'amgx'."
            )
        precision = precision or default_precision()
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        preconditioner = self._resolve_pc(preconditioner)

        increments = torch.tensor([0.0, 1.0]) if increments is None else torch.as_tensor(increments)
        incs = [float(v) for v in increments]
        N = len(incs)

        if differentiable_parameters is None:
            params = ()
        elif isinstance(differentiable_parameters, Tensor):
            params = (differentiable_parameters,)
        else:
            params = tuple(differentiable_parameters)
        track = any(p.requires_grad for p in params)

        if finite:
            cfg = dict(precision=precision, pc=preconditioner, stol=stol, rtol=rtol, atol=atol,
                       max_iter=max_iter, inner_rtol=inner_rtol)
            self._K_cache = None
            u_all, f_all, q_all, g_all, s_all = solve_finite(
                self, incs, cfg, nlgeom, track, cutback_factor, growth_factor,
                max_cutbacks, aggregate_integration_points, verbose)
            if not track:
                u_all, f_all, q_all, g_all = (t.detach() for t in (u_all, f_all, q_all, g_all))
            if return_intermediate:
                return u_all, f_all, q_all, g_all, s_all
            return u_all[-1], f_all[-1], q_all[-1], g_all[-1], s_all[-1]

        kind, p0, p1, scale = self._material_params()
        cfg = dict(kind=kind, precision=precision, pc=preconditioner, stol=stol, rtol=rtol,
                   atol=atol, max_iter=max_iter, inner_rtol=inner_rtol)
        cached = self.cached_solve if use_cached_solve else None
        if cached is not None:
            cfg.update(cached=cached, update_cache=True, x0=cached.previous_x,
                       x0_adjoint=cached.previous_grad)
        self.last_solve_info = {}
        self._K_cache = None

        lam0 = incs[0]
        F_unit = self._neumann.reshape(-1)
        D_unit = self._dirichlet.reshape(-1)
        if not track:
            F_unit, D_unit = F_unit.detach(), D_unit.detach()
            p0 = p0.detach()
            p1 = p1.detach() if p1 is not None else None
            scale = scale.detach() if scale is not None else None

        # unit-load solutions: forces scale with the load factor, prescribed
        # values with its change since the first increment (base.py:761-770)
        with torch.set_grad_enabled(track):
            split = lam0 != 0.0 and bool(D_unit.abs().max() > 0) and bool(F_unit.abs().max() > 0)
            if split:
                uF = _StaticSolve.apply(self, cfg, F_unit, torch.zeros_like(D_unit), p0, p1, scale)
                uD = _StaticSolve.apply(self, cfg, torch.zeros_like(F_unit), D_unit, p0, p1, scale)
            else:
                uF = _StaticSolve.apply(self, cfg, F_unit, D_unit, p0, p1, scale)
                uD = None
            fF, gF, qF = _Integrate.apply(self, kind, uF, p0, p1, scale)
            if uD is not None:
                fD, gD, qD = _Integrate.apply(self, kind, uD, p0, p1, scale)

            nd = self.n_dof_per_node
            flux_shape = (self.n_elem, *self.n_flux)
            u_out, f_out, q_out, g_out = [], [], [], []
            init = self.initial_grad.to(self.device, torch.float64)
            d_only = (not split) and lam0 != 0.0 and bool(D_unit.abs().max() > 0)
            for n in range(N):
                # forces are totals of the load factor, prescribed values grow
                # with its change since the first increment
                a = 0.0 if n == 0 else (incs[n] - lam0 if d_only else incs[n])
                b = 0.0 if n == 0 else incs[n] - lam0
                u = a * uF
                f = a * fF
                g = a * gF
                q = a * qF
                if uD is not None:
                    u = u + b * uD
                    f = f + b * fD
                    g = g + b * gD
                    q = q + b * qD
                u_out.append(u.reshape(-1, nd))
                f_out.append(f.reshape(-1, nd))
                g_out.append(init + g.reshape(flux_shape))
                q_out.append(q.reshape(flux_shape))

        u_all = torch.stack(u_out)
        f_all = torch.stack(f_out)
        q_all = torch.stack(q_out)
        g_all = torch.stack(g_out)
        s_all = torch.zeros(N, self.n_elem, self.n_state, dtype=torch.float64, device=self.device)
        if not aggregate_integration_points:
            # a linear model from a zero state has the same flux at every
            # Gauss point only for constant-strain elements; report per point
            q_all, g_all = self._per_gp_outputs(u_all, init)
            s_all = torch.zeros(N, self.n_int, self.n_elem, self.n_state, dtype=torch.float64,
                                device=self.device)
        q_all = q_all.squeeze((-2, -1)) if q_all.dim() > 2 else q_all
        g_all = g_all.squeeze((-2, -1)) if g_all.dim() > 2 else g_all
        if not track:
            u_all, f_all, q_all, g_all = (t.detach() for t in (u_all, f_all, q_all, g_all))
        if verbose:
            self._print_report(precision, preconditioner)
        if return_intermediate:
            return u_all, f_all, q_all, g_all, s_all
        return u_all[-1], f_all[-1], q_all[-1], g_all[-1], s_all[-1]

    def integrate_mass(self) -> Tensor:
        """Consistent element mass matrices [n_elem, npe*ndof, npe*ndof]
        (base.py:376-396): sum over Gauss points of w detJ rho (thickness)
        N_I N_J per DOF component, one HIP block per element."""
        if self.material is None:
            raise ValueError("this model has no material")
        rho = torch.as_tensor(self.material.rho).to(self.device, torch.float64)
        rho = torch.broadcast_to(rho.reshape(-1) if rho.dim() else rho, (self.n_elem,)).contiguous()
        scale = self._scale_tensor()
        pat = self._pattern
        W = pat.npe * pat.ndof
        m = torch.empty(self.n_elem, W, W, dtype=torch.float64, device=self.device)
        bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        L.check(L.lib().tfm_element_mass(pat.stream(), C.byref(pat.mesh), L.ptr(rho),
                                         L.ptr(scale.detach() if scale is not None else None),
                                         L.ptr(m), L.ptr(bad)))
        self._check_jacobian(bad)
        return m

    def time_integration(
        self,
        t_output: Tensor | None = None,
        delta_t: float = 1.0e-1,
        max_iter: int = 100,
        verbose: bool = False,
        rtol: float = 1e-8,
        atol: float = 1e-6,
        stol: float = 1e-10,
        device: str | None = None,
        method: str | None = None,
        aggregate_integration_points: bool = True,
        use_cached_solve: bool = False,
        differentiable_parameters: Tensor | Iterable[Tensor] | None = None,
        precision: str | None = None,
        preconditioner: str | None = None,
        inner_rtol: float | None = None,
    ):
        """Implicit time integration from an equilibrium state
        (base.py:1288-1530): the state at t = 0 is the solution with every
        DOF at its prescribed value; each internal step of at most delta_t
        (the output times are hit exactly) solves
        M du + dt/2 (f_int_old + f_int(du) + f_ext) = 0 by Newton with
        (M + dt/2 K) du = -R on the node-block PCG, M being the consistent
        mass (integrate_mass, a HIP kernel) assembled like K. Returns
        temperature (or displacement), internal vector, flux, gradient and
        state at the output times. Gradients through the time steps are not
        part of this build."""
        t_output = torch.tensor([0.0, 1.0]) if t_output is None else torch.as_tensor(t_output)
        t_output = t_output.to(torch.float64).reshape(-1)
        if t_output.numel() == 0:
            raise ValueError("t_output must contain at least one time.")
        if t_output.min() < 0.0:
            raise ValueError("t_output must not contain negative times.")
        if (t_output[1:] <= t_output[:-1]).any():
            raise ValueError("t_output must be strictly increasing.")
        if differentiable_parameters is not None:
            params = ((differentiable_parameters,) if isinstance(differentiable_parameters, Tensor)
                      else tuple(differentiable_parameters))
            if any(p.requires_grad for p in params):
                raise NotImplementedError(
                    "gradients through time integration are not part of this build")
        if getattr(self.material, "incremental", False) or self.n_state != 0:
            raise NotImplementedError("time integration supports linear materials in this build")
        precision = precision or default_precision()
        preconditioner = self._resolve_pc(preconditioner)
        cfg = dict(precision=precision, pc=preconditioner, stol=stol, inner_rtol=inner_rtol)
        dev, dt64 = self.device, torch.float64
        nd = self.n_dof_per_node

        # equilibrium at t = 0 with every DOF prescribed
        bc = self._constraints.clone()
        self._constraints = torch.ones_like(bc)
        try:
            T0 = self.solve(precision=precision, stol=stol)[0].reshape(-1)
        finally:
            self._constraints = bc
        shape = (self.n_int, self.n_elem, *self.n_flux)
        zero_q = torch.zeros(shape, dtype=dt64, device=dev)
        g_init = self.initial_grad.to(dev, dt64).expand(shape).clone()
        s0 = torch.zeros(self.n_int, self.n_elem, self.n_state, dtype=dt64, device=dev)
        _, _, g0, q0, _ = self.integrate_material(None, g_init, zero_q, s0, T0, None, 0, False,
                                                  compute_stiffness=False)

        # internal times: every interval between output times in equal steps
        # of at most delta_t, starting at t = 0
        knots = [float(v) for v in t_output]
        if knots[0] > 0.0:
            knots = [0.0] + knots
        times, rows = [knots[0]], ([] if float(t_output[0]) > 0.0 else [0])
        for t0, t1 in zip(knots[:-1], knots[1:]):
            ratio = (t1 - t0) / delta_t
            n_sub = max(1, math.ceil(ratio - 1e-9 * max(1.0, ratio)))
            times += [t0 + (t1 - t0) * (k + 1) / n_sub for k in range(n_sub - 1)] + [t1]
            rows.append(len(times) - 1)

        con = self._con_mask().bool()
        f_ext = self._neumann.reshape(-1).to(dt64)
        m_el = self.integrate_mass()
        k_el = self.k0()
        from ._finite import TangentSystem

        M = TangentSystem(self, m_el, "fp64")  # M with the Dirichlet rows and columns
        systems: dict = {}
        u = [T0]
        f = [torch.zeros_like(T0)]
        q = [q0]
        g = [g0]
        st = [s0]
        t_start = time.perf_counter()
        info = []
        for n in range(1, len(times)):
            dt = times[n] - times[n - 1]
            key = round(dt, 15)
            if key not in systems:
                systems[key] = TangentSystem(self, m_el + 0.5 * dt * k_el, precision)
            A = systems[key]
            guess = u[-1].clone()
            res0 = None
            for it in range(max_iter):
                du = guess - u[-1]
                _, fe, g_new, q_new, s_new = self.integrate_material(
                    u[-1], g[-1], q[-1], st[-1], du, None, it, False, compute_stiffness=False)
                f_int = self.assemble_rhs(fe)
                Mdu = torch.empty_like(du)
                L.check(L.lib().tfm_spmv(self._pattern.stream(), C.byref(self._pattern.mesh), 0,
                                         L.ptr(M.val64), L.ptr(du.contiguous()), L.ptr(Mdu), 1))
                res = Mdu + 0.5 * dt * (f[-1] + f_int + f_ext)
                res = torch.where(con, torch.zeros_like(res), res)
                rn = float(torch.linalg.norm(res))
                res0 = rn if res0 is None else res0
                if rn < rtol * res0 or rn < atol:
                    break
                guess = guess + self._pcg(A, -res[None], cfg)[0]
            else:
                raise RuntimeError("Newton-Raphson iteration did not converge.")
            info.append(dict(time=times[n], newton_iterations=it, res=rn))
            if verbose:
                print(f"time step {n}: t = {times[n]:.6g} | newton {it} | res {rn:.3e}", flush=True)
            u.append(guess)
            f.append(f_int)
            q.append(q_new)
            g.append(g_new)
            st.append(s_new)
        self.last_solve_info = dict(time_integration=info, time=time.perf_counter() - t_start)
        idx = torch.tensor(rows)
        out_u = torch.stack(u)[idx].reshape(len(rows), -1, nd)
        out_f = torch.stack(f)[idx].reshape(len(rows), -1, nd)
        out_q = torch.stack(q)[idx]
        out_g = torch.stack(g)[idx]
        out_s = torch.stack(st)[idx]
        if aggregate_integration_points:
            out_q, out_g, out_s = out_q.mean(dim=1), out_g.mean(dim=1), out_s.mean(dim=1)
        return out_u, out_f, out_q.squeeze((-2, -1)), out_g.squeeze((-2, -1)), out_s

    def solve_load_cases(
        self,
        forces: Tensor,
        rtol: float = 1e-8,
        atol: float = 1e-6,
        stol: float = 1e-10,
        max_iter: int = 10,
        differentiable_parameters: Tensor | Iterable[Tensor] | None = None,
        precision: str | None = None,
        preconditioner: str | None = None,
        inner_rtol: float | None = None,
    ) -> tuple[Tensor, Tensor]:
        """Solves several load cases of one model at once (no counterpart in the
        reference, whose solve() takes one load state): `forces` [nrhs, n_nod,
        n_dof] with the model's own prescribed values and constraints. The
        cases share one assembled matrix and one batched multi-RHS PCG; their
        adjoints are solved together the same way. Returns (u, f), each
        [nrhs, n_nod, n_dof]."""
        if self.n_state != 0:
            raise NotImplementedError("stateful materials are not part of this build")
        forces = torch.as_tensor(forces).to(self.device)
        if forces.dim() != 3 or forces.shape[1:] != (self.n_nod, self.n_dof_per_node):
            raise ValueError("forces must have shape [n_cases, n_nod, n_dof_per_node]")
        if not 1 <= forces.shape[0] <= 16:
            raise ValueError("between 1 and 16 load cases can be solved at once")
        precision = precision or default_precision()
        if differentiable_parameters is None:
            params = ()
        elif isinstance(differentiable_parameters, Tensor):
            params = (differentiable_parameters,)
        else:
            params = tuple(differentiable_parameters)
        track = any(p.requires_grad for p in params)
        preconditioner = self._resolve_pc(preconditioner)
        kind, p0, p1, scale = self._material_params()
        cfg = dict(kind=kind, precision=precision, pc=preconditioner, stol=stol, rtol=rtol,
                   atol=atol, max_iter=max_iter, inner_rtol=inner_rtol)
        self.last_solve_info = {}
        self._K_cache = None
        F = forces.reshape(forces.shape[0], -1).to(torch.float64)
        D = self._dirichlet.reshape(-1)
        if not track:
            F, D, p0 = F.detach(), D.detach(), p0.detach()
            p1 = p1.detach() if p1 is not None else None
            scale = scale.detach() if scale is not None else None
        with torch.set_grad_enabled(track):
            u = _StaticSolve.apply(self, cfg, F, D, p0, p1, scale)
            f, _, _ = _Integrate.apply(self, kind, u, p0, p1, scale)
        shape = (forces.shape[0], self.n_nod, self.n_dof_per_node)
        u, f = u.reshape(shape), f.reshape(shape)
        if not track:
            u, f = u.detach(), f.detach()
        return u, f

    def _per_gp_outputs(self, u_all: Tensor, init: Tensor):
        kind, p0, p1, scale = self._material_params()
        h = _MatHolder(kind, p0.detach(), p1.detach() if p1 is not None else None,
                       scale.detach() if scale is not None else None)
        qs, gs = [], []
        shape = (self.n_int, self.n_elem, *self.n_flux)
        for u in u_all:
            _, _, _, _, gg, qg = self._forces(h.mat, u.detach(), per_gp=True)
            gs.append(init + gg.reshape(shape))
            qs.append(qg.reshape(shape))
        return torch.stack(qs), torch.stack(gs)

    def _print_report(self, precision, pc):
        info = self.last_solve_info.get("forward", {})
        its = [p.get("iterations") for p in info.get("pcg", [])]
        print(f"torch-fem (MI355X) | {type(self).__name__} | {self.n_elem:,} elem | "
              f"{self.n_dofs:,} dof | {precision} PCG ({pc}) | newton {info.get('newton_iterations')}"
              f" | pcg iterations {its} | res {info.get('res', float('nan')):.3e}")


class _Gather(torch.autograd.Function):
    """assemble_rhs with autograd: F = sum over elements; dF/df = index select."""

    @staticmethod
    def forward(ctx, fem, f):
        ctx.fem = fem
        return fem.assemble_rhs(f.detach())

    @staticmethod
    def backward(ctx, g):
        fem = ctx.fem
        return None, g[fem.idx.long()]


class Mechanics(FEM, ABC):
    """Solid and structural mechanics (base.py:926-1126)."""

    @property
    def n_dof_per_node(self) -> int:
        return self.nodes.shape[1]

    @property
    def initial_grad(self) -> Tensor:
        return torch.eye(self.n_flux[0], dtype=torch.float64, device=self.device)

    @property
    def forces(self) -> Tensor:
        return self._neumann

    @forces.setter
    def forces(self, value: Tensor):
        if not value.shape == (self.n_nod, self.n_dof_per_node):
            raise ValueError("Forces must have the same shape as nodes.")
        if not torch.is_floating_point(value):
            raise TypeError("Forces must be a floating-point tensor.")
        self._neumann = value.to(self.device)

    @property
    def displacements(self) -> Tensor:
        return self._dirichlet

    @displacements.setter
    def displacements(self, value: Tensor):
        if not value.shape == (self.n_nod, self.n_dof_per_node):
            raise ValueError("Displacements must have the same shape as nodes.")
        if not torch.is_floating_point(value):
            raise TypeError("Displacements must be a floating-point tensor.")
        self._dirichlet = value.to(self.device)

    @property
    def ext_strain(self) -> Tensor:
        return self._external_gradient

    @ext_strain.setter
    def ext_strain(self, value: Tensor):
        if not value.shape == (self.n_elem, self.n_dof_per_node, self.n_dim):
            raise ValueError("External strain must have the same shape as strains.")
        if not torch.is_floating_point(value):
            raise TypeError("External strain must be a floating-point tensor.")
        self._external_gradient = value.to(self.device)


class Heat(FEM, ABC):
    """Steady heat conduction (base.py:1129-1286)."""

    @property
    def n_dof_per_node(self) -> int:
        return 1

    @property
    def n_flux(self) -> list[int]:
        return [1, self.n_dim]

    @property
    def initial_grad(self) -> Tensor:
        return torch.zeros(1, dtype=torch.float64, device=self.device)

    @property
    def heat_flux(self) -> Tensor:
        return self._neumann

    @heat_flux.setter
    def heat_flux(self, value: Tensor):
        if not value.shape == (self.n_nod, 1):
            raise ValueError("Heat flux must have the same shape as nodes.")
        if not torch.is_floating_point(value):
            raise TypeError("Heat flux must be a floating-point tensor.")
        self._neumann = value.to(self.device)

    @property
    def temperatures(self) -> Tensor:
        return self._dirichlet

    @temperatures.setter
    def temperatures(self, value: Tensor):
        if not value.shape == (self.n_nod, self.n_dof_per_node):
            raise ValueError("Temperatures must have the same shape as nodes.")
        if not torch.is_floating_point(value):
            raise TypeError("Temperatures must be a floating-point tensor.")
        self._dirichlet = value.to(self.device)
