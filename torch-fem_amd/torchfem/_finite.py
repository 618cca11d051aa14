"""Incremental Newton path of FEM.solve for hyperelastic and plastic materials.

The reference solves every model by load increments with a Newton loop whose
residual is integrate_material + assemble_rhs and whose tangent is
integrate_material's k assembled by assemble_matrix (base.py:700-880,
982-1092; sparse.py:541-639); its gradients come from the implicit adjoint of
each converged increment, chained through the previous increment's state
(NewtonRaphsonAdjoint.backward, sparse.py:641-724).

Here the kinematics and the assembly of every Newton evaluation are HIP
kernels: the displacement gradients per Gauss point and the internal forces
int B^T P are the internal-force kernel (tfm_internal_force, with the material
contribution switched off and P passed as the previous flux), the element
matrices are the per-point tangent kernel (tfm_element_matrices_gp) and the
CSR values, FP32 copy and nodal diagonal blocks come from the deterministic
assembly (tfm_assemble_elements); the Newton corrections and the adjoints are
the library's PCG, or its MINRES where the tangent is indefinite. Only psi and its derivatives run in torch, as the user
wrote them. The two HIP operators carry their own adjoints (each one is the
other's transpose up to the quadrature weights), so the residual stays
differentiable with respect to the previous deformation gradient, the loads
and the material parameters without a dense B in memory.
"""

from __future__ import annotations

import ctypes as C
import math
import time

import torch
from torch import Tensor

from . import _lib as L


class Kinematics:
    """HIP operators of one model: u -> H_g = u (x) grad N per Gauss point and
    P_g -> F_int = sum_g w_g detJ_g B_g^T P_g, and per-point tangents ->
    element matrices."""

    def __init__(self, fem):
        from .base import _MatHolder

        self.fem = fem
        pat = fem._pattern
        dev = fem.device
        E = fem.n_elem
        scale = fem._scale_tensor()
        zero = torch.zeros(E, dtype=torch.float64, device=dev)
        # an isotropic material of zero moduli: the kernel's flux is then the
        # previous flux alone, so it integrates a given stress field
        self._zero = _MatHolder(L.MAT_ISO, zero, zero, scale)
        _, _, detJ = fem.eval_shape_functions(fem.etype.ipoints)
        w = fem.etype.iweights.to(device=dev, dtype=torch.float64)
        wd = detJ * w[:, None]
        if scale is not None:
            wd = wd * scale[None, :]
        self.wdet = wd.contiguous()  # [n_gp, n_elem]
        self._u0 = torch.zeros(fem.n_dofs, dtype=torch.float64, device=dev)
        self.shape = (fem.n_int, E, pat.ndof, pat.dim)

    def grad(self, u: Tensor) -> Tensor:
        """Displacement gradients H [n_gp, n_elem, ndof, dim] of nodal values u."""
        _, _, _, _, gg, _ = self.fem._forces(self._zero.mat, u.detach(), per_gp=True)
        return gg.reshape(self.shape)

    def force(self, P: Tensor) -> Tensor:
        """Assembled internal forces [n_dofs] of a per-point stress [n_gp, n_elem, ndof, dim]."""
        F, _, _, _, _, _ = self.fem._forces(self._zero.mat, self._u0,
                                            flux_prev=P.detach().contiguous())
        return F

    def element_forces(self, P: Tensor) -> Tensor:
        """Element force vectors [n_elem, npe*ndof] of a per-point stress."""
        _, work, _, _, _, _ = self.fem._forces(self._zero.mat, self._u0,
                                               flux_prev=P.detach().contiguous())
        return work.reshape(self.fem.n_elem, -1)

    def matrices(self, A: Tensor) -> Tensor:
        """Element matrices [n_elem, W, W] of per-point tangents [n_gp, n_elem, 3, 3, 3, 3]."""
        fem = self.fem
        pat = fem._pattern
        A = A.detach().to(torch.float64).contiguous()
        W = pat.npe * pat.ndof
        k = torch.empty(fem.n_elem, W, W, dtype=torch.float64, device=fem.device)
        bad = torch.zeros(1, dtype=torch.int32, device=fem.device)
        scale = fem._scale_tensor()
        L.check(L.lib().tfm_element_matrices_gp(pat.stream(), C.byref(pat.mesh), L.ptr(A),
                                                L.ptr(scale), L.ptr(k), L.ptr(bad)))
        fem._check_jacobian(bad)
        return k


class _GradOp(torch.autograd.Function):
    """H = grad(u) on the HIP kernel; its adjoint sum_g B_g^T Hbar_g is the
    force kernel applied to Hbar_g / (w_g detJ_g)."""

    @staticmethod
    def forward(ctx, kin: Kinematics, u: Tensor) -> Tensor:
        ctx.kin = kin
        return kin.grad(u)

    @staticmethod
    def backward(ctx, gH):
        kin = ctx.kin
        return None, kin.force(gH / kin.wdet[:, :, None, None])


class _ForceOp(torch.autograd.Function):
    """F_int = sum_g w_g detJ_g B_g^T P_g on the HIP kernel; its adjoint is
    w_g detJ_g grad(Fbar)_g."""

    @staticmethod
    def forward(ctx, kin: Kinematics, P: Tensor) -> Tensor:
        ctx.kin = kin
        return kin.force(P)

    @staticmethod
    def backward(ctx, gF):
        kin = ctx.kin
        return None, kin.grad(gF.contiguous()) * kin.wdet[:, :, None, None]


class TangentSystem:
    """Assembled Newton tangent: FP64 (and FP32 for mixed precision) CSR
    values with the Dirichlet rows and columns replaced by the identity, and
    the nodal diagonal blocks, in the layout the PCG reads."""

    def __init__(self, fem, k: Tensor, precision: str):
        pat = fem._pattern
        dev = fem.device
        self.fem = fem
        self.con = fem._con_mask()
        self.val64 = torch.empty(pat.nnz, dtype=torch.float64, device=dev)
        self.val32 = (torch.empty(pat.nnz, dtype=torch.float32, device=dev)
                      if precision == "mixed" else None)
        self.diag = torch.empty(pat.n_nod * pat.ndof**2, dtype=torch.float64, device=dev)
        L.check(L.lib().tfm_assemble_elements(pat.stream(), C.byref(pat.mesh), L.ptr(k),
                                              L.ptr(self.con), L.ptr(self.val64),
                                              L.ptr(self.val32), L.ptr(self.diag),
                                              L.FP32_DIFF))
        self.sys = L.System(mesh=C.pointer(pat.mesh), mat=None, con=self.con.data_ptr(),
                            val64=L.ptr(self.val64), val32=L.ptr(self.val32),
                            diag_blocks=self.diag.data_ptr(), val32_form=L.FP32_DIFF)


class FiniteStrainRun:
    """State of one finite-strain solve: the model, its operators, the
    constraint mask, the solver settings and the last assembled tangent."""

    def __init__(self, fem, cfg: dict, nlgeom: bool, verbose: bool = False):
        self.fem = fem
        self.verbose = verbose
        self.mat = fem.material
        self.kin = Kinematics(fem)
        self.con = fem.constraints.reshape(-1).to(fem.device)
        self.cfg = cfg
        self.nlgeom = nlgeom
        self.system: TangentSystem | None = None
        self.log: list[dict] = []
        self._krylov = None
        self._csr = None
        self.last_iter = 0

    def solve_linear(self, system: TangentSystem, rhs: Tensor,
                     adjoint: bool = False) -> tuple[Tensor, dict]:
        """K x = rhs with the configured PCG; a tangent that PCG cannot solve
        (indefinite, as a Neo-Hookean tangent becomes where lam ln J > mu)
        goes to the library's MINRES on the same FP64 values, which needs
        symmetry only (the reference's direct or MINRES solves serve the same
        role, sparse.py:271-347)."""
        fem = self.fem
        try:
            x = fem._pcg(system, rhs.reshape(1, -1), self.cfg, adjoint=adjoint)[0]
            return x, dict(fem._last_pcg, solver="pcg")
        except RuntimeError:
            pass
        from .sparse import KrylovSolver

        if self._krylov is None:
            rows, cols = fem.glob_idx
            n = fem.n_dofs
            rowptr = torch.zeros(n + 1, dtype=torch.int64, device=fem.device)
            rowptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
            self._csr = (rowptr.contiguous(), cols.to(torch.int32).contiguous())
            self._krylov = KrylovSolver(n, 1, fem.device)
        b = rhs.reshape(-1).to(torch.float64).contiguous()
        x = torch.zeros_like(b)
        info = self._krylov.solve("minres", *self._csr, system.val64, b, x,
                                  (0.7 if adjoint else 0.25) * self.cfg["stol"],
                                  max(1000, 20 * fem.n_dofs))
        return x, dict(info, solver="minres")

    def close(self) -> None:
        if self._krylov is not None:
            self._krylov.close()
            self._krylov = None

    def update(self, H: Tensor, grad_prev: Tensor, flux_prev: Tensor, state_prev: Tensor,
               it: int, tangent: bool = True, mat=None):
        """Material update at every Gauss point from the increment H = grad du:
        (stress, state, tangent or None). A hyperelastic material evaluates
        P(F) and dP/dF at F = grad_prev + H; other incremental materials run
        their step from (grad_prev, flux_prev, state_prev) (materials/base.py:48-91)."""
        mat = self.mat if mat is None else mat
        if getattr(mat, "finite_strain", False):
            F = grad_prev + H
            if tangent:
                P, A = mat.stress_tangent(F)
                return P, state_prev, A
            return mat.stress(F), state_prev, None
        return mat.step(H, grad_prev, flux_prev, state_prev, None, None, it)

    def residual(self, du: Tensor, grad_prev: Tensor, flux_prev: Tensor, state_prev: Tensor,
                 F_ext: Tensor, DU: Tensor, it: int = 0, mat=None) -> Tensor:
        """Differentiable residual F_int(du with du[con] = DU[con]) - F_ext,
        zero on constrained DOFs (the eval_residual of base.py:707-740)."""
        du_bc = torch.where(self.con, DU, du)
        H = _GradOp.apply(self.kin, du_bc)
        P, _, _ = self.update(H, grad_prev, flux_prev, state_prev, it, tangent=False, mat=mat)
        F_int = _ForceOp.apply(self.kin, P)
        return torch.where(self.con, torch.zeros_like(F_int), F_int - F_ext)

    def newton(self, du: Tensor, grad_prev: Tensor, flux_prev: Tensor, state_prev: Tensor,
               F_ext: Tensor, DU: Tensor) -> Tensor:
        """Newton iterations du -= K^{-1} R until |R| < rtol |R_0| or |R| < atol
        (NewtonRaphsonAdjoint.forward, sparse.py:572-623). The tangent is
        reassembled at every evaluation with nlgeom or an internal state and
        kept from the first evaluation of the solve otherwise
        (base.py:1031-1035)."""
        cfg, kin, fem = self.cfg, self.kin, self.fem
        t0 = time.perf_counter()
        du = du.detach().clone()
        zero = torch.zeros_like(du)
        res0 = None
        pcg = []
        hist = []
        for i in range(cfg["max_iter"]):
            du_bc = torch.where(self.con, DU, du)
            need_k = self.system is None or self.nlgeom or self.mat.n_state > 0
            P, _, A = self.update(kin.grad(du_bc), grad_prev, flux_prev, state_prev, i,
                                  tangent=need_k)
            if need_k:
                self.system = TangentSystem(fem, kin.matrices(A), cfg["precision"])
            self.last_iter = i
            res = torch.where(self.con, zero, kin.force(P) - F_ext)
            rn = float(torch.linalg.norm(res))
            hist.append(rn)
            if i == 0:
                res0 = rn
            if rn < cfg["rtol"] * res0 or rn < cfg["atol"]:
                self.log.append(dict(newton_iterations=i, res0=res0, res=rn, pcg=pcg,
                                     time=time.perf_counter() - t0))
                return du
            if not math.isfinite(rn):
                break
            d, info = self.solve_linear(self.system, res)
            pcg.append(info)
            if self.verbose:
                its = info["iterations"]
                print(f"  newton {i}: |R| {rn:.3e}, {info['solver']} "
                      f"{its[0] if isinstance(its, list) else its} iterations", flush=True)
            du = du - d
        raise RuntimeError("Newton-Raphson iteration did not converge. Residual norms: "
                           + ", ".join(f"{r:.3e}" for r in hist[-4:]))


class NewtonStep(torch.autograd.Function):
    """One converged increment
    du = Newton(du0; grad_prev, flux_prev, state_prev, F_ext, DU, material)
    with the implicit adjoint of NewtonRaphsonAdjoint.backward
    (sparse.py:672-724): K lambda = dL/ddu with the last assembled tangent,
    then -lambda pulled back through the recomputed residual to the previous
    deformation gradient, stress and state, the loads and the tensors the
    material's update reads (its `tensors()`). Those tensors are inputs of
    this node, so autograd carries their gradients on to whatever they were
    computed from (the user's differentiable parameters); the pull-back
    itself runs on detached copies and never re-enters the caller's graph."""

    @staticmethod
    def forward(ctx, run: FiniteStrainRun, names: tuple, du0, grad_prev, flux_prev, state_prev,
                F_ext, DU, *mat_tensors):
        du = run.newton(du0, grad_prev, flux_prev, state_prev, F_ext, DU)
        ctx.run = run
        ctx.names = names
        ctx.system = run.system
        ctx.it = run.last_iter
        ctx.save_for_backward(du, grad_prev, flux_prev, state_prev, F_ext, DU, *mat_tensors)
        return du

    @staticmethod
    def backward(ctx, gdu):
        run = ctx.run
        du, *prev = ctx.saved_tensors[:6]
        mt = ctx.saved_tensors[6:]
        lam, _ = run.solve_linear(ctx.system, gdu, adjoint=True)
        needs = ctx.needs_input_grad
        with torch.enable_grad():
            locs = [t.detach().requires_grad_(True) for t in prev]
            mloc = [t.detach().requires_grad_(True) for t in mt]
            mat = run.mat.replace(**dict(zip(ctx.names, mloc)))
            r = run.residual(du.detach(), *locs, it=ctx.it, mat=mat)
            grads = torch.autograd.grad(r, locs + mloc, grad_outputs=-lam, allow_unused=True)
        return (None, None, None,
                *[g if needs[3 + k] else None for k, g in enumerate(grads)])


def solve_finite(fem, incs: list[float], cfg: dict, nlgeom: bool, track: bool,
                 cutback_factor: float, growth_factor: float, max_cutbacks: int,
                 aggregate: bool, verbose: bool):
    """Incremental loading with automatic cutback (base.py:742-880): every
    substep runs NewtonStep from the previous substep's increment, then
    evaluates the converged state; results are stored at the requested
    increments."""
    run = FiniteStrainRun(fem, cfg, nlgeom, verbose)
    kin = run.kin
    dev = fem.device
    dt = torch.float64
    N = len(incs)
    nd = fem.n_dof_per_node
    con = run.con
    F_unit = fem._neumann.reshape(-1)
    D_unit = fem._dirichlet.reshape(-1)
    if not track:
        F_unit, D_unit = F_unit.detach(), D_unit.detach()

    grad0 = fem.initial_grad.to(dev, dt).expand(fem.n_int, fem.n_elem, *fem.n_flux)
    u_cur = torch.zeros(fem.n_dofs, dtype=dt, device=dev)
    f_cur = torch.zeros(fem.n_dofs, dtype=dt, device=dev)
    grad_cur = grad0.clone()
    flux_cur = torch.zeros_like(grad_cur)
    state_cur = torch.zeros(fem.n_int, fem.n_elem, fem.n_state, dtype=dt, device=dev)
    du = torch.zeros(fem.n_dofs, dtype=dt, device=dev)
    out_u, out_f, out_q, out_g, out_s = [u_cur], [f_cur], [flux_cur], [grad_cur], [state_cur]

    mt = run.mat.tensors()
    names, mat_tensors = tuple(mt), tuple(mt.values())
    lam = incs[0]
    step_frac = 1.0
    t0 = time.perf_counter()
    with torch.set_grad_enabled(track):
        for n in range(1, N):
            target = incs[n]
            span = target - lam
            direction = math.copysign(1.0, span)
            step_size = step_frac * abs(span)
            min_step = abs(span) * cutback_factor**max_cutbacks
            cutbacks = 0
            while abs(target - lam) > 1e-12 * max(1.0, abs(target)):
                step = direction * min(step_size, abs(target - lam))
                F_ext = (lam + step) * F_unit
                DU = step * D_unit
                try:
                    du = NewtonStep.apply(run, names, du.detach(), grad_cur, flux_cur,
                                          state_cur, F_ext, DU, *mat_tensors)
                except RuntimeError as err:
                    step_size = cutback_factor * abs(step)
                    cutbacks += 1
                    if step_size < min_step:
                        raise RuntimeError(
                            f"Newton-Raphson did not converge in increment {n} "
                            f"after {max_cutbacks} cutbacks."
                        ) from err
                    continue
                # converged state: forces, deformation gradient, stress, state
                du_eval = torch.where(con, DU, du)
                H = _GradOp.apply(kin, du_eval)
                P, state_cur, _ = run.update(H, grad_cur, flux_cur, state_cur, cfg["max_iter"],
                                             tangent=False)
                F = grad_cur + H
                f_cur = _ForceOp.apply(kin, P)
                u_cur = u_cur + du_eval
                grad_cur = F
                if nlgeom:
                    flux_cur = F @ P / torch.linalg.det(F)[..., None, None]
                else:
                    flux_cur = P
                du = du_eval
                lam += step
                step_size = min(growth_factor * step_size, abs(span))
            step_frac = min(step_size / abs(span), 1.0) if span != 0.0 else 1.0
            out_u.append(u_cur)
            out_f.append(f_cur)
            out_q.append(flux_cur)
            out_g.append(grad_cur)
            out_s.append(state_cur)
            if verbose:
                last = run.log[-1] if run.log else {}
                print(f"increment {n}: load factor {lam:.6g} | newton "
                      f"{last.get('newton_iterations')} | res {last.get('res', float('nan')):.3e}"
                      f" | cutbacks {cutbacks} | {time.perf_counter() - t0:.2f} s", flush=True)

    u_all = torch.stack(out_u).reshape(N, -1, nd)
    f_all = torch.stack(out_f).reshape(N, -1, nd)
    q_all = torch.stack(out_q)
    g_all = torch.stack(out_g)
    s_all = torch.stack(out_s)
    if aggregate:
        q_all, g_all, s_all = q_all.mean(dim=1), g_all.mean(dim=1), s_all.mean(dim=1)
    run.close()  # a backward that needs MINRES opens its own workspace
    fem._last_system = None
    fem._K_cache = None
    if run.system is not None:
        fem._finite_system = run.system
    fem.last_solve_info = dict(forward=dict(increments=run.log,
                                            time=time.perf_counter() - t0))
    return u_all, f_all, q_all, g_all, s_all
