"""Benchmark: assembly + solve + adjoint gradient on the Hexa1 64^3 cube.

Workload (one step): a Hexa1 cube on a structured 64^3 mesh over [0,1]^3
(274,625 nodes, 823,875 DOFs), isotropic elasticity E=1000, nu=0.3 with a
per-element modulus tensor, z=0 clamped, u_z=-0.01 prescribed on z=1, 2x2x2
Gauss points. One step is FEM.solve() (fused integration + assembly, PCG
solve to 1e-8 relative residual) followed by the adjoint gradient of sum(u)
with respect to the per-element moduli and the nodal forces (the published
cube benchmark differentiates sum(u) w.r.t. forces). The pattern setup is
timed separately, as in the reference's benchmark protocol.

Prints one JSON line (stdout); progress goes to stderr.
"""

from __future__ import annotations

import argparse
import contextlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "torch-fem_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = "solve+adjoint-grad time & DOFs/s, Hexa1 cube, mixed-prec vs FP64 @1e-8 rel"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per cube edge")
    ap.add_argument("--stol", type=float, default=1e-8)
    ap.add_argument("--preconditioner", default="block_jacobi")
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="mixed precision: residual drop between reliable updates "
                         "(default: 0.1 with block-Jacobi, 0.03 with two-level)")
    ap.add_argument("--fp64-steps", type=int, default=5, help="0 skips the FP64 timing")
    ap.add_argument("--two-level-steps", type=int, default=5,
                    help="timed mixed steps with preconditioner='two_level' beside the "
                         "block-Jacobi headline (0 skips them)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-stride", type=int, default=32,
                    help="HIP-event timing of the SpMV kernel on every k-th PCG iteration "
                         "(each timed launch adds an event pair to the stream; 0 disables)")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the tight FP64 solve used for the accuracy report")
    ap.add_argument("--workload", default="hexa1_64",
                    choices=["hexa1_64", "simp48", "tetra80", "hexa2_32x16", "hyper105"],
                    help="hexa1_64: the headline Hexa1 64^3 cube (default); the others time the "
                         "SIMP, thermal and multi-load-case problems the same way (no roofline "
                         "or CPU baseline)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB buffer 5 times (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_probe(mib: int) -> None:
    """Five device-to-device copies of a known size: the reference that scales
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 5}), flush=True)


def build_model(n: int):
    import torch

    from torchfem import Solid
    from torchfem.materials import IsotropicElasticity3D
    from torchfem.mesh import cube_hexa

    nodes, elements = cube_hexa(n + 1, n + 1, n + 1)
    t0 = time.perf_counter()
    model = Solid(nodes.cuda(), elements.cuda(), IsotropicElasticity3D(1000.0, 0.3))
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    X = model.nodes
    model.constraints[X[:, 2] == 0.0, :] = True
    model.constraints[X[:, 2] == 1.0, 2] = True
    model.displacements[X[:, 2] == 1.0, 2] = -0.01
    return model, setup


def make_step(model, precision: str, stol: float, pc: str, inner_rtol: float | None):
    import torch

    from torchfem.materials import IsotropicElasticity3D

    E0 = torch.full((model.n_elem,), 1000.0, dtype=torch.float64, device=model.device)

    def step():
        E = E0.clone().requires_grad_(True)
        F = torch.zeros(model.n_nod, 3, dtype=torch.float64, device=model.device, requires_grad=True)
        model.material = IsotropicElasticity3D(E, 0.3)
        model.forces = F
        u, f, sigma, _, _ = model.solve(differentiable_parameters=(E, F), precision=precision,
                                        stol=stol, preconditioner=pc, inner_rtol=inner_rtol)
        gE, gF = torch.autograd.grad(u.sum(), (E, F))
        return u.detach(), gE, gF, f.detach(), sigma.detach()

    return step


def measured_traffic():
    """HBM bytes per launch of the FP32 SpMV kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a copy of known size);
    None when that summary is absent."""
    path = ROOT / "profiles" / "pcg_hexa1_64.md"
    try:
        for line in path.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def timed(step, steps: int, warmup: int, label: str):
    import torch

    for i in range(warmup):
        step()
        torch.cuda.synchronize()
        log(f"{label}: warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for i in range(steps):
        out = step()
        log(f"{label}: step {i + 1}/{steps} done")
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(n: int, iters: tuple[int, int], slab: int = 8) -> dict:
    """The CPU oracle (FP64 torch einsum integration, index_add_ assembly,
    SciPy Jacobi-CG, autograd implicit adjoint) priced on the Hexa1 n^3
    headline workload from a bounded sample of it.

    A full oracle run at 64^3 takes about ten minutes (most of it SciPy CG on
    65M nonzeros), so the sample is an n x n x `slab` slab of the same cube
    (the same elements, material and solver), whose phases are timed and
    scaled to the full cube the way they grow: integration, assembly,
    outputs and the backward's residual pass by the element count, one
    Jacobi-CG iteration by the number of stored entries, and the iterations
    by the counts of the same Jacobi-PCG (FP64, same stol) on the full cube,
    `iters` = (forward, adjoint), measured on the GPU. The timing split
    follows the reference's benchmarks/utils.py:47-82 (setup excluded)."""
    import numpy as np
    import scipy.sparse as sp
    import torch

    import oracle as O

    threads = min(os.cpu_count() or 1, 16)  # the host share of one GPU
    torch.set_num_threads(threads)
    nodes, el = O.cube_hexa(n + 1, n + 1, slab + 1)
    nodes[:, 2] *= slab / n  # a slab of the unit cube: cells of edge 1/n
    E = torch.full((len(el),), 1000.0, dtype=torch.float64, requires_grad=True)
    m = O.OracleModel(nodes, el, O.Hexa1, "solid", O.iso3d_C(E, torch.tensor(0.3, dtype=torch.float64)))
    con = torch.zeros(len(nodes), 3, dtype=torch.bool)
    con[nodes[:, 2] == 0.0, :] = True
    cidx = torch.nonzero(con.ravel()).ravel()
    u = 1e-3 * torch.randn(m.n_dofs, generator=torch.Generator().manual_seed(0), dtype=torch.float64)

    def best(fn, reps=2):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return min(ts), out

    # forward: integration + assembly (base.py:1031-1090, 398-426)
    def assemble():
        with torch.no_grad():
            k, _, _, _ = m.integrate(u, need_k=True)
            return m.csr(m.assemble_matrix(k, cidx))
    t_asm, K = best(assemble)
    # outputs: one integration pass without the stiffness (base.py:844-867)
    t_out, _ = best(lambda: m.integrate(u.detach(), need_k=False))
    # backward: the residual re-evaluation with autograd (sparse.py:687-717)
    def residual_grad():
        with torch.enable_grad():
            # the tangent is rebuilt from E per pass, as its graph is freed by grad
            m.mat = O.iso3d_C(E, torch.tensor(0.3, dtype=torch.float64))
            _, f, _, _ = m.integrate(u, need_k=False)
            return torch.autograd.grad(m.assemble_rhs(f).sum(), E)
    t_bwd, _ = best(residual_grad)
    # one Jacobi-CG iteration of SciPy (sparse.py:406-421 / 505-512): timed
    # as a fixed number of iterations of scipy's cg
    Kc = K.tocsr()
    Minv = sp.diags(1.0 / Kc.diagonal())
    b = np.ones(m.n_dofs)
    n_it = 30
    def cg_iters():
        import scipy.sparse.linalg as spla
        spla.cg(Kc, b, rtol=1e-30, atol=0.0, M=Minv, maxiter=n_it)
    t_cg, _ = best(cg_iters)
    t_iter = t_cg / n_it

    el_scale = n / slab
    nnz_full = 9 * (3 * (n + 1) - 2) ** 3  # the Hexa1 cube's pattern size (closed form)
    nz_scale = nnz_full / Kc.nnz
    dofs = 3 * (n + 1) ** 3
    fwd = el_scale * (t_asm + t_out) + iters[0] * t_iter * nz_scale
    bwd = el_scale * t_bwd + iters[1] * t_iter * nz_scale
    return {
        "value": dofs / (fwd + bwd),
        "unit": "DOFs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"Hexa1 cube {n}^3 ({dofs} DOFs) priced from a {n}x{n}x{slab} slab of it "
                  f"({len(el)} elements, nnz {Kc.nnz}): integration+assembly {t_asm:.2f}s, "
                  f"outputs {t_out:.2f}s, backward residual pass {t_bwd:.2f}s (x{el_scale:g} "
                  f"elements), SciPy Jacobi-CG {t_iter * 1e3:.1f} ms/iteration (x{nz_scale:.2f} "
                  f"nnz) x {iters[0]} + {iters[1]} iterations (FP64 Jacobi-PCG on the full cube, "
                  f"same stol); estimate fwd {fwd:.1f}s bwd {bwd:.1f}s; {threads} torch threads "
                  f"(SciPy's SpMV runs on one) on {cpu_model()}",
    }


def other_workload(args) -> dict:
    """Mixed-precision and FP64 step times of the other problems of the path,
    with the errors of the mixed results against an FP64 solve at 1e-11. A
    step is FEM.solve() (or solve_load_cases) plus the adjoint gradient."""
    import torch

    from torchfem import Solid, SolidHeat
    from torchfem.elements import linear_to_quadratic
    from torchfem.materials import IsotropicConductivity3D, IsotropicElasticity3D
    from torchfem.mesh import cube_hexa, cube_tetra6

    dt = torch.float64
    nrhs = 1
    if args.workload == "simp48":
        # Hexa1 48^3, rho ~ U(0.3, 1) (seed 0), E = 1000 rho^3, clamped at x=0,
        # total -z load 1 split equally over the x=1 face; dC/drho of C = F.u
        nodes, el = cube_hexa(49, 49, 49)
        rho0 = (0.3 + 0.7 * torch.rand(len(el), generator=torch.Generator().manual_seed(0),
                                       dtype=dt)).cuda()
        model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
        X = model.nodes
        model.constraints[X[:, 0] == 0.0, :] = True
        face = X[:, 0] == 1.0
        model.forces[face, 2] = -1.0 / int(face.sum())
        F = model.forces.clone()
        desc = "SIMP Hexa1 48^3, compliance gradient w.r.t. rho"

        def step(precision, stol):
            rho = rho0.clone().requires_grad_(True)
            model.material = IsotropicElasticity3D(1000.0 * rho**3, 0.3)
            u, _, _, _, _ = model.solve(differentiable_parameters=rho, precision=precision,
                                        stol=stol, preconditioner=args.preconditioner,
                                        inner_rtol=args.inner_rtol)
            (g,) = torch.autograd.grad(torch.inner(F.ravel(), u.ravel()), rho)
            return u.detach(), g
    elif args.workload == "tetra80":
        # Tetra1 Kuhn split of 80^3 cells, jitter 0.1 h (seed 0), kappa =
        # exp(N(0, 0.5^2)) (seed 1), T = 0 at x=0 and 1 at x=1; d(mean T)/d kappa
        nodes, el = cube_tetra6(80, jitter=0.1, seed=0)
        k0 = torch.exp(0.5 * torch.randn(len(el), generator=torch.Generator().manual_seed(1),
                                         dtype=dt)).cuda()
        model = SolidHeat(nodes.cuda(), el.cuda(), IsotropicConductivity3D(1.0))
        X = model.nodes
        model.constraints[X[:, 0] == 0.0] = True
        model.constraints[X[:, 0] == 1.0] = True
        model.temperatures[X[:, 0] == 1.0] = 1.0
        desc = "thermal Tetra1 80^3 x 6, gradient of mean T w.r.t. kappa"

        def step(precision, stol):
            kap = k0.clone().requires_grad_(True)
            model.material = IsotropicConductivity3D(kap)
            T, _, _, _, _ = model.solve(differentiable_parameters=kap, precision=precision,
                                        stol=stol, preconditioner=args.preconditioner,
                                        inner_rtol=args.inner_rtol)
            (g,) = torch.autograd.grad(T.mean(), kap)
            return T.detach(), g
    else:
        # Hexa2 32^3, clamped at z=0, 16 load cases: tractions ~N(0,1) per face
        # of z=1 (seed k) as consistent nodal loads of unit total magnitude;
        # one batched solve and one batched adjoint of sum_k C_k w.r.t. E
        nodes, el = linear_to_quadratic(*cube_hexa(33, 33, 33))
        model = Solid(nodes.cuda(), el.cuda(), IsotropicElasticity3D(1000.0, 0.3))
        X = model.nodes
        model.constraints[X[:, 2] == 0.0, :] = True
        top = X[:, 2] == 1.0
        loads = []
        for k in range(16):
            trac = torch.randn(32 * 32, 3, generator=torch.Generator().manual_seed(k), dtype=dt)
            f = model.integrate_surface_load(top, trac.cuda())
            loads.append(f / f.sum(0).norm())
        Fb = torch.stack(loads)
        nrhs = 16
        desc = "Hexa2 32^3, 16 load cases, batched solve + adjoint of sum_k C_k w.r.t. E"

        def step(precision, stol):
            E = torch.full((model.n_elem,), 1000.0, dtype=dt, device="cuda", requires_grad=True)
            model.material = IsotropicElasticity3D(E, 0.3)
            u, _ = model.solve_load_cases(Fb, differentiable_parameters=E, precision=precision,
                                          stol=stol, preconditioner=args.preconditioner,
                                          inner_rtol=args.inner_rtol)
            (g,) = torch.autograd.grad((Fb * u).sum(), E)
            return u.detach(), g

    log(f"{desc}: {model.n_dofs} DOFs, {nrhs} load case(s)")
    ms_mixed, (u_mix, g_mix) = timed(lambda: step("mixed", args.stol), args.steps, args.warmup,
                                     "mixed")
    info = dict(model.last_solve_info)
    ms_fp64 = None
    info64 = None
    if args.fp64_steps > 0:
        ms_fp64, _ = timed(lambda: step("fp64", args.stol), args.fp64_steps, 1, "fp64")
        info64 = dict(model.last_solve_info)
    acc = {}
    if not args.no_reference:
        u_ref, g_ref = step("fp64", 1e-11)
        rel = lambda a, b: float((a - b).norm() / b.norm())  # noqa: E731
        acc = {"u_rel_err_vs_fp64": rel(u_mix, u_ref), "grad_rel_err_vs_fp64": rel(g_mix, g_ref)}
    fwd = info["forward"]["pcg"]
    return {
        "metric": f"solve+adjoint-grad time & DOFs/s, {desc}, mixed-prec vs FP64 @{args.stol:g} rel",
        "value": model.n_dofs * nrhs / ms_mixed,
        "unit": "DOFs/s (per load case, summed)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed * 1e3,
        "higher_is_better": True,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_dofs": model.n_dofs, "load_cases": nrhs,
                   "nnz": model._pattern.nnz, "n_adj": int(model._pattern.adj.numel()),
                   "adj16": model._pattern.adj16 is not None, "stol": args.stol,
                   "preconditioner": args.preconditioner, "inner_rtol": args.inner_rtol},
        "fp64_ms_per_step": ms_fp64 * 1e3 if ms_fp64 else None,
        "speedup_vs_fp64": ms_fp64 / ms_mixed if ms_fp64 else None,
        "pcg_iterations": {
            "mixed_fwd": [sum(p["iterations"][r] for p in fwd) for r in range(nrhs)],
            "mixed_adj": info["adjoint"]["iterations"][:nrhs],
            "fp64_fwd": ([sum(p["iterations"][r] for p in info64["forward"]["pcg"])
                          for r in range(nrhs)] if info64 else None),
            "fp64_adj": info64["adjoint"]["iterations"][:nrhs] if info64 else None},
        "accuracy": acc,
    }


def neo_hookean(F, params):
    """Compressible Neo-Hookean energy mu/2 (tr C - 3) - mu ln J + lbd/2 (ln J)^2."""
    import torch

    C = F.transpose(-1, -2) @ F
    ln_j = 0.5 * torch.logdet(C)
    return params[0] / 2 * (torch.trace(C) - 3.0) - params[0] * ln_j + params[1] / 2 * ln_j**2


def stretch_closed_form(mu: float, lbd: float, lam: float):
    """Uniaxial stretch lam of a Neo-Hookean block with free lateral faces:
    J from mu (J / lam - 1) + lbd ln J = 0, the nominal stress
    R = mu (lam - 1 / lam) + lbd ln J / lam and dR/d(mu, lbd)."""
    import math

    J = 1.0
    for _ in range(100):
        J -= (mu * (J / lam - 1.0) + lbd * math.log(J)) / (mu / lam + lbd / J)
    den = mu / lam + lbd / J
    dj_dmu, dj_dlbd = -(J / lam - 1.0) / den, -math.log(J) / den
    R = mu * (lam - 1.0 / lam) + lbd * math.log(J) / lam
    return R, (lam - 1.0 / lam + lbd / (lam * J) * dj_dmu,
               math.log(J) / lam + lbd / (lam * J) * dj_dlbd)


def hyper_workload(args) -> dict:
    """Finite-strain Newton path: a Neo-Hookean box of 4 x 104 x 104 cubic
    Hexa1 cells stretched to 10x its length along x in 10 geometric
    increments (nlgeom), then d(reaction)/d(mu, lbd) through every increment
    (the reference's hyperelasticity benchmark problem at its largest size,
    benchmarks/hyperelasticity.py). The homogeneous solution gives the exact
    reaction and gradient to check against."""
    import math

    import torch

    from torchfem import Solid
    from torchfem.materials import Hyperelastic3D
    from torchfem.mesh import cube_hexa

    N, stretch, n_inc = 105, 10.0, 11
    E, nu = 1000.0, 0.3
    lbd, mu = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))
    lx = 4.0 / (N - 1)
    nodes, elements = cube_hexa(5, N, N, lx, 1.0, 1.0)
    params0 = torch.tensor([mu, lbd], dtype=torch.float64)
    box = Solid(nodes.cuda(), elements.cuda(), Hyperelastic3D(neo_hookean, params0))
    X = box.nodes
    right = X[:, 0] == lx
    box.constraints[X[:, 0] == 0.0, 0] = True
    box.constraints[right, 0] = True
    box.constraints[X[:, 1] == 0.5, 1] = True
    box.constraints[X[:, 2] == 0.5, 2] = True
    box.displacements[right, 0] = (stretch - 1.0) * lx
    lam = torch.logspace(0, math.log10(stretch), n_inc, dtype=torch.float64)
    incs = (lam - 1.0) / (stretch - 1.0)
    R_exact, dR_exact = stretch_closed_form(mu, lbd, stretch)
    area = 1.0  # the reaction acts on the unit x-face
    log(f"hyperelastic stretch: {box.n_dofs} DOFs, {box.n_elem} Hexa1, {n_inc - 1} increments")

    def run(precision):
        p = params0.clone().requires_grad_(True)
        box.material = Hyperelastic3D(neo_hookean, p).vectorize(box.n_elem)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # the per-increment report goes to stderr with the other progress lines
        with contextlib.redirect_stdout(sys.stderr):
            _, f, _, _, _ = box.solve(increments=incs, nlgeom=True, differentiable_parameters=p,
                                      precision=precision, verbose=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        reaction = f[right, 0].sum()
        (g,) = torch.autograd.grad(reaction, p)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        info = box.last_solve_info["forward"]["increments"]
        solvers = [q.get("solver") for inc in info for q in inc["pcg"]]
        return {"fwd_s": t1 - t0, "bwd_s": t2 - t1,
                "newton_iterations": [inc["newton_iterations"] for inc in info],
                "linear_solves": {k: solvers.count(k) for k in set(solvers)},
                "reaction_rel_err_vs_closed_form": abs(float(reaction) - R_exact * area) / R_exact,
                "grad_rel_err_vs_closed_form": float(
                    (g - torch.tensor(dR_exact, dtype=torch.float64)).norm()
                    / torch.tensor(dR_exact).norm())}

    for i in range(args.warmup):
        run("mixed")
        log(f"hyper: warmup {i + 1}/{args.warmup} done")
    res = {}
    for prec in ("mixed", "fp64") if args.fp64_steps > 0 else ("mixed",):
        runs = [run(prec) for _ in range(max(1, args.steps if prec == "mixed" else args.fp64_steps))]
        log(f"hyper {prec}: {[round(r['fwd_s'] + r['bwd_s'], 3) for r in runs]} s")
        best = min(runs, key=lambda r: r["fwd_s"] + r["bwd_s"])
        res[prec] = dict(best, fwd_s_all=[r["fwd_s"] for r in runs],
                         bwd_s_all=[r["bwd_s"] for r in runs])
    m = res["mixed"]
    total = m["fwd_s"] + m["bwd_s"]
    return {
        "metric": "Neo-Hookean stretch (10x, 10 increments, nlgeom) solve + adjoint "
                  "d(reaction)/d(mu, lbd) time & DOFs/s",
        "value": box.n_dofs / total,
        "unit": "DOFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total * 1e3,
        "higher_is_better": True,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "hyper105", "N": N, "n_dofs": box.n_dofs, "n_elem": box.n_elem,
                   "increments": n_inc - 1, "stretch": stretch,
                   "newton": "rtol 1e-8, atol 1e-6, PCG stol 1e-10 (solve defaults)"},
        "results": res,
        "reference_published": {
            "source": "benchmarks/results/hyperelasticity_rtx5090_cuda.json (N=105)",
            "hardware": "RTX 5090, AmgX, FP64", "fwd_s": 39.6331, "bwd_s": 7.5877},
    }


def main() -> None:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding the path over "
              "several GPUs is not built", file=sys.stderr)
        sys.exit(2)
    import torch

    # FP64 default dtype, as the reference's benchmark harness sets it
    # (benchmarks/utils.py:59): material constants given as Python floats
    # become FP64 tensors
    torch.set_default_dtype(torch.float64)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "hyper105":
        print(json.dumps(hyper_workload(args)), flush=True)
        return
    if args.workload != "hexa1_64":
        print(json.dumps(other_workload(args)), flush=True)
        return
    log(f"building Hexa1 cube {args.n}^3")
    model, setup = build_model(args.n)
    pat = model._pattern
    log(f"setup {setup:.2f}s | {model.n_dofs} DOFs | nnz {pat.nnz}")

    step_mixed = make_step(model, "mixed", args.stol, args.preconditioner, args.inner_rtol)
    solver = None
    # warmup, then the timed steps with SpMV event timing on the solve stream
    for i in range(args.warmup):
        step_mixed()
        torch.cuda.synchronize()
        log(f"mixed: warmup {i + 1}/{args.warmup} done")
    solver = model._solver()
    solver.timing(args.timing_stride)
    ms_mixed, (u_mix, gE_mix, gF_mix, f_mix, s_mix) = timed(step_mixed, args.steps, 0, "mixed")
    spmv_ms, spmv_n = solver.timing(0)
    info_mixed = dict(model.last_solve_info)

    ms_fp64, info_fp64, u64, gE64 = None, None, None, None
    if args.fp64_steps > 0:
        step_fp64 = make_step(model, "fp64", args.stol, args.preconditioner, args.inner_rtol)
        ms_fp64, (u64, gE64, _, _, _) = timed(step_fp64, args.fp64_steps, 1, "fp64")
        info_fp64 = dict(model.last_solve_info)

    # the same step with the two-level preconditioner (coarse correction on
    # node aggregates), reported beside the block-Jacobi headline
    two, u_two, gE_two = None, None, None
    if args.two_level_steps > 0 and args.preconditioner != "two_level":
        step_two = make_step(model, "mixed", args.stol, "two_level", args.inner_rtol)
        ms_two, (u_two, gE_two, _, _, _) = timed(step_two, args.two_level_steps, 1, "two-level")
        info_two = dict(model.last_solve_info)
        two = {
            "preconditioner": "two_level (block-Jacobi + coarse correction on node aggregates)",
            "ms_per_step": ms_two * 1e3,
            "value": model.n_dofs / ms_two,
            "unit": "DOFs/s",
            "speedup_vs_block_jacobi": ms_mixed / ms_two,
            "pcg_iterations": {
                "mixed_fwd": sum(p["iterations"][0] for p in info_two["forward"]["pcg"]),
                "mixed_adj": info_two["adjoint"]["iterations"][0]},
        }
        log(f"two-level {two}")

    # FP64 reference of the same problem solved tightly, for the error report
    acc = {
        "relres_forward": info_mixed["forward"]["pcg"][-1]["relres"][0],
        "relres_adjoint": info_mixed["adjoint"]["relres"][0],
    }
    if not args.no_reference:
        step_ref = make_step(model, "fp64", 1e-11, args.preconditioner, args.inner_rtol)
        u_ref, gE_ref, gF_ref, f_ref, s_ref = step_ref()
        rel = lambda a, b: float((a - b).norm() / b.norm())  # noqa: E731
        acc.update({
            "u_rel_err_vs_fp64": rel(u_mix, u_ref),
            "f_rel_err_vs_fp64": rel(f_mix, f_ref),
            "sigma_rel_err_vs_fp64": rel(s_mix, s_ref),
            "gradE_rel_err_vs_fp64": rel(gE_mix, gE_ref),
            "gradF_rel_err_vs_fp64": rel(gF_mix, gF_ref),
            "reference": "all-FP64 GPU PCG to 1e-11, whose parity with the CPU oracle is tested",
        })
        if u64 is not None:  # the FP64 path at the same tolerance, for scale
            acc["fp64_same_stol_u_rel_err"] = rel(u64, u_ref)
            acc["fp64_same_stol_gradE_rel_err"] = rel(gE64, gE_ref)
        if two is not None:
            two["u_rel_err_vs_fp64"] = rel(u_two, u_ref)
            two["gradE_rel_err_vs_fp64"] = rel(gE_two, gE_ref)
    log(f"accuracy {acc}")

    # roofline of the dominant kernel: the FP32 node-block SpMV fused with the
    # direction update (reads K, the adjacency, u once, p and s; writes p and s)
    n_adj = int(pat.adj.numel())
    idx_bytes = 2 if pat.adj16 is not None else 4  # 16-bit neighbour offsets when they fit
    bytes_launch = pat.nnz * 4 + n_adj * idx_bytes + (pat.n_nod + 1) * (4 + 8) + model.n_dofs * 4 * 5
    avg_s = spmv_ms / max(spmv_n, 1) / 1e3
    achieved = bytes_launch / avg_s / 1e9 if spmv_n else None
    roof = {
        # the launcher runs the 6-wave variant of this kernel on gfx950 (the
        # 8-wave one would spill); the committed kernel trace names it
        "kernel": "k_pcg_spmv_dir_w6<float,3,16,1,false,true>",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved else None,
        "traffic": measured_traffic(),
        "avg_launch_us": avg_s * 1e6,
        "launches": spmv_n,
        "timing": f"HIP events on the solver stream around every {args.timing_stride}th "
                  "iteration's SpMV launch during the timed steps",
        "algorithmic_bytes": bytes_launch,
    }
    it_fwd = sum(p["iterations"][0] for p in info_mixed["forward"]["pcg"])
    it_adj = info_mixed["adjoint"]["iterations"][0]
    it64_fwd = sum(p["iterations"][0] for p in info_fp64["forward"]["pcg"]) if info_fp64 else None
    it64_adj = info_fp64["adjoint"]["iterations"][0] if info_fp64 else None
    result = {
        "metric": METRIC,
        "value": model.n_dofs / ms_mixed,
        "unit": "DOFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"Hexa1 cube {args.n}^3 (z=0 clamped, u_z=-0.01 on z=1), assembly + "
                        "mixed-precision PCG to 1e-8 + adjoint d(sum u)/d(E_e, F)",
            "n_dofs": model.n_dofs,
            "nnz": pat.nnz,
            "precision": "mixed (FP32 inner PCG, FP64 refinement)",
            "preconditioner": args.preconditioner,
            "stol": args.stol,
        },
        "setup_s": setup,
        "fp64_ms_per_step": ms_fp64 * 1e3 if ms_fp64 else None,
        "speedup_vs_fp64": ms_fp64 / ms_mixed if ms_fp64 else None,
        "pcg_iterations": {"mixed_fwd": it_fwd, "mixed_adj": it_adj,
                           "mixed_outer_fwd": info_mixed["forward"]["pcg"][-1]["outer"],
                           "mixed_outer_adj": info_mixed["adjoint"]["outer"],
                           "fp64_fwd": it64_fwd, "fp64_adj": it64_adj},
        "phase_s": {"mixed_fwd_pcg": info_mixed["forward"]["pcg"][-1]["time"],
                    "mixed_adj_pcg": info_mixed["adjoint"]["time"]},
        "accuracy": acc,
        "two_level": two,
        "roofline": roof,
        "cpu_baseline": None,
        "published_context": {
            "note": "no published number for this metric; the nearest reference rows are "
                    "FP64 solve()+backward() of its cube benchmark (load u_x=0.1 at x=1, "
                    "clamped x=0), DOFs/s derived as dofs/(fwd+bwd)",
            "rtx5090_amgx_cube_n60_dofs_per_s": 220957,
            "rtx5090_amgx_cube_n70_dofs_per_s": 237441,
            "threadripper_9975wx_cube_n60_dofs_per_s": 43461,
        },
    }
    if not args.no_cpu_baseline:
        # iteration counts of the oracle's solver (Jacobi-CG, FP64, same stol)
        # on the full cube, from the same algorithm on the GPU
        step_j = make_step(model, "fp64", args.stol, "jacobi", args.inner_rtol)
        step_j()
        info_j = model.last_solve_info
        iters_j = (sum(p["iterations"][0] for p in info_j["forward"]["pcg"]),
                   info_j["adjoint"]["iterations"][0])
        log(f"CPU baseline sample: {args.n}x{args.n}x8 slab with the oracle; Jacobi-CG "
            f"iterations on the full cube {iters_j}")
        result["cpu_baseline"] = cpu_baseline(args.n, iters_j)
        log(f"CPU baseline {result['cpu_baseline']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
