"""Linear, stateless materials of the assemble-and-solve path.

Same classes, attributes and `step` contract as the reference
(src/torchfem/materials/base.py:29-91, elasticity.py:12-182,
conductivity.py:11-98). Each material also tells the HIP kernels how to read
its parameters: isotropic materials pass two (or one) numbers per element
(lam, G or kappa) instead of a materialised 81-entry tangent; assigning a
tangent (`material.C = ...` or `material.KAPPA = ...`, e.g. a SIMP-scaled C)
switches to the general layout, which reads the full tensor per element.
Autograd links from these parameters to user design variables are kept, so
the adjoint gradients flow back to them.
"""

from __future__ import annotations

import copy
import math
from abc import ABC, abstractmethod

import torch
from torch import Tensor

from .. import _lib as L

__all__ = [
    "Material",
    "IsotropicElasticity3D",
    "IsotropicElasticityPlaneStress",
    "IsotropicConductivity3D",
    "IsotropicConductivity2D",
    "Hyperelastic3D",
    "HyperelasticPlaneStrain",
    "HyperelasticPlaneStress",
    "IsotropicPlasticity3D",
]


def _t(x) -> Tensor:
    return torch.as_tensor(x, dtype=torch.get_default_dtype()) if not isinstance(x, Tensor) else x


class Material(ABC):
    """Base class (materials/base.py:10-103)."""

    n_state: int = 0
    is_vectorized: bool = False

    def __init__(self):
        self.n_state = 0
        self.is_vectorized = False
        self.rho = torch.tensor(1.0)

    def vectorize(self, n_elem: int) -> "Material":
        """Returns the material batched over `n_elem` elements (base.py:29-46)."""
        if self.is_vectorized:
            return self
        new = copy.copy(self)
        for key, value in list(vars(new).items()):
            if isinstance(value, Tensor):
                setattr(new, key, value.repeat(n_elem, *(value.dim() * [1])))
        new.is_vectorized = True
        return new

    @abstractmethod
    def step(self, H_inc, F, stress, state, de0, cl, iter):
        """Incremental update (base.py:48-91): (stress, state, tangent)."""

    def rotate(self, R: Tensor) -> "Material":
        return self

    def tensors(self) -> dict:
        """The tensors an incremental update reads (the inputs through which
        gradients reach the material)."""
        return {}

    def replace(self, **tensors) -> "Material":
        """A shallow copy with the named tensors replaced."""
        new = copy.copy(self)
        for key, value in tensors.items():
            setattr(new, key, value)
        return new

    # ------------------------------------------------------------ kernels
    @abstractmethod
    def kernel_params(self, n_elem: int, device) -> tuple[int, Tensor, Tensor | None]:
        """(layout, p0, p1) for libtfm_hip: FP64 per-element tensors on `device`."""


def _per_element(x: Tensor, n_elem: int, device) -> Tensor:
    x = torch.as_tensor(x).to(device=device, dtype=torch.float64)
    return torch.broadcast_to(x.reshape(-1) if x.dim() else x, (n_elem,)).contiguous()


class IsotropicElasticity3D(Material):
    """Isotropic linear elasticity in 3D (elasticity.py:12-127).

    lam = E nu / ((1 + nu)(1 - 2 nu)), G = E / (2 (1 + nu)),
    C_ijkl = lam d_ij d_kl + G (d_ik d_jl + d_il d_jk).
    """

    dim = 3

    def __init__(self, E: Tensor | float, nu: Tensor | float, rho: Tensor | float = 1.0):
        self.E = _t(E)
        self.nu = _t(nu)
        self.rho = _t(rho)
        self.n_state = 0
        self.is_vectorized = self.E.dim() > 0
        self.lbd = self.E * self.nu / ((1.0 + self.nu) * (1.0 - 2.0 * self.nu))
        self.G = self.E / (2.0 * (1.0 + self.nu))
        self._C: Tensor | None = None

    def _iso_lambda(self) -> Tensor:
        return self.lbd

    @property
    def C(self) -> Tensor:
        """Stiffness tensor [..., d, d, d, d]; assigned values override (lam, G)."""
        if self._C is not None:
            return self._C
        d = self.dim
        I2 = torch.eye(d, dtype=self.G.dtype, device=self.G.device)
        I4 = torch.einsum("ij,kl->ijkl", I2, I2)
        I4S = torch.einsum("ik,jl->ijkl", I2, I2) + torch.einsum("il,jk->ijkl", I2, I2)
        lam = self._iso_lambda()[..., None, None, None, None]
        return lam * I4 + self.G[..., None, None, None, None] * I4S

    @C.setter
    def C(self, value: Tensor) -> None:
        self._C = value

    def step(self, H_inc, F, stress, state, de0, cl, iter):
        """sigma_{n+1} = sigma_n + C : (sym(H_inc) - de0), tangent C (elasticity.py:119-127)."""
        de = 0.5 * (H_inc.transpose(-1, -2) + H_inc)
        C = self.C
        stress_new = stress + torch.einsum("...ijkl,...kl->...ij", C, de - de0)
        return stress_new, state, C

    def kernel_params(self, n_elem, device):
        if self._C is None:
            return (
                L.MAT_ISO,
                _per_element(self._iso_lambda(), n_elem, device),
                _per_element(self.G, n_elem, device),
            )
        d = self.dim
        C = torch.as_tensor(self._C).to(device=device, dtype=torch.float64)
        C = torch.broadcast_to(C, (n_elem, d, d, d, d)).reshape(n_elem, d**4).contiguous()
        return L.MAT_GENERAL, C, None


class IsotropicElasticityPlaneStress(IsotropicElasticity3D):
    """Plane stress (elasticity.py:130-182): C0000 = C1111 = E/(1-nu^2),
    C0011 = C1100 = E nu/(1-nu^2), C0101 = C0110 = C1001 = C1010 = E/(2(1+nu)),
    i.e. the isotropic form with lam' = E nu / (1 - nu^2)."""

    dim = 2

    def _iso_lambda(self) -> Tensor:
        return self.E * self.nu / (1.0 - self.nu**2)


class IsotropicConductivity3D(Material):
    """Isotropic heat conduction (conductivity.py:11-98): KAPPA = kappa I,
    flux q = q_n + KAPPA (grad T_inc - de0)."""

    dim = 3

    def __init__(self, kappa: Tensor | float, rho: Tensor | float = 1.0):
        self.kappa = _t(kappa)
        self.rho = _t(rho)
        self.n_state = 0
        self.is_vectorized = self.kappa.dim() > 0
        self._KAPPA: Tensor | None = None

    @property
    def KAPPA(self) -> Tensor:
        if self._KAPPA is not None:
            return self._KAPPA
        return self.kappa[..., None, None] * torch.eye(self.dim, dtype=self.kappa.dtype,
                                                       device=self.kappa.device)

    @KAPPA.setter
    def KAPPA(self, value: Tensor) -> None:
        self._KAPPA = value

    def step(self, H_inc, F, stress, state, de0, cl, iter):
        K = self.KAPPA
        q = stress + torch.einsum("...ij,...kj->...ki", K, H_inc - de0)
        return q, state, K

    def kernel_params(self, n_elem, device):
        if self._KAPPA is None:
            return L.MAT_HEAT_ISO, _per_element(self.kappa, n_elem, device), None
        d = self.dim
        K = torch.as_tensor(self._KAPPA).to(device=device, dtype=torch.float64)
        K = torch.broadcast_to(K, (n_elem, d, d)).reshape(n_elem, d * d).contiguous()
        return L.MAT_HEAT_GENERAL, K, None


class IsotropicConductivity2D(IsotropicConductivity3D):
    """Planar isotropic heat conduction (conductivity.py, 2D variant)."""

    dim = 2


class Hyperelastic3D(Material):
    """Hyperelastic material in 3D (hyperelasticity.py:12-127): a strain
    energy density psi(F, params) supplies the first Piola-Kirchhoff stress
    P = dpsi/dF and the tangent A = d2psi/dF2, both by torch.func automatic
    differentiation, per Gauss point. It has no constant tangent, so models
    with it are solved by the finite-strain Newton path of FEM.solve, whose
    displacement gradients, internal forces and element matrices are HIP
    kernels; psi itself runs as the user wrote it."""

    dim = 3
    finite_strain = True
    incremental = True

    def __init__(self, psi, params, rho: Tensor | float = 1.0):
        self.psi = psi
        self.params = torch.as_tensor(params)
        self.n_state = 0
        self.rho = _t(rho)
        self.is_vectorized = self.params.dim() > 1

    def vectorize(self, n_elem: int) -> "Hyperelastic3D":
        """Parameters repeated per element, [n_elem, p] (hyperelasticity.py:73-80
        with base.py:29-46); the autograd link to the given tensor is kept."""
        if self.is_vectorized:
            return self
        new = copy.copy(self)
        new.params = self.params.repeat(n_elem, 1)
        new.rho = self.rho.repeat(n_elem)
        new.is_vectorized = True
        return new

    def tensors(self) -> dict:
        return {"params": self.params}

    def _batched(self, F: Tensor, params: Tensor | None):
        p = self.params if params is None else params
        p = p.to(device=F.device, dtype=torch.float64)
        lead = F.shape[:-2]
        # (p,) or [n_elem, p] against F [..., n_elem, 3, 3]
        p = p.expand(*lead, p.shape[-1])
        return F.reshape(-1, 3, 3), p.reshape(-1, p.shape[-1]), lead

    def stress(self, F: Tensor, params: Tensor | None = None) -> Tensor:
        """P(F) for deformation gradients [..., 3, 3]."""
        from torch.func import jacrev, vmap

        Ff, pf, lead = self._batched(F, params)
        P = vmap(jacrev(self.psi))(Ff, pf)
        return P.reshape(*lead, 3, 3)

    def stress_tangent(self, F: Tensor, params: Tensor | None = None):
        """(P, A) with A[..., i, J, k, L] = d2psi / dF_iJ dF_kL."""
        from torch.func import jacrev, vmap

        Ff, pf, lead = self._batched(F, params)
        dpsi = jacrev(self.psi)

        def both(f, q):
            P = dpsi(f, q)
            return P, P

        A, P = vmap(jacrev(both, has_aux=True))(Ff, pf)
        return P.reshape(*lead, 3, 3), A.reshape(*lead, 3, 3, 3, 3)

    def step(self, H_inc, F, stress, state, de0, cl, iter):
        """(P, state, A) at F + H_inc (hyperelasticity.py:83-127)."""
        P, A = self.stress_tangent(F + H_inc)
        return P, state, A

    def kernel_params(self, n_elem, device):
        raise TypeError(
            "a hyperelastic material has no constant tangent: its models are solved "
            "by the finite-strain Newton path of FEM.solve"
        )


def _embed(F2: Tensor, f33) -> Tensor:
    """diag(F2, f33) as a 3-D deformation gradient [..., 3, 3]."""
    F3 = torch.zeros(*F2.shape[:-2], 3, 3, dtype=F2.dtype, device=F2.device)
    F3[..., :2, :2] = F2
    F3[..., 2, 2] = f33
    return F3


class HyperelasticPlaneStrain(Hyperelastic3D):
    """Plane-strain hyperelasticity (hyperelasticity.py:272-330): the 3-D
    energy at F = diag(F2, 1); P and the tangent are the in-plane blocks."""

    dim = 2

    def stress(self, F: Tensor, params: Tensor | None = None) -> Tensor:
        return super().stress(_embed(F, 1.0), params)[..., :2, :2]

    def stress_tangent(self, F: Tensor, params: Tensor | None = None):
        P, A = super().stress_tangent(_embed(F, 1.0), params)
        return P[..., :2, :2], A[..., :2, :2, :2, :2]


class HyperelasticPlaneStress(Hyperelastic3D):
    """Plane-stress hyperelasticity (hyperelasticity.py:130-269): the
    out-of-plane stretch lam3 (state: lam3 - 1) solves P33 = 0 by a local
    Newton iteration from the previous increment's value, and the in-plane
    tangent is condensed, A_ijkl - A_ij33 A_33kl / A_3333. It carries an
    internal state, so the incremental path runs its `step`."""

    dim = 2
    finite_strain = False

    def __init__(self, psi, params, rho: Tensor | float = 1.0, tolerance: float = 1e-5,
                 max_iter: int = 10):
        super().__init__(psi, params, rho)
        self.tolerance = tolerance
        self.max_iter = max_iter
        self.n_state = 1

    def step(self, H_inc, F, stress, state, de0, cl, iter):
        F2 = F + H_inc
        lam3 = 1.0 + state[..., 0]
        for _ in range(self.max_iter):
            P, A = Hyperelastic3D.stress_tangent(self, _embed(F2, lam3))
            res = P[..., 2, 2]
            lam3 = lam3 - res / A[..., 2, 2, 2, 2]
            if not bool((res.abs() >= self.tolerance).any()):
                break
        cond = torch.einsum("...ij,...kl->...ijkl", A[..., :2, :2, 2, 2], A[..., 2, 2, :2, :2])
        A2 = A[..., :2, :2, :2, :2] - cond / A[..., 2, 2, 2, 2][..., None, None, None, None]
        return P[..., :2, :2], (lam3 - 1.0)[..., None], A2


class IsotropicPlasticity3D(IsotropicElasticity3D):
    """Small-strain J2 (von Mises) plasticity with isotropic hardening
    (plasticity.py:24-215): yield stress sigma_f(q) of the equivalent plastic
    strain q (one state variable), radial return with a local Newton solve for
    the plastic multiplier, and the consistent tangent. Solved by the
    incremental Newton path of FEM.solve, which assembles the per-point
    tangents with the HIP kernels every iteration."""

    incremental = True

    def __init__(self, E, nu, sigma_f, sigma_f_prime, tolerance: float = 1e-5,
                 max_iter: int = 10, rho: Tensor | float = 1.0):
        super().__init__(E, nu, rho)
        self.sigma_f = sigma_f
        self.sigma_f_prime = sigma_f_prime
        self.tolerance = tolerance
        self.max_iter = max_iter
        self.n_state = 1

    def vectorize(self, n_elem: int) -> "IsotropicPlasticity3D":
        if self.is_vectorized:
            return self
        new = copy.copy(self)
        for key in ("E", "nu", "rho", "lbd", "G"):
            v = getattr(self, key)
            setattr(new, key, v.repeat(n_elem) if v.dim() == 0 else v)
        new.is_vectorized = True
        return new

    def tensors(self) -> dict:
        return {"lbd": self.lbd, "G": self.G}

    def step(self, H_inc, F, stress, state, de0, cl, iter):
        """One strain increment from (stress, state) for a batch [..., 3, 3]
        whose trailing batch dimension runs over the elements. Returns the
        stress, the state and the tangent [..., 3, 3, 3, 3]."""
        dev_t = H_inc.device
        lam = self.lbd.to(dev_t, torch.float64)
        G = self.G.to(dev_t, torch.float64)
        I = torch.eye(3, dtype=torch.float64, device=dev_t)
        eps = 0.5 * (H_inc + H_inc.transpose(-1, -2))
        if de0 is not None:
            eps = eps - de0
        tr = eps.diagonal(dim1=-2, dim2=-1).sum(-1)
        trial = stress + lam[..., None, None] * tr[..., None, None] * I + 2.0 * G[..., None, None] * eps
        s_dev = trial - trial.diagonal(dim1=-2, dim2=-1).sum(-1)[..., None, None] / 3.0 * I
        norm = torch.sqrt((s_dev * s_dev).sum((-1, -2)))
        q0 = state[..., 0]
        k = math.sqrt(2.0 / 3.0)
        yielding = norm - k * self.sigma_f(q0) > 0.0
        # local Newton on the consistency condition
        #   |s_dev| - 2 G dg - sqrt(2/3) sigma_f(q0 + sqrt(2/3) dg) = 0
        dg = torch.zeros_like(norm)
        q = q0
        for _ in range(self.max_iter):
            res = norm - 2.0 * G * dg - k * self.sigma_f(q)
            step = res / (2.0 * G + (2.0 / 3.0) * self.sigma_f_prime(q))
            dg = torch.where(yielding, dg + step, dg)
            q = q0 + k * dg
            if not bool((res.abs() * yielding).max() >= self.tolerance):
                break
        safe = torch.where(norm > 0.0, norm, torch.ones_like(norm))
        n = s_dev / safe[..., None, None]
        y = yielding[..., None, None]
        new_stress = torch.where(y, trial - (2.0 * G * dg)[..., None, None] * n, trial)
        new_state = torch.where(yielding[..., None], q[..., None], state)
        # consistent tangent: C - a n(x)n - b (I_dev - n(x)n)
        I4 = torch.einsum("ij,kl->ijkl", I, I)
        I4s = 0.5 * (torch.einsum("ik,jl->ijkl", I, I) + torch.einsum("il,jk->ijkl", I, I))
        C = (lam[..., None, None, None, None] * I4 + 2.0 * G[..., None, None, None, None] * I4s)
        C = C.expand(*new_stress.shape[:-2], 3, 3, 3, 3)
        a = 2.0 * G / (1.0 + self.sigma_f_prime(q) / (3.0 * G))
        b = 4.0 * G * G * dg / safe
        nn = torch.einsum("...ij,...kl->...ijkl", n, n)
        soft = a[..., None, None, None, None] * nn + b[..., None, None, None, None] * (
            I4s - I4 / 3.0 - nn)
        tangent = torch.where(yielding[..., None, None, None, None], C - soft, C)
        return new_stress, new_state, tangent

    def kernel_params(self, n_elem, device):
        raise TypeError(
            "a plastic material has no constant tangent: its models are solved by the "
            "incremental Newton path of FEM.solve"
        )
