"""Solid models (src/torchfem/solid.py:15-271): Solid and SolidHeat."""

from __future__ import annotations

from functools import cached_property

import torch
from torch import Tensor

from .base import FEM, Heat, Mechanics
from .elements import Element, Hexa1, Hexa2, Tetra1


class SolidGeometry(FEM):
    """Three-dimensional continuum discretisation shared by Solid/SolidHeat."""

    def __repr__(self) -> str:
        return (f"<torch-fem solid ({self.n_nod} nodes, {self.n_elem} "
                f"{self.etype.__name__} elements)>")

    @property
    def etype(self) -> type[Element]:
        """Element type from the nodes per element (solid.py:33-44)."""
        n = self.elements.shape[1]
        if n == 4:
            return Tetra1
        if n == 8:
            return Hexa1
        if n == 20:
            return Hexa2
        raise ValueError("Element type not supported.")

    @cached_property
    def char_lengths(self) -> Tensor:
        return self.integrate_field() ** (1 / 3)

    def compute_k(self, detJ: Tensor, BCB: Tensor) -> Tensor:
        """Element stiffness contribution detJ * BCB (solid.py:52-54)."""
        return torch.einsum("j,jkl->jkl", detJ, BCB)

    def compute_f(self, detJ: Tensor, B: Tensor, S: Tensor) -> Tensor:
        """Element internal force (solid.py:56-58)."""
        return torch.einsum("..., ...iI,...Ai->...IA", detJ, B, S)


class Solid(SolidGeometry, Mechanics):
    """Solid mechanics in 3D (solid.py:240-254)."""

    @property
    def n_flux(self) -> list[int]:
        return [3, 3]


class SolidHeat(SolidGeometry, Heat):
    """Solid heat conduction (solid.py:257-271)."""
