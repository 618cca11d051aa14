"""Planar models (src/torchfem/planar.py:17-339): Planar and PlanarHeat."""

from __future__ import annotations

from functools import cached_property

import torch
from torch import Tensor

from .base import FEM, Heat, Mechanics
from .elements import Element, Quad1
from .materials import Material


class PlanarGeometry(FEM):
    """Surface in the z = 0 plane with per-element thickness (planar.py:17-96)."""

    def __init__(self, nodes: Tensor, elements: Tensor, material: Material,
                 thickness: Tensor | float = 1.0):
        super().__init__(nodes, elements, material)
        if isinstance(thickness, float):
            self.thickness = torch.full((self.n_elem,), thickness, dtype=torch.float64,
                                        device=self.device)
        else:
            self.thickness = torch.as_tensor(thickness)

    def __repr__(self) -> str:
        return (f"<torch-fem planar ({self.n_nod} nodes, {self.n_elem} "
                f"{self.etype.__name__} elements)>")

    @property
    def etype(self) -> type[Element]:
        """Element type from the nodes per element (planar.py:62-75); the
        Quad1 element is the one on the assemble-and-solve path."""
        if self.elements.shape[1] == 4:
            return Quad1
        raise ValueError("Element type not supported.")

    @cached_property
    def char_lengths(self) -> Tensor:
        return self.integrate_field() ** (1 / 2)

    @property
    def volume_scale(self) -> Tensor:
        return self.thickness

    def _scale_tensor(self) -> Tensor:
        t = torch.as_tensor(self.thickness)
        return torch.broadcast_to(t.to(device=self.device, dtype=torch.float64),
                                  (self.n_elem,)).contiguous()

    def compute_k(self, detJ: Tensor, BCB: Tensor):
        """thickness * detJ * BCB (planar.py:86-88)."""
        return torch.einsum("...,...,...kl->...kl", self.thickness, detJ, BCB)

    def compute_f(self, detJ: Tensor, B: Tensor, S: Tensor):
        """planar.py:90-92."""
        return torch.einsum("...,...,...iI,...Ai->...IA", self.thickness, detJ, B, S)


class Planar(PlanarGeometry, Mechanics):
    """Plane stress / plane strain mechanics (planar.py:306-320)."""

    @property
    def n_flux(self) -> list[int]:
        return [2, 2]


class PlanarHeat(PlanarGeometry, Heat):
    """Planar heat conduction (planar.py:323-339)."""
