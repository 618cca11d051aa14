"""torchfem for MI355X: the assemble-and-solve path of torch-fem on gfx950.

Drop-in names for the reference's models (Solid, SolidHeat, Planar,
PlanarHeat), materials, elements, meshes and sparse solves; the compute runs
in the hand-written HIP library libtfm_hip.so.
"""

from .planar import Planar, PlanarHeat  # noqa: F401
from .solid import Solid, SolidHeat  # noqa: F401

__all__ = ["Planar", "PlanarHeat", "Solid", "SolidHeat"]
