"""CPU FP64 restatement of torch-fem's assemble-and-solve path (test oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Every function cites the reference lines it restates (paths relative to the
torch-fem v0.9.0 source tree). The restatement keeps the reference's own
operations (einsum integration, index_add_ assembly, sorted-unique pattern,
SciPy solves, autograd through the residual for the adjoint) so that it can
stand in for the reference, which cannot be imported here (its package
imports pyvista and pyamg, both absent).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch
from torch import Tensor

DT = torch.float64
SQ3 = math.sqrt(3.0)

__all__ = [
    "Bar1", "Quad1", "Quad2", "Tria1", "Tetra1", "Hexa1", "Hexa2", "ELEMENTS",
    "cube_hexa", "rect_quad", "cube_tetra6", "linear_to_quadratic",
    "build_pattern", "iso3d_C", "plane_stress_C", "iso_KAPPA",
    "OracleModel", "simp_compliance_sensitivity", "hexa_nnz_closed_form",
    "neo_hookean_psi", "hyper_step", "oracle_integrate_hyper", "oracle_solve_hyper",
    "oracle_integrate_incremental", "oracle_solve_incremental", "hyper_material", "j2_step",
    "oracle_mass", "oracle_time_integration", "hyper_plane_strain", "hyper_plane_stress",
]


# --------------------------------------------------------------------------
# Elements (src/torchfem/elements.py)
# --------------------------------------------------------------------------
class Bar1:
    """elements.py:172-213 Bar1: 2-node line, 1-point rule (facet of Quad1)."""

    nodes, iso_dim = 2, 1

    @staticmethod
    def N(xi):
        return torch.stack([0.5 * (1 - xi[..., 0]), 0.5 * (1 + xi[..., 0])], dim=-1)

    @staticmethod
    def B(xi):
        one = torch.ones_like(xi[..., 0])
        return torch.stack([-0.5 * one, 0.5 * one], dim=-1).unsqueeze(-2)

    ipoints = torch.tensor([[0.0]], dtype=DT)
    iweights = torch.tensor([2.0], dtype=DT)


class Tria1:
    """elements.py:310-360 Tria1 (facet of Tetra1): 3-node triangle, 1-point rule."""

    nodes, iso_dim = 3, 2

    @staticmethod
    def N(xi):
        return torch.stack([1 - xi[..., 0] - xi[..., 1], xi[..., 0], xi[..., 1]], dim=-1)

    @staticmethod
    def B(xi):
        b = torch.tensor([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]], dtype=DT)
        return b.expand(*xi.shape[:-1], 2, 3)

    ipoints = torch.tensor([[1 / 3, 1 / 3]], dtype=DT)
    iweights = torch.tensor([0.5], dtype=DT)


class Quad1:
    """elements.py:491-567."""

    nodes, iso_dim, iso_volume = 4, 2, 4.0
    iso_coords = torch.tensor([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]], dtype=DT)
    facets = torch.tensor([[0, 1], [1, 2], [2, 3], [3, 0]])
    facet_type = Bar1

    @staticmethod
    def N(xi):  # elements.py:525-530
        x, y = xi[..., 0], xi[..., 1]
        return 0.25 * torch.stack(
            [(1 - x) * (1 - y), (1 + x) * (1 - y), (1 + x) * (1 + y), (1 - x) * (1 + y)], dim=-1
        )

    @staticmethod
    def B(xi):  # elements.py:532-553
        x, y = xi[..., 0], xi[..., 1]
        return 0.25 * torch.stack(
            [
                torch.stack([-(1 - y), (1 - y), (1 + y), -(1 + y)], dim=-1),
                torch.stack([-(1 - x), -(1 + x), (1 + x), (1 - x)], dim=-1),
            ],
            dim=xi.dim() - 1,
        )

    # elements.py:555-566: integer weights, points ordered xi_1 fastest
    iweights = torch.tensor([1, 1, 1, 1])
    ipoints = torch.tensor(
        [[a / SQ3, b / SQ3] for b in [-1, 1] for a in [-1, 1]], dtype=DT
    )


class Quad2:
    """elements.py:599-695 (facet of Hexa2): 8-node serendipity quad."""

    nodes, iso_dim = 8, 2

    @staticmethod
    def N(xi):  # elements.py:637-646
        x, y = xi[..., 0], xi[..., 1]
        return 0.25 * torch.stack(
            [
                -(1 - x) * (1 - y) * (1 + x + y),
                -(1 + x) * (1 - y) * (1 - x + y),
                -(1 + x) * (1 + y) * (1 - x - y),
                -(1 - x) * (1 + y) * (1 + x - y),
                2 * (1 - x) * (1 - y) * (1 + x),
                2 * (1 + x) * (1 - y) * (1 + y),
                2 * (1 - x) * (1 + y) * (1 + x),
                2 * (1 - x) * (1 - y) * (1 + y),
            ],
            dim=-1,
        )

    @staticmethod
    def B(xi):  # elements.py:648-680
        x, y = xi[..., 0], xi[..., 1]
        return 0.25 * torch.stack(
            [
                torch.stack(
                    [
                        -(y - 1) * (2 * x + y), -(y - 1) * (2 * x - y),
                        (y + 1) * (2 * x + y), (y + 1) * (2 * x - y),
                        4 * x * (y - 1), 2 - 2 * y**2, -4 * x * (y + 1), -2 + 2 * y**2,
                    ],
                    dim=-1,
                ),
                torch.stack(
                    [
                        -(x - 1) * (x + 2 * y), -(x + 1) * (x - 2 * y),
                        (x + 1) * (x + 2 * y), (x - 1) * (x - 2 * y),
                        -2 + 2 * x**2, -4 * y * (x + 1), 2 - 2 * x**2, 4 * (x - 1) * y,
                    ],
                    dim=-1,
                ),
            ],
            dim=xi.dim() - 1,
        )

    iweights = torch.tensor([1, 1, 1, 1])
    ipoints = torch.tensor([[a / SQ3, b / SQ3] for b in [-1, 1] for a in [-1, 1]], dtype=DT)


class Tetra1:
    """elements.py:727-790."""

    nodes, iso_dim, iso_volume = 4, 3, 1.0 / 6.0
    iso_coords = torch.tensor(
        [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]], dtype=DT
    )
    facets = torch.tensor([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])
    facet_type = Tria1

    @staticmethod
    def N(xi):  # elements.py:763-768
        return torch.stack(
            [1.0 - xi[..., 0] - xi[..., 1] - xi[..., 2], xi[..., 0], xi[..., 1], xi[..., 2]],
            dim=-1,
        )

    @staticmethod
    def B(xi):  # elements.py:770-780
        b = torch.tensor(
            [[-1.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 1.0, 0.0], [-1.0, 0.0, 0.0, 1.0]], dtype=DT
        )
        return b if xi.dim() == 1 else b.repeat(xi.shape[0], 1, 1)

    iweights = torch.tensor([1.0 / 6.0], dtype=DT)
    ipoints = torch.tensor([[0.25, 0.25, 0.25]], dtype=DT)


class Hexa1:
    """elements.py:936-1075."""

    nodes, iso_dim, iso_volume = 8, 3, 8.0
    iso_coords = torch.tensor(
        [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
         [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=DT
    )
    facets = torch.tensor(
        [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]
    )
    edges = torch.tensor(
        [[0, 1], [1, 2], [2, 3], [3, 0], [4, 5], [5, 6], [6, 7], [7, 4],
         [0, 4], [1, 5], [2, 6], [3, 7]]
    )
    facet_type = Quad1
    iweights = torch.ones(8, dtype=DT)
    # elements.py:1000-1009: xi_1 fastest
    ipoints = torch.tensor(
        [[a / SQ3, b / SQ3, c / SQ3] for c in [-1, 1] for b in [-1, 1] for a in [-1, 1]],
        dtype=DT,
    )

    @staticmethod
    def N(xi):  # elements.py:1011-1021
        x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]
        return 0.125 * torch.stack(
            [
                (1 - x) * (1 - y) * (1 - z), (1 + x) * (1 - y) * (1 - z),
                (1 + x) * (1 + y) * (1 - z), (1 - x) * (1 + y) * (1 - z),
                (1 - x) * (1 - y) * (1 + z), (1 + x) * (1 - y) * (1 + z),
                (1 + x) * (1 + y) * (1 + z), (1 - x) * (1 + y) * (1 + z),
            ],
            dim=-1,
        )

    @staticmethod
    def B(xi):  # elements.py:1023-1075
        x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]
        return 0.125 * torch.stack(
            [
                torch.stack(
                    [-(1 - y) * (1 - z), (1 - y) * (1 - z), (1 + y) * (1 - z), -(1 + y) * (1 - z),
                     -(1 - y) * (1 + z), (1 - y) * (1 + z), (1 + y) * (1 + z), -(1 + y) * (1 + z)],
                    dim=-1,
                ),
                torch.stack(
                    [-(1 - x) * (1 - z), -(1 + x) * (1 - z), (1 + x) * (1 - z), (1 - x) * (1 - z),
                     -(1 - x) * (1 + z), -(1 + x) * (1 + z), (1 + x) * (1 + z), (1 - x) * (1 + z)],
                    dim=-1,
                ),
                torch.stack(
                    [-(1 - x) * (1 - y), -(1 + x) * (1 - y), -(1 + x) * (1 + y), -(1 - x) * (1 + y),
                     (1 - x) * (1 - y), (1 + x) * (1 - y), (1 + x) * (1 + y), (1 - x) * (1 + y)],
                    dim=-1,
                ),
            ],
            dim=xi.dim() - 1,
        )


class Hexa2:
    """elements.py:1078-1406 (20-node serendipity, reduced 2x2x2 rule)."""

    nodes, iso_dim, iso_volume = 20, 3, 8.0
    iso_coords = torch.tensor(
        [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1],
         [1, 1, 1], [-1, 1, 1], [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1],
         [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1], [-1, -1, 0], [1, -1, 0],
         [1, 1, 0], [-1, 1, 0]], dtype=DT
    )
    facets = torch.tensor(
        [[0, 3, 2, 1, 11, 10, 9, 8], [4, 5, 6, 7, 12, 13, 14, 15],
         [0, 1, 5, 4, 8, 17, 12, 16], [1, 2, 6, 5, 9, 18, 13, 17],
         [2, 3, 7, 6, 10, 19, 14, 18], [3, 0, 4, 7, 11, 16, 15, 19]]
    )
    facet_type = Quad2
    iweights = Hexa1.iweights
    ipoints = Hexa1.ipoints

    @staticmethod
    def N(xi):  # elements.py:1162-1253
        x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]
        return 0.125 * torch.stack(
            [
                (1 - x) * (1 - y) * (1 - z) * (-x - y - z - 2),
                (1 + x) * (1 - y) * (1 - z) * (x - y - z - 2),
                (1 + x) * (1 + y) * (1 - z) * (x + y - z - 2),
                (1 - x) * (1 + y) * (1 - z) * (-x + y - z - 2),
                (1 - x) * (1 - y) * (1 + z) * (-x - y + z - 2),
                (1 + x) * (1 - y) * (1 + z) * (x - y + z - 2),
                (1 + x) * (1 + y) * (1 + z) * (x + y + z - 2),
                (1 - x) * (1 + y) * (1 + z) * (-x + y + z - 2),
                2 * (1 - x**2) * (1 - y) * (1 - z),
                2 * (1 + x) * (1 - y**2) * (1 - z),
                2 * (1 - x**2) * (1 + y) * (1 - z),
                2 * (1 - x) * (1 - y**2) * (1 - z),
                2 * (1 - x**2) * (1 - y) * (1 + z),
                2 * (1 + x) * (1 - y**2) * (1 + z),
                2 * (1 - x**2) * (1 + y) * (1 + z),
                2 * (1 - x) * (1 - y**2) * (1 + z),
                2 * (1 - x) * (1 - y) * (1 - z**2),
                2 * (1 + x) * (1 - y) * (1 - z**2),
                2 * (1 + x) * (1 + y) * (1 - z**2),
                2 * (1 - x) * (1 + y) * (1 - z**2),
            ],
            dim=-1,
        )

    @staticmethod
    def B(xi):  # elements.py:1255-1380
        x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]
        dx = [
            +(y - 1) * (z - 1) * (+2 * x + y + z + 1),
            -(y - 1) * (z - 1) * (-2 * x + y + z + 1),
            -(y + 1) * (z - 1) * (+2 * x + y - z - 1),
            +(y + 1) * (z - 1) * (-2 * x + y - z - 1),
            -(y - 1) * (z + 1) * (+2 * x + y - z + 1),
            +(y - 1) * (z + 1) * (-2 * x + y - z + 1),
            +(y + 1) * (z + 1) * (+2 * x + y + z - 1),
            -(y + 1) * (z + 1) * (-2 * x + y + z - 1),
            -4 * x * (y - 1) * (z - 1),
            +2 * (y**2 - 1) * (z - 1),
            +4 * x * (y + 1) * (z - 1),
            -2 * (y**2 - 1) * (z - 1),
            +4 * x * (y - 1) * (z + 1),
            -2 * (y**2 - 1) * (z + 1),
            -4 * x * (y + 1) * (z + 1),
            +2 * (y**2 - 1) * (z + 1),
            -2 * (y - 1) * (z**2 - 1),
            +2 * (y - 1) * (z**2 - 1),
            -2 * (y + 1) * (z**2 - 1),
            +2 * (y + 1) * (z**2 - 1),
        ]
        dy = [
            +(x - 1) * (z - 1) * (x + 2 * y + z + 1),
            +(x + 1) * (z - 1) * (x - 2 * y - z - 1),
            -(x + 1) * (z - 1) * (x + 2 * y - z - 1),
            -(x - 1) * (z - 1) * (x - 2 * y + z + 1),
            -(x - 1) * (z + 1) * (x + 2 * y - z + 1),
            -(x + 1) * (z + 1) * (x - 2 * y + z - 1),
            +(x + 1) * (z + 1) * (x + 2 * y + z - 1),
            +(x - 1) * (z + 1) * (x - 2 * y - z + 1),
            -2 * (x**2 - 1) * (z - 1),
            +4 * y * (x + 1) * (z - 1),
            +2 * (x**2 - 1) * (z - 1),
            -4 * y * (x - 1) * (z - 1),
            +2 * (x**2 - 1) * (z + 1),
            -4 * y * (x + 1) * (z + 1),
            -2 * (x**2 - 1) * (z + 1),
            +4 * y * (x - 1) * (z + 1),
            -2 * (x - 1) * (z**2 - 1),
            +2 * (x + 1) * (z**2 - 1),
            -2 * (x + 1) * (z**2 - 1),
            +2 * (x - 1) * (z**2 - 1),
        ]
        dz = [
            +(x - 1) * (y - 1) * (x + y + 2 * z + 1),
            +(x + 1) * (y - 1) * (x - y - 2 * z - 1),
            -(x + 1) * (y + 1) * (x + y - 2 * z - 1),
            -(x - 1) * (y + 1) * (x - y + 2 * z + 1),
            -(x - 1) * (y - 1) * (x + y - 2 * z + 1),
            -(x + 1) * (y - 1) * (x - y + 2 * z - 1),
            +(x + 1) * (y + 1) * (x + y + 2 * z - 1),
            +(x - 1) * (y + 1) * (x - y - 2 * z + 1),
            -2 * (x**2 - 1) * (y - 1),
            +2 * (x + 1) * (y**2 - 1),
            +2 * (x**2 - 1) * (y + 1),
            -2 * (x - 1) * (y**2 - 1),
            +2 * (x**2 - 1) * (y - 1),
            -2 * (x + 1) * (y**2 - 1),
            -2 * (x**2 - 1) * (y + 1),
            +2 * (x - 1) * (y**2 - 1),
            -4 * (x - 1) * (y - 1) * z,
            +4 * (x + 1) * (y - 1) * z,
            -4 * (x + 1) * (y + 1) * z,
            +4 * (x - 1) * (y + 1) * z,
        ]
        return 0.125 * torch.stack(
            [torch.stack(dx, dim=-1), torch.stack(dy, dim=-1), torch.stack(dz, dim=-1)],
            dim=xi.dim() - 1,
        )


ELEMENTS = {"Quad1": Quad1, "Tetra1": Tetra1, "Hexa1": Hexa1, "Hexa2": Hexa2}


# --------------------------------------------------------------------------
# Meshes (src/torchfem/mesh.py, elements.py:1438-1494)
# --------------------------------------------------------------------------
def cube_hexa(Nx, Ny, Nz, Lx=1.0, Ly=1.0, Lz=1.0):
    """mesh.py:8-46: node id i*Ny*Nz + j*Nz + k, element [n0,n1,n3,n2,n4,n5,n7,n6]."""
    X = torch.linspace(0, Lx, Nx, dtype=DT)
    Y = torch.linspace(0, Ly, Ny, dtype=DT)
    Z = torch.linspace(0, Lz, Nz, dtype=DT)
    x, y, z = torch.meshgrid(X, Y, Z, indexing="ij")
    nodes = torch.vstack([x.ravel(), y.ravel(), z.ravel()]).T
    ind = torch.arange(Nx * Ny * Nz).reshape((Nx, Ny, Nz))
    n0 = ind[:-1, :-1, :-1].ravel()
    n1 = ind[1:, :-1, :-1].ravel()
    n2 = ind[:-1, 1:, :-1].ravel()
    n3 = ind[1:, 1:, :-1].ravel()
    n4 = ind[:-1, :-1, 1:].ravel()
    n5 = ind[1:, :-1, 1:].ravel()
    n6 = ind[:-1, 1:, 1:].ravel()
    n7 = ind[1:, 1:, 1:].ravel()
    return nodes, torch.vstack([n0, n1, n3, n2, n4, n5, n7, n6]).T


def rect_quad(Nx, Ny, Lx=1.0, Ly=1.0):
    """mesh.py:103-135."""
    X = torch.linspace(0, Lx, Nx, dtype=DT)
    Y = torch.linspace(0, Ly, Ny, dtype=DT)
    x, y = torch.meshgrid(X, Y, indexing="ij")
    nodes = torch.vstack([x.ravel(), y.ravel()]).T
    ind = torch.arange(Nx * Ny).reshape((Nx, Ny))
    n0 = ind[:-1, :-1].ravel()
    n1 = ind[1:, :-1].ravel()
    n2 = ind[:-1, 1:].ravel()
    n3 = ind[1:, 1:].ravel()
    return nodes, torch.vstack([n0, n1, n3, n2]).T


def cube_tetra6(N, jitter=0.0, seed=0):
    """Six-tetrahedron (Kuhn) split of an N^3-cell unit cube (no reference
    generator exists; the 5-tet split of mesh.py:49-100 is a different mesh).

    Nodes follow cube_hexa's numbering; each cell is cut along its 0-6 main
    diagonal into the 6 tets [v0, v0+e_a, v0+e_a+e_b, v6] over the axis
    permutations (a, b, c), with odd permutations reordered so detJ > 0.
    Interior nodes move by U(-jitter*h, jitter*h) per coordinate with a torch
    generator seeded with `seed` (h = 1/N).
    """
    nodes, _ = cube_hexa(N + 1, N + 1, N + 1)
    n1 = N + 1
    ii, jj, kk = torch.meshgrid(torch.arange(N), torch.arange(N), torch.arange(N), indexing="ij")
    base = (ii * n1 * n1 + jj * n1 + kk).ravel()
    step = [n1 * n1, n1, 1]
    perms = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)]
    tets = []
    for p_idx, (a, b, c) in enumerate(perms):
        v1 = base + step[a]
        v2 = v1 + step[b]
        v3 = v2 + step[c]
        t = [base, v1, v2, v3] if p_idx < 3 else [base, v2, v1, v3]
        tets.append(torch.stack(t, dim=1))
    elements = torch.stack(tets, dim=1).reshape(-1, 4)
    if jitter > 0.0:
        h = 1.0 / N
        g = torch.Generator().manual_seed(seed)
        d = (torch.rand(nodes.shape, generator=g, dtype=DT) * 2.0 - 1.0) * jitter * h
        interior = ((nodes > 0.0) & (nodes < 1.0)).all(dim=1)
        nodes = nodes + d * interior[:, None]
    return nodes, elements


def linear_to_quadratic(nodes, elements):
    """elements.py:1438-1494 (Hexa1 -> Hexa2): unique sorted edges via numpy."""
    edges = Hexa1.edges
    ev = elements[:, edges].reshape(-1, 2)
    ev, _ = torch.sort(ev, dim=1)
    evn = ev.numpy()
    dt = np.dtype((np.void, evn.dtype.itemsize * ev.shape[1]))
    b = np.ascontiguousarray(evn).view(dt)
    _, uidx, inv = np.unique(b, return_index=True, return_inverse=True)
    uedges = ev[uidx]
    mid = (nodes[uedges[:, 0]] + nodes[uedges[:, 1]]) / 2.0
    new_nodes = torch.cat([nodes, mid], dim=0)
    mid_map = torch.as_tensor(inv).view(elements.shape[0], -1) + nodes.shape[0]
    return new_nodes, torch.cat([elements, mid_map], dim=1)


def hexa_nnz_closed_form(n):
    """nnz of the 3-DOF Hexa1 pattern on an n x n x n node grid: 9 (3n-2)^3."""
    return 9 * (3 * n - 2) ** 3


# --------------------------------------------------------------------------
# Sparsity pattern (base.py:72-122)
# --------------------------------------------------------------------------
def build_pattern(elements: Tensor, n_nod: int, ndof: int):
    """base.py:72-122: idx, glob_idx [2, nnz] int64, k_map int32, diag_map int32."""
    n_elem = len(elements)
    idx = (ndof * elements).unsqueeze(-1) + torch.arange(ndof)
    idx = idx.reshape(n_elem, -1)
    n = idx.shape[1]
    chunk = max(1, min(n_elem, (16 * 1024 * 1024) // (n * n)))
    parts = []
    for s in range(0, n_elem, chunk):
        ic = idx[s:s + chunk]
        parts.append(torch.unique(((ic.unsqueeze(-1) << 32) | ic.unsqueeze(1)).reshape(-1)))
    diag = torch.arange(ndof * n_nod, dtype=torch.int64)
    parts.append((diag << 32) | diag)
    packed = torch.unique(torch.cat(parts))
    k_parts = []
    for s in range(0, n_elem, chunk):
        ic = idx[s:s + chunk]
        k_parts.append(
            torch.searchsorted(packed, ((ic.unsqueeze(-1) << 32) | ic.unsqueeze(1)).reshape(-1)).to(
                torch.int32
            )
        )
    k_map = torch.cat(k_parts) if k_parts else torch.zeros(0, dtype=torch.int32)
    diag_map = torch.searchsorted(packed, (diag << 32) | diag).to(torch.int32)
    glob_idx = torch.stack([torch.div(packed, 2**32, rounding_mode="floor"), packed % 2**32])
    return idx, glob_idx, k_map, diag_map


# --------------------------------------------------------------------------
# Materials (materials/elasticity.py:44-70, 164-182; conductivity.py:30-45)
# --------------------------------------------------------------------------
def iso3d_C(E, nu):
    """elasticity.py:59-70: C = lam I(x)I + G (I4S), batched over E."""
    E = torch.as_tensor(E, dtype=DT)
    nu = torch.as_tensor(nu, dtype=DT)
    lbd = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    I2 = torch.eye(3, dtype=DT)
    I4 = torch.einsum("ij,kl->ijkl", I2, I2)
    I4S = torch.einsum("ik,jl->ijkl", I2, I2) + torch.einsum("il,jk->ijkl", I2, I2)
    return lbd[..., None, None, None, None] * I4 + G[..., None, None, None, None] * I4S


def plane_stress_C(E, nu):
    """elasticity.py:164-182."""
    E = torch.as_tensor(E, dtype=DT)
    nu = torch.as_tensor(nu, dtype=DT)
    fac = E / (1.0 - nu**2)
    C = torch.zeros(*E.shape, 2, 2, 2, 2, dtype=DT)
    C = C + 0.0 * fac[..., None, None, None, None]  # keep the autograd link to E
    C[..., 0, 0, 0, 0] = fac
    C[..., 0, 0, 1, 1] = fac * nu
    C[..., 1, 1, 0, 0] = fac * nu
    C[..., 1, 1, 1, 1] = fac
    C[..., 0, 1, 0, 1] = fac * 0.5 * (1.0 - nu)
    C[..., 0, 1, 1, 0] = fac * 0.5 * (1.0 - nu)
    C[..., 1, 0, 0, 1] = fac * 0.5 * (1.0 - nu)
    C[..., 1, 0, 1, 0] = fac * 0.5 * (1.0 - nu)
    return C


def iso_KAPPA(kappa, dim=3):
    """conductivity.py:41-45: KAPPA = kappa I."""
    kappa = torch.as_tensor(kappa, dtype=DT)
    return kappa[..., None, None] * torch.eye(dim, dtype=DT)


# --------------------------------------------------------------------------
# Model: integration, assembly, linear solve and the implicit adjoint
# --------------------------------------------------------------------------
@dataclass
class OracleModel:
    """A linear, stateless torch-fem model restated on the CPU.

    kind: "solid" (Solid, ndof = 3), "planar" (Planar, ndof = 2, thickness) or
    "heat" (SolidHeat/PlanarHeat, ndof = 1). `mat` is the per-element tangent:
    C [E, d, d, d, d] for mechanics, KAPPA [E, d, d] for heat (may carry autograd
    links to design parameters).
    """

    nodes: Tensor
    elements: Tensor
    etype: type
    kind: str
    mat: Tensor
    thickness: Tensor | None = None
    pattern: tuple | None = field(default=None, repr=False)

    def __post_init__(self):
        self.n_nod, self.n_dim = self.nodes.shape
        self.n_elem = len(self.elements)
        self.ndof = 1 if self.kind == "heat" else self.n_dim
        self.n_dofs = self.n_nod * self.ndof
        if self.pattern is None:
            self.pattern = build_pattern(self.elements, self.n_nod, self.ndof)
        self.idx, self.glob_idx, self.k_map, self.diag_map = self.pattern

    # base.py:293-314
    def eval_shape_functions(self, xi):
        nodes = self.nodes[self.elements, :]
        b = self.etype.B(xi)
        J = torch.einsum("...iN, ANj -> ...Aij", b, nodes)
        detJ = torch.linalg.det(J)
        if torch.any(detJ <= 0.0):
            raise ValueError("Negative Jacobian. Check element numbering.")
        B = torch.einsum("...Eij,...jN->...EiN", torch.linalg.inv(J), b)
        return self.etype.N(xi), B, detJ

    def _scale(self, detJ):
        return detJ * self.thickness if self.kind == "planar" or (
            self.kind == "heat" and self.thickness is not None
        ) else detJ

    def integrate(self, u: Tensor, need_k: bool = True, mat: Tensor | None = None):
        """Mechanics/Heat.integrate_material for a total field u from a zero state
        (base.py:982-1092, 1177-1286; compute_k/compute_f of solid.py:52-58 and
        planar.py:86-92). Returns k [E,W,W] (or None), f [E,W], grad, flux
        (both per Gauss point)."""
        mat = self.mat if mat is None else mat
        _, B, detJ = self.eval_shape_functions(self.etype.ipoints)
        npe, nd, E = self.etype.nodes, self.ndof, self.n_elem
        f = torch.zeros(E, nd * npe, dtype=DT)
        k = torch.zeros(E, nd * npe, nd * npe, dtype=DT) if need_k else None
        grads, fluxes = [], []
        if self.kind != "heat":
            due = u.view(-1, nd)[self.elements].reshape(E, -1, nd).transpose(-1, -2)
        else:
            due = u.view(-1, 1)[self.elements].reshape(E, -1, 1)
        for i, w in enumerate(self.etype.iweights):
            sdet = self._scale(detJ[i])
            if self.kind != "heat":
                H = due @ B[i].transpose(-1, -2)  # base.py:1052
                eps = 0.5 * (H.transpose(-1, -2) + H)  # elasticity.py:121
                P = torch.einsum("...ijkl,...kl->...ij", mat, eps)
                grads.append(torch.eye(self.n_dim, dtype=DT) + H)
                fluxes.append(P)
                fc = torch.einsum("..., ...iI,...Ai->...IA", sdet, B[i], P)  # solid.py:56-58
                f = f + w * fc.reshape(-1, nd * npe)
                if need_k:
                    BCB = torch.einsum("...Jp,...iJkL,...Lq->...piqk", B[i], mat, B[i])
                    BCB = BCB.reshape(-1, nd * npe, nd * npe)
                    k = k + torch.einsum("j,jkl->jkl", sdet, BCB) * w
            else:
                tg = torch.einsum("...ij,...jk->...ki", B[i], due)  # base.py:1250
                q = torch.einsum("...ij,...kj->...ki", mat, tg)  # conductivity.py:91-93
                grads.append(tg)
                fluxes.append(q)
                fc = torch.einsum("..., ...iI,...Ai->...IA", sdet, B[i], q)
                f = f + w * fc.reshape(-1, npe)
                if need_k:
                    BCB = torch.einsum("...ij,...iN,...jM->...NM", mat, B[i], B[i])
                    k = k + torch.einsum("j,jkl->jkl", sdet, BCB) * w
        return k, f, torch.stack(grads), torch.stack(fluxes)

    # base.py:398-426
    def assemble_matrix(self, k: Tensor, con: Tensor):
        val = torch.zeros(self.glob_idx.shape[1], dtype=DT)
        val.index_add_(0, self.k_map.long(), k.ravel())
        is_con = torch.zeros(self.n_dofs, dtype=torch.bool)
        is_con[con] = True
        mask = is_con[self.glob_idx[0]] | is_con[self.glob_idx[1]]
        val[mask] = 0.0
        val[self.diag_map[con].long()] = 1.0
        return val

    def csr(self, val: Tensor):
        r, c = self.glob_idx.numpy()
        return sp.csr_matrix((val.detach().numpy(), (r, c)), shape=(self.n_dofs, self.n_dofs))

    # base.py:428-445
    def assemble_rhs(self, f: Tensor):
        F = torch.zeros(self.n_dofs, dtype=f.dtype)
        return F.index_add(0, self.idx.ravel(), f.ravel())

    def solve_linear(self, K, rhs, method="auto", rtol=1e-13):
        """sparse.py:447-514 restated: spsolve below 10k DOFs (sparse.py:78-81),
        otherwise a tightly converged Jacobi-preconditioned CG (pyamg is absent)."""
        rhs = np.asarray(rhs, dtype=np.float64)
        if method == "spsolve" or (method == "auto" and K.shape[0] < 10000):
            return spla.spsolve(K.tocsc(), rhs)
        d = K.diagonal()
        M = sp.diags(1.0 / d)
        x, info = spla.cg(K, rhs, rtol=rtol, atol=0.0, M=M, maxiter=100000)
        if info != 0:
            raise RuntimeError(f"oracle CG failed ({info})")
        return x

    def solve(self, constraints: Tensor, forces: Tensor, displacements: Tensor,
              method="auto", rtol=1e-13, params=(), max_iter=10, newton_rtol=1e-8,
              newton_atol=1e-6):
        """FEM.solve for one increment [0, 1] (base.py:597-923) with the
        implicit-function backward of NewtonRaphsonAdjoint (sparse.py:517-724).

        `params` are tensors the material tangent / thickness / loads depend on;
        gradients flow to them through torch autograd exactly as in the
        reference. Returns u [n_nod, nd], f [n_nod, nd], flux mean, grad mean.
        """
        con = torch.nonzero(constraints.ravel(), as_tuple=False).ravel()
        F_ext = forces.ravel()
        DU = displacements.ravel()
        model = self
        track = any(p.requires_grad for p in params)

        def eval_residual(du, need_k):
            du_bc = du.clone()
            du_bc[con] = DU[con]
            k, f_i, _, _ = model.integrate(du_bc, need_k=need_k)
            F_int = model.assemble_rhs(f_i)
            res = F_int - F_ext
            res[con] = 0.0
            return res, k

        class _Newton(torch.autograd.Function):
            @staticmethod
            def forward(ctx, du0, *ps):
                du = du0.detach().clone()
                K = None
                res_norm0 = None
                for i in range(max_iter):
                    with torch.no_grad():
                        res, k = eval_residual(du, need_k=K is None)
                    if K is None:
                        K = model.csr(model.assemble_matrix(k, con))
                    rn = torch.linalg.norm(res)
                    if i == 0:
                        res_norm0 = rn
                    if rn < newton_rtol * res_norm0 or rn < newton_atol:
                        break
                    du_i = model.solve_linear(K, res.numpy(), method, rtol)
                    du = du - torch.as_tensor(du_i, dtype=DT)
                else:
                    raise RuntimeError("Newton-Raphson iteration did not converge.")
                ctx.K = K
                ctx.save_for_backward(du, *ps)
                return du

            @staticmethod
            def backward(ctx, grad_du):
                du, *ps = ctx.saved_tensors
                lam = model.solve_linear(ctx.K.T.tocsr(), grad_du.numpy(), method, rtol)
                lam = torch.as_tensor(lam, dtype=DT)
                du_l = du.detach().requires_grad_(True)
                with torch.enable_grad():
                    res, _ = eval_residual(du_l, need_k=False)
                grads = torch.autograd.grad(res, ps, grad_outputs=-lam, allow_unused=True)
                return (None, *grads)

        du0 = torch.zeros(self.n_dofs, dtype=DT)
        du = _Newton.apply(du0, *params) if track else _Newton.apply(du0)
        du_eval = du.clone()
        du_eval[con] = DU[con]
        _, f_i, grad, flux = self.integrate(du_eval, need_k=False)
        F_int = self.assemble_rhs(f_i)
        nd = self.ndof
        u = du_eval.reshape(-1, nd)
        f = F_int.reshape(-1, nd)
        return u, f, flux.mean(dim=0), grad.mean(dim=0)

    def k0(self):
        """base.py:228-245."""
        k, _, _, _ = self.integrate(torch.zeros(self.n_dofs, dtype=DT), need_k=True)
        return k

    # base.py:452-539 (surface loads, used for the multi-RHS tractions)
    def boundary_facets(self, mask: Tensor):
        table = self.etype.facets
        facets = self.elements[:, table].reshape(-1, table.shape[1])
        facets = facets[mask[facets].all(dim=1)]
        _, inv, count = torch.unique(
            facets.sort(dim=1).values, dim=0, return_inverse=True, return_counts=True
        )
        first = torch.full((len(count),), len(inv))
        first.scatter_reduce_(0, inv, torch.arange(len(inv)), reduce="amin")
        return facets[first[count == 1]]

    def integrate_surface_load(self, mask: Tensor, load: Tensor):
        conn = self.boundary_facets(mask)
        ftype = self.etype.facet_type
        xi = ftype.ipoints
        J = torch.einsum("iaN,eNj->ieaj", ftype.B(xi), self.nodes[conn])
        detJ = torch.sqrt(torch.linalg.det(J @ J.transpose(-1, -2)))
        per_facet = load if load.dim() == 2 else load.reshape(1, -1)
        val = per_facet.expand(len(conn), -1).expand(len(xi), -1, -1)
        contrib = torch.einsum("i,in,ie,iek->enk", ftype.iweights.to(DT), ftype.N(xi), detJ, val)
        out = torch.zeros(self.n_nod, contrib.shape[-1], dtype=DT)
        return out.index_add_(0, conn.ravel(), contrib.flatten(0, 1)), conn


def simp_compliance_sensitivity(model: OracleModel, u: Tensor, rho: Tensor, E0: float, p=3.0):
    """examples/optimization/solid/topology.ipynb cell 8: dC/drho_e =
    -p rho_e^{p-1} E0 u_e^T k0_e u_e with k0 the unit-modulus element matrix."""
    k = model.k0()  # model.mat holds the unit-modulus tangent
    ue = u.reshape(-1)[model.idx]
    return -p * rho ** (p - 1) * E0 * torch.einsum("ei,eij,ej->e", ue, k, ue)


# --------------------------------------------------------------------------
# Hyperelastic materials and the nonlinear solve
# (src/torchfem/materials/hyperelasticity.py:12-140; base.py:597-923, 982-1092)
# --------------------------------------------------------------------------
def neo_hookean_psi(F: Tensor, params: Tensor) -> Tensor:
    """Compressible Neo-Hookean energy psi = mu/2 (tr C - 3) - mu ln J +
    lam/2 (ln J)^2 with params = (mu, lam) and ln J = 1/2 logdet(C), as the
    reference's own test writes it (tests/test_gradients.py:89-94; its second
    derivatives are finite at F = I, where those of det's SVD-based backward
    are not)."""
    mu, lam = params[0], params[1]
    C = F.T @ F
    lnJ = 0.5 * torch.logdet(C)
    return 0.5 * mu * (torch.trace(C) - 3.0) - mu * lnJ + 0.5 * lam * lnJ**2


def hyper_step(psi, F: Tensor, params: Tensor):
    """Hyperelastic3D.step (hyperelasticity.py:83-132): the first
    Piola-Kirchhoff stress P = dpsi/dF and the tangent d2psi/dF2 by automatic
    differentiation, for a batch of deformation gradients [..., 3, 3] and
    per-item parameters [..., p]."""
    from torch.func import jacrev, vmap

    lead = F.shape[:-2]
    Ff = F.reshape(-1, 3, 3)
    pf = params.expand(*lead, params.shape[-1]).reshape(-1, params.shape[-1])
    with torch.enable_grad():
        P = vmap(jacrev(psi))(Ff, pf)
        A = vmap(jacrev(jacrev(psi)))(Ff, pf)
    return P.reshape(*lead, 3, 3), A.reshape(*lead, 3, 3, 3, 3)


def oracle_integrate_incremental(model: "OracleModel", step, grad_prev: Tensor, flux_prev: Tensor,
                                 state_prev: Tensor, du: Tensor, need_k: bool = True):
    """Mechanics.integrate_material for an incremental material
    (base.py:1036-1092): per Gauss point H = du B^T, (stress, state, A) =
    step(H, grad_prev, flux_prev, state_prev), f += w detJ B^T stress,
    k += w detJ B^T A B, F = grad_prev + H. Returns k, f, F, stress, state."""
    _, B, detJ = model.eval_shape_functions(model.etype.ipoints)
    E, npe, d = model.n_elem, model.etype.nodes, model.ndof
    due = du.view(-1, d)[model.elements].reshape(E, -1, d).transpose(-1, -2)
    f = torch.zeros(E, d * npe, dtype=DT)
    k = torch.zeros(E, d * npe, d * npe, dtype=DT) if need_k else None
    Fs, Ps, Ss = [], [], []
    for i, w in enumerate(model.etype.iweights):
        H = due @ B[i].transpose(-1, -2)
        P, state, A = step(H, grad_prev[i], flux_prev[i], state_prev[i])
        Fs.append(grad_prev[i] + H)
        Ps.append(P)
        Ss.append(state)
        sdet = model._scale(detJ[i])  # detJ, times the thickness for planar models
        fc = torch.einsum("..., ...iI,...Ai->...IA", sdet, B[i], P)
        f = f + float(w) * fc.reshape(-1, d * npe)
        if need_k:
            BCB = torch.einsum("...Jp,...iJkL,...Lq->...piqk", B[i], A, B[i])
            k = k + torch.einsum("j,jkl->jkl", sdet, BCB.reshape(-1, d * npe, d * npe)) * float(w)
    return k, f, torch.stack(Fs), torch.stack(Ps), torch.stack(Ss)


def hyper_material(psi, params: Tensor):
    """Hyperelastic3D.step (hyperelasticity.py:83-127) as an oracle step."""
    def step(H, F_prev, stress, state):
        P, A = hyper_step(psi, F_prev + H, params)
        return P, state, A
    return step


def _embed2(F2: Tensor, f33) -> Tensor:
    F3 = torch.zeros(*F2.shape[:-2], 3, 3, dtype=DT)
    F3[..., :2, :2] = F2
    F3[..., 2, 2] = f33
    return F3


def hyper_plane_strain(psi, params: Tensor):
    """HyperelasticPlaneStrain (hyperelasticity.py:272-330): the 3-D energy at
    diag(F2, 1), in-plane blocks of P and of the tangent."""
    def step(H, F_prev, stress, state):
        P, A = hyper_step(psi, _embed2(F_prev + H, 1.0), params)
        return P[..., :2, :2], state, A[..., :2, :2, :2, :2]
    return step


def hyper_plane_stress(psi, params: Tensor, tol: float = 1e-12, max_iter: int = 50):
    """HyperelasticPlaneStress.step (hyperelasticity.py:191-269): Newton on
    the out-of-plane stretch for P33 = 0 from the stored lam3 - 1, then the
    condensed tangent A_ijkl - A_ij33 A_33kl / A_3333."""
    def step(H, F_prev, stress, state):
        F2 = F_prev + H
        lam3 = 1.0 + state[..., 0]
        for _ in range(max_iter):
            P, A = hyper_step(psi, _embed2(F2, lam3), params)
            r = P[..., 2, 2]
            lam3 = lam3 - r / A[..., 2, 2, 2, 2]
            if float(r.abs().max()) < tol:
                break
        A2 = A[..., :2, :2, :2, :2] - torch.einsum(
            "...ij,...kl->...ijkl", A[..., :2, :2, 2, 2], A[..., 2, 2, :2, :2]) / A[
            ..., 2, 2, 2, 2][..., None, None, None, None]
        return P[..., :2, :2], (lam3 - 1.0)[..., None], A2
    return step


def oracle_integrate_hyper(model: "OracleModel", psi, params: Tensor, grad_prev: Tensor,
                           du: Tensor, need_k: bool = True):
    """oracle_integrate_incremental for a hyperelastic material; params [E, p].
    Returns k, f, F, P."""
    n_gp = len(model.etype.iweights)
    zero = torch.zeros(n_gp, model.n_elem, 3, 3, dtype=DT)
    k, f, F, P, _ = oracle_integrate_incremental(
        model, hyper_material(psi, params), grad_prev, zero,
        torch.zeros(n_gp, model.n_elem, 0, dtype=DT), du, need_k)
    return k, f, F, P


def j2_step(E: float, nu: float, sigma_f, sigma_f_prime, tol: float = 1e-5, max_iter: int = 10):
    """IsotropicPlasticity3D.step (plasticity.py:67-215), restated point by
    point: trial stress sigma_n + C : sym(H), von Mises check
    |dev| > sqrt(2/3) sigma_f(q), the scalar Newton for the plastic
    multiplier dg on |dev| - 2G dg - sqrt(2/3) sigma_f(q + sqrt(2/3) dg) = 0,
    the return sigma = trial - 2G dg n and the consistent tangent
    C - 2G/(1 + sigma_f'/(3G)) n(x)n - 4G^2 dg/|dev| (I_dev - n(x)n)."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    r23 = math.sqrt(2.0 / 3.0)
    I = torch.eye(3, dtype=DT)
    I4 = torch.einsum("ij,kl->ijkl", I, I)
    I4s = 0.5 * (torch.einsum("ik,jl->ijkl", I, I) + torch.einsum("il,jk->ijkl", I, I))
    C = lam * I4 + 2.0 * G * I4s

    def point(H, stress, q):
        eps = 0.5 * (H + H.T)
        trial = stress + torch.einsum("ijkl,kl->ij", C, eps)
        dev = trial - torch.trace(trial) / 3.0 * I
        nrm = float(torch.linalg.norm(dev))
        if nrm - r23 * float(sigma_f(torch.tensor(q, dtype=DT))) <= 0.0:
            return trial, q, C.clone()
        dg, qn = 0.0, q
        for _ in range(max_iter):
            res = nrm - 2.0 * G * dg - r23 * float(sigma_f(torch.tensor(qn, dtype=DT)))
            dg += res / (2.0 * G + 2.0 / 3.0 * float(sigma_f_prime(torch.tensor(qn, dtype=DT))))
            qn = q + r23 * dg
            if abs(res) < tol:
                break
        n = dev / nrm
        nn = torch.einsum("ij,kl->ijkl", n, n)
        a = 2.0 * G / (1.0 + float(sigma_f_prime(torch.tensor(qn, dtype=DT))) / (3.0 * G))
        b = 4.0 * G * G * dg / nrm
        return trial - 2.0 * G * dg * n, qn, C - a * nn - b * (I4s - I4 / 3.0 - nn)

    def step(H, F_prev, stress, state):
        lead = H.shape[:-2]
        Hf, Sf, qf = H.reshape(-1, 3, 3), stress.reshape(-1, 3, 3), state.reshape(-1)
        out = [point(Hf[t], Sf[t], float(qf[t])) for t in range(Hf.shape[0])]
        P = torch.stack([o[0] for o in out]).reshape(*lead, 3, 3)
        q = torch.tensor([o[1] for o in out], dtype=DT).reshape(*lead, 1)
        A = torch.stack([o[2] for o in out]).reshape(*lead, 3, 3, 3, 3)
        return P, q, A
    return step


def oracle_solve_incremental(model: "OracleModel", step, n_state: int, constraints: Tensor,
                             forces: Tensor, displacements: Tensor, increments,
                             nlgeom: bool = True, rtol: float = 1e-10, atol: float = 1e-10,
                             max_iter: int = 30):
    """FEM.solve's incremental Newton loop (base.py:700-880) without cutbacks:
    per increment n the loads are lam_n F_ext and (lam_n - lam_{n-1}) U_D,
    Newton starts from the previous increment (du = du_{n-1}), K is
    reassembled at every evaluation with nlgeom or internal state and kept
    from the first evaluation otherwise (base.py:1031-1033), and
    du -= K^{-1} R until |R| < rtol |R_0| or |R| < atol (sparse.py:572-623).
    The stored flux is the stress the step returns (Cauchy F P / J with
    nlgeom) and is passed to the next increment's step. Returns the final u,
    f (internal forces incl. reactions), the Gauss-point means of the flux,
    of F and of the state."""
    con = torch.nonzero(constraints.ravel(), as_tuple=False).ravel()
    n = model.n_dofs
    n_gp = len(model.etype.iweights)
    d = model.ndof
    grad = torch.eye(d, dtype=DT).expand(n_gp, model.n_elem, d, d).clone()
    flux = torch.zeros(n_gp, model.n_elem, d, d, dtype=DT)
    state = torch.zeros(n_gp, model.n_elem, n_state, dtype=DT)
    u = torch.zeros(n, dtype=DT)
    du = torch.zeros(n, dtype=DT)
    K = None
    lam_prev = float(increments[0])
    f_int = torch.zeros(n, dtype=DT)
    for lam_n in [float(v) for v in increments[1:]]:
        F_ext = lam_n * forces.ravel()
        DU = (lam_n - lam_prev) * displacements.ravel()
        res0 = None
        for it in range(max_iter):
            du_bc = du.clone()
            du_bc[con] = DU[con]
            need_k = K is None or nlgeom or n_state > 0
            k, fe, _, _, _ = oracle_integrate_incremental(model, step, grad, flux, state, du_bc,
                                                          need_k)
            if need_k:
                K = model.csr(model.assemble_matrix(k, con))
            R = model.assemble_rhs(fe) - F_ext
            R[con] = 0.0
            rn = float(torch.linalg.norm(R))
            res0 = rn if res0 is None else res0
            if rn < rtol * res0 or rn < atol:
                break
            du = du - torch.as_tensor(spla.spsolve(K.tocsc(), R.numpy()), dtype=DT)
        else:
            raise RuntimeError("Newton-Raphson iteration did not converge.")
        du[con] = DU[con]
        _, fe, grad, P, state = oracle_integrate_incremental(model, step, grad, flux, state, du,
                                                             need_k=False)
        flux = grad @ P / torch.linalg.det(grad)[..., None, None] if nlgeom else P
        f_int = model.assemble_rhs(fe)
        u = u + du
        lam_prev = lam_n
    return u.reshape(-1, d), f_int.reshape(-1, d), flux.mean(0), grad.mean(0), state.mean(0)


def oracle_solve_hyper(model: "OracleModel", psi, params: Tensor, constraints: Tensor,
                       forces: Tensor, displacements: Tensor, increments, nlgeom: bool = True,
                       rtol: float = 1e-10, atol: float = 1e-10, max_iter: int = 30):
    """oracle_solve_incremental for a hyperelastic material; params [E, p].
    Returns u, f, mean flux, mean F."""
    u, f, flux, grad, _ = oracle_solve_incremental(
        model, hyper_material(psi, params), 0, constraints, forces, displacements, increments,
        nlgeom, rtol, atol, max_iter)
    return u, f, flux, grad


# --------------------------------------------------------------------------
# Mass matrices and implicit time integration
# (src/torchfem/base.py:376-396, 1288-1530)
# --------------------------------------------------------------------------
def oracle_mass(model: "OracleModel", rho: Tensor) -> Tensor:
    """FEM.integrate_mass (base.py:376-396): m += w (rho detJ [thickness])
    N_I N_J per DOF component. rho [E]."""
    N, _, detJ = model.eval_shape_functions(model.etype.ipoints)
    nd, npe = model.ndof, model.etype.nodes
    m = torch.zeros(model.n_elem, nd * npe, nd * npe, dtype=DT)
    I_nd = torch.eye(nd, dtype=DT)
    for i, w in enumerate(model.etype.iweights):
        ms = torch.einsum("N,M,E->ENM", N[i], N[i], rho * model._scale(detJ[i]))
        m = m + float(w) * torch.einsum("Enm,ij->Enimj", ms, I_nd).reshape(m.shape)
    return m


def oracle_time_integration(model: "OracleModel", rho: Tensor, constraints: Tensor,
                            prescribed: Tensor, f_ext: Tensor, t_output: Tensor, delta_t: float,
                            rtol: float = 1e-8, atol: float = 1e-6, max_iter: int = 100):
    """FEM.time_integration (base.py:1288-1530) for a linear material: the
    state at t = 0 is `prescribed` everywhere; internal steps of at most
    delta_t subdivide every interval between the output times (and 0); each
    step solves M du + dt/2 (f_old + f_int + f_ext) = 0 by Newton with
    (M + dt/2 K)_bc du = -R, where M and K carry the Dirichlet identity rows
    and columns, f_int = K u (unconstrained) and f_old of the first step is
    zero (f[0] is never filled, base.py:1410-1420). Returns u at the output
    times [n_out, n_dofs]."""
    con = torch.nonzero(constraints.ravel(), as_tuple=False).ravel()
    K_el = model.k0()
    M_el = oracle_mass(model, rho)
    Kf = model.csr(model.assemble_matrix(K_el, torch.zeros(0, dtype=torch.int64)))
    Mb = model.csr(model.assemble_matrix(M_el, con))
    knots = [float(v) for v in t_output]
    if knots[0] > 0.0:
        knots = [0.0] + knots
    times = [knots[0]]
    rows = [] if float(t_output[0]) > 0.0 else [0]
    for t0, t1 in zip(knots[:-1], knots[1:]):
        ratio = (t1 - t0) / delta_t
        n_sub = max(1, math.ceil(ratio - 1e-9 * max(1.0, ratio)))
        times += list(np.linspace(t0, t1, n_sub + 1)[1:-1]) + [t1]
        rows.append(len(times) - 1)
    u = [prescribed.ravel().clone()]
    f = [torch.zeros(model.n_dofs, dtype=DT)]
    for n in range(1, len(times)):
        dt = times[n] - times[n - 1]
        A = model.csr(model.assemble_matrix(M_el + 0.5 * dt * K_el, con)).tocsc()
        guess = u[-1].clone()
        res0 = None
        for _ in range(max_iter):
            du = guess - u[-1]
            f_int = torch.as_tensor(Kf @ guess.numpy(), dtype=DT)
            R = torch.as_tensor(Mb @ du.numpy(), dtype=DT) + 0.5 * dt * (f[-1] + f_int + f_ext.ravel())
            R[con] = 0.0
            rn = float(torch.linalg.norm(R))
            res0 = rn if res0 is None else res0
            if rn < rtol * res0 or rn < atol:
                break
            guess = guess + torch.as_tensor(spla.spsolve(A, -R.numpy()), dtype=DT)
        else:
            raise RuntimeError("Newton-Raphson iteration did not converge.")
        u.append(guess)
        f.append(f_int)
    return torch.stack(u)[torch.tensor(rows)]
