"""Structured meshes (src/torchfem/mesh.py) plus the six-tetrahedron split.

`cube_hexa` and `rect_quad` number nodes and elements exactly like the
reference (mesh.py:8-46, 103-135), so results can be compared node by node.
"""

from __future__ import annotations

import torch
from torch import Tensor


def cube_hexa(Nx: int, Ny: int, Nz: int, Lx: float = 1.0, Ly: float = 1.0,
              Lz: float = 1.0) -> tuple[Tensor, Tensor]:
    """Structured Hexa1 mesh; node (i, j, k) is i*Ny*Nz + j*Nz + k and each
    cell lists its corners as [n0, n1, n3, n2, n4, n5, n7, n6] in terms of the
    lexicographic corner index (x fastest)."""
    axes = [torch.linspace(0, L, N, dtype=torch.float64) for N, L in ((Nx, Lx), (Ny, Ly), (Nz, Lz))]
    grid = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)
    nodes = grid.reshape(-1, 3)
    ind = torch.arange(Nx * Ny * Nz).reshape(Nx, Ny, Nz)
    c = {(a, b, d): ind[a:Nx - 1 + a, b:Ny - 1 + b, d:Nz - 1 + d].reshape(-1)
         for a in (0, 1) for b in (0, 1) for d in (0, 1)}
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    return nodes, torch.stack([c[o] for o in order], dim=1)


def rect_quad(Nx: int, Ny: int, Lx: float = 1.0, Ly: float = 1.0) -> tuple[Tensor, Tensor]:
    """Structured Quad1 mesh; node (i, j) is i*Ny + j, cells counter-clockwise."""
    axes = [torch.linspace(0, L, N, dtype=torch.float64) for N, L in ((Nx, Lx), (Ny, Ly))]
    nodes = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1).reshape(-1, 2)
    ind = torch.arange(Nx * Ny).reshape(Nx, Ny)
    c = {(a, b): ind[a:Nx - 1 + a, b:Ny - 1 + b].reshape(-1) for a in (0, 1) for b in (0, 1)}
    return nodes, torch.stack([c[(0, 0)], c[(1, 0)], c[(1, 1)], c[(0, 1)]], dim=1)


def cube_tetra6(N: int, jitter: float = 0.0, seed: int = 0) -> tuple[Tensor, Tensor]:
    """Kuhn split of an N^3-cell unit cube into 6 conforming Tetra1 per cell.

    The reference's cube_tetra (mesh.py:49-100) uses 5 tetrahedra per cell;
    this split instead cuts every cell along the same main diagonal, one
    tetrahedron per ordering of the three axes, so the mesh is conforming
    without a checkerboard. Nodes are numbered as in cube_hexa(N+1, N+1, N+1);
    interior nodes are moved by U(-jitter*h, jitter*h) per coordinate (h = 1/N)
    from a torch generator seeded with `seed`.
    """
    nodes, _ = cube_hexa(N + 1, N + 1, N + 1)
    n1 = N + 1
    stride = (n1 * n1, n1, 1)
    ii, jj, kk = torch.meshgrid(torch.arange(N), torch.arange(N), torch.arange(N), indexing="ij")
    v0 = (ii * stride[0] + jj * stride[1] + kk * stride[2]).reshape(-1)
    tets = []
    # even permutations keep the orientation positive; odd ones swap two nodes
    for a, b, c, odd in [(0, 1, 2, 0), (1, 2, 0, 0), (2, 0, 1, 0), (0, 2, 1, 1), (2, 1, 0, 1),
                         (1, 0, 2, 1)]:
        v1 = v0 + stride[a]
        v2 = v1 + stride[b]
        v3 = v2 + stride[c]
        tets.append(torch.stack([v0, v2, v1, v3] if odd else [v0, v1, v2, v3], dim=1))
    elements = torch.stack(tets, dim=1).reshape(-1, 4)
    if jitter > 0.0:
        g = torch.Generator().manual_seed(seed)
        shift = (torch.rand(nodes.shape, generator=g, dtype=torch.float64) * 2.0 - 1.0) * jitter * (1.0 / N)
        interior = ((nodes > 0.0) & (nodes < 1.0)).all(dim=1)
        nodes = nodes + shift * interior[:, None]
    return nodes, elements
