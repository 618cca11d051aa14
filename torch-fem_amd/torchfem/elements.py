"""Element types of the assemble-and-solve path (src/torchfem/elements.py).

The class names, node orderings, Gauss rules and facet tables are those of the
reference (Quad1 elements.py:491-567, Tetra1 727-790, Hexa1 936-1075, Hexa2
1078-1406, and their facet types Bar1 172-213, Tria1 310-360, Quad2 599-695),
so a mesh built for the reference works unchanged. The shape functions are
written once per family from the nodal reference coordinates:

* tensor-product vertices: N = prod_k (1 + x_k c_k) / 2^d,
* serendipity vertices: N = prod_k (1 + x_k c_k) (sum_k x_k c_k - (d-1)) / 2^d,
* serendipity mid-sides (c_k = 0 on one axis): N = prod_k f_k / 2^(d-1) with
  f_k = 1 - x_k^2 on that axis,
* simplices: barycentric coordinates.

These host-side tables serve the API (eval_shape_functions, facet loads); the
HIP kernels carry their own compile-time tables of the same functions.
"""

from __future__ import annotations

import math

import numpy as np
import torch
from torch import Tensor

_G = 1.0 / math.sqrt(3.0)


def _f64(x) -> Tensor:
    return torch.tensor(x, dtype=torch.float64)


def _gauss_tensor(dim: int) -> Tensor:
    pts = [[0.0]]
    for _ in range(dim):
        pts = [p + [s * _G] for s in (-1.0, 1.0) for p in pts]
    # the first coordinate varies fastest, as in the reference's ipoints
    return _f64([[p[k] for k in range(1, dim + 1)] for p in pts])


def _axis(x: Tensor, c: float) -> tuple[Tensor, Tensor]:
    """Factor along one axis and its derivative."""
    if c == 0.0:
        return 1.0 - x * x, -2.0 * x
    return 1.0 + x * c, torch.full_like(x, c)


class Element:
    """Isoparametric element with N(xi) [.., nodes] and B(xi) = dN/dxi [.., iso_dim, nodes]."""

    iso_volume: float
    iso_dim: int
    nodes: int
    facet_type: type["Element"]
    serendipity = False
    simplex = False

    @classmethod
    def _coords(cls) -> list[list[float]]:
        return cls.iso_coords.tolist()

    @classmethod
    def N(cls, xi: Tensor) -> Tensor:
        xi = torch.as_tensor(xi)
        if cls.simplex:
            rest = 1.0 - xi.sum(dim=-1, keepdim=True)
            return torch.cat([rest, xi], dim=-1)
        d = cls.iso_dim
        out = []
        for c in cls._coords():
            f = [_axis(xi[..., k], c[k])[0] for k in range(d)]
            prod = f[0]
            for k in range(1, d):
                prod = prod * f[k]
            if cls.serendipity and all(v != 0.0 for v in c):
                s = sum(xi[..., k] * c[k] for k in range(d)) - (d - 1)
                out.append(prod * s / 2**d)
            elif cls.serendipity:
                out.append(prod / 2 ** (d - 1))
            else:
                out.append(prod / 2**d)
        return torch.stack(out, dim=-1)

    @classmethod
    def B(cls, xi: Tensor) -> Tensor:
        xi = torch.as_tensor(xi)
        d = cls.iso_dim
        if cls.simplex:
            b = torch.cat([-torch.ones(d, 1), torch.eye(d)], dim=1).to(xi.dtype)
            return b if xi.dim() == 1 else b.expand(*xi.shape[:-1], d, d + 1).clone()
        rows = [[] for _ in range(d)]
        for c in cls._coords():
            fd = [_axis(xi[..., k], c[k]) for k in range(d)]
            vertex = all(v != 0.0 for v in c)
            for k in range(d):
                prod_other = torch.ones_like(xi[..., 0])
                for m in range(d):
                    if m != k:
                        prod_other = prod_other * fd[m][0]
                if cls.serendipity and vertex:
                    s = sum(xi[..., m] * c[m] for m in range(d)) - (d - 1)
                    val = (fd[k][1] * prod_other * s + fd[k][0] * prod_other * c[k]) / 2**d
                elif cls.serendipity:
                    val = fd[k][1] * prod_other / 2 ** (d - 1)
                else:
                    val = fd[k][1] * prod_other / 2**d
                rows[k].append(val)
        return torch.stack([torch.stack(r, dim=-1) for r in rows], dim=-2)


class Bar1(Element):
    """Two-node line (elements.py:172-213); 1-point rule."""

    iso_volume, iso_dim, nodes = 2.0, 1, 2
    iso_coords = _f64([[-1.0], [1.0]])
    ipoints = _f64([[0.0]])
    iweights = _f64([2.0])


class Tria1(Element):
    """Three-node triangle (elements.py:310-360); 1-point rule."""

    iso_volume, iso_dim, nodes = 0.5, 2, 3
    simplex = True
    iso_coords = _f64([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    facets = torch.tensor([[0, 1], [1, 2], [2, 0]])
    ipoints = _f64([[1.0 / 3.0, 1.0 / 3.0]])
    iweights = _f64([0.5])
    facet_type = Bar1


class Quad1(Element):
    """Four-node bilinear quadrilateral (elements.py:491-567)."""

    iso_volume, iso_dim, nodes = 4.0, 2, 4
    iso_coords = _f64([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
    facets = torch.tensor([[0, 1], [1, 2], [2, 3], [3, 0]])
    edges = facets
    ipoints = _gauss_tensor(2)
    iweights = torch.tensor([1, 1, 1, 1])  # integer weights, as in the reference
    facet_type = Bar1


class Quad2(Quad1):
    """Eight-node serendipity quadrilateral (elements.py:599-695)."""

    nodes = 8
    serendipity = True
    iso_coords = _f64(
        [[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0],
         [0.0, -1.0], [1.0, 0.0], [0.0, 1.0], [-1.0, 0.0]]
    )
    facets = torch.tensor([[0, 1, 4], [1, 2, 5], [2, 3, 6], [3, 0, 7]])


class Tetra1(Element):
    """Four-node linear tetrahedron (elements.py:727-790)."""

    iso_volume, iso_dim, nodes = 1.0 / 6.0, 3, 4
    simplex = True
    iso_coords = _f64(
        [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    )
    facets = torch.tensor([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])
    edges = torch.tensor([[0, 1], [1, 2], [0, 2], [3, 0], [1, 3], [2, 3]])
    ipoints = _f64([[0.25, 0.25, 0.25]])
    iweights = _f64([1.0 / 6.0])
    facet_type = Tria1


class Hexa1(Element):
    """Eight-node trilinear hexahedron (elements.py:936-1075)."""

    iso_volume, iso_dim, nodes = 8.0, 3, 8
    iso_coords = _f64(
        [[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [1.0, 1.0, -1.0], [-1.0, 1.0, -1.0],
         [-1.0, -1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, 1.0], [-1.0, 1.0, 1.0]]
    )
    facets = torch.tensor(
        [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]
    )
    edges = torch.tensor(
        [[0, 1], [1, 2], [2, 3], [3, 0], [4, 5], [5, 6], [6, 7], [7, 4],
         [0, 4], [1, 5], [2, 6], [3, 7]]
    )
    ipoints = _gauss_tensor(3)
    iweights = torch.ones(8, dtype=torch.float64)
    facet_type = Quad1


class Hexa2(Hexa1):
    """Twenty-node serendipity hexahedron (elements.py:1078-1406), 2x2x2 rule."""

    nodes = 20
    serendipity = True
    iso_coords = _f64(
        [[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [1.0, 1.0, -1.0], [-1.0, 1.0, -1.0],
         [-1.0, -1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, 1.0], [-1.0, 1.0, 1.0],
         [0.0, -1.0, -1.0], [1.0, 0.0, -1.0], [0.0, 1.0, -1.0], [-1.0, 0.0, -1.0],
         [0.0, -1.0, 1.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [-1.0, 0.0, 1.0],
         [-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]]
    )
    facets = torch.tensor(
        [[0, 3, 2, 1, 11, 10, 9, 8], [4, 5, 6, 7, 12, 13, 14, 15],
         [0, 1, 5, 4, 8, 17, 12, 16], [1, 2, 6, 5, 9, 18, 13, 17],
         [2, 3, 7, 6, 10, 19, 14, 18], [3, 0, 4, 7, 11, 16, 15, 19]]
    )
    facet_type = Quad2


def linear_etype(nodes: Tensor, elements: Tensor) -> type[Element]:
    """The linear element type of a mesh (elements.py:1409-1435)."""
    n, dim = elements.shape[1], nodes.shape[1]
    if n == 3 and dim == 2:
        return Tria1
    if n == 4 and dim == 2:
        return Quad1
    if n == 4 and dim == 3:
        return Tetra1
    if n == 8 and dim == 3:
        return Hexa1
    raise Exception("The element type is not supported. Maybe the element is already quadratic?")


def linear_to_quadratic(nodes: Tensor, elements: Tensor) -> tuple[Tensor, Tensor]:
    """Hexa1 -> Hexa2 mesh upgrade (elements.py:1438-1494): one new node per
    unique edge at its midpoint, appended after the existing nodes.

    The new nodes are numbered in the order the reference's CPU path gives
    them: unique (low, high) node pairs ordered by their raw int64 bytes, as
    numpy orders a void view of the pairs, so meshes match node for node.
    """
    etype = linear_etype(nodes, elements)
    if etype is not Hexa1:
        raise NotImplementedError("only Hexa1 -> Hexa2 is part of this build")
    pairs = torch.sort(elements[:, etype.edges].reshape(-1, 2), dim=1).values
    pn = np.ascontiguousarray(pairs.cpu().numpy().astype(np.int64))
    rows = pn.view(np.dtype((np.void, 16))).ravel()
    _, first, inv = np.unique(rows, return_index=True, return_inverse=True)
    a = torch.as_tensor(pn[first, 0], device=nodes.device)
    b = torch.as_tensor(pn[first, 1], device=nodes.device)
    mid = (nodes[a] + nodes[b]) / 2.0
    new_nodes = torch.cat([nodes, mid], dim=0)
    mid_map = torch.as_tensor(inv.reshape(elements.shape[0], -1), device=elements.device)
    return new_nodes, torch.cat([elements, mid_map + nodes.shape[0]], dim=1)
