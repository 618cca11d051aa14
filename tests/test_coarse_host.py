"""Host index logic of the two-level preconditioner's coarse space
(torchfem._coarse.aggregates), on the CPU: every node in one aggregate, the
aggregate lists and centroids against brute force, the 27-slot neighbour
table consistent in both directions, and every node block of the adjacency
in exactly one (aggregate, slot) group whose slot names the aggregate of its
column, in ascending order."""

import itertools

import pytest
import torch

import oracle as O
from torchfem._coarse import aggregates

DT = torch.float64


def adjacency(n_nod, elements):
    """Node adjacency (with the node itself), rows ascending, from the
    element connectivity by brute force."""
    nb = [set([i]) for i in range(n_nod)]
    for e in elements.tolist():
        for a, b in itertools.product(e, e):
            nb[a].add(b)
    ptr = [0]
    adj = []
    for s in nb:
        adj.extend(sorted(s))
        ptr.append(len(adj))
    return torch.tensor(ptr, dtype=torch.int64), torch.tensor(adj, dtype=torch.int32)


def meshes():
    nodes, el = O.cube_hexa(9, 7, 6, 2.0, 1.0, 0.8)
    inner = ((nodes > 0) & (nodes < nodes.max(0).values)).all(dim=1)
    yield "hexa", nodes + 0.02 * torch.sin(5.0 * nodes[:, [1, 2, 0]]) * inner[:, None], el, 3
    yield "tetra-heat", *O.cube_tetra6(6, jitter=0.1, seed=4), 1
    yield "quad", *O.rect_quad(31, 9, 3.0, 1.0), 2


@pytest.mark.parametrize("case", list(range(3)))
@pytest.mark.parametrize("size", [None, 60])
def test_aggregates(case, size):
    name, X, el, ndof = list(meshes())[case]
    X = X.to(DT)
    n_nod = X.shape[0]
    ptr, adj = adjacency(n_nod, el)
    a = aggregates(X, el, ptr, adj, n_nod * ndof, ndof, size)
    agg = a["agg"].long()
    agg_ptr = a["agg_ptr"].long()
    n_agg = agg_ptr.numel() - 1
    assert a["nm"] == {3: 6, 2: 3, 1: 1}[ndof]
    assert n_agg >= 1 and int(agg.min()) == 0 and int(agg.max()) == n_agg - 1
    # aggregate lists: the nodes of aggregate I ascending, every node once
    nodes_of = [[] for _ in range(n_agg)]
    for i, g in enumerate(agg.tolist()):
        nodes_of[g].append(i)
    assert a["agg_nodes"].tolist() == [i for lst in nodes_of for i in lst]
    assert agg_ptr.tolist() == [0] + list(itertools.accumulate(len(lst) for lst in nodes_of))
    for g, lst in enumerate(nodes_of):
        assert torch.allclose(a["cen"][g], X[lst].mean(0), rtol=0, atol=1e-12)
    # the slot table: slot 13 is the aggregate itself, and J in slot s of I
    # means I in slot 26 - s of J
    slot_agg = a["slot_agg"].long().reshape(n_agg, 27)
    assert slot_agg[:, 13].tolist() == list(range(n_agg))
    for I in range(n_agg):
        for s in range(27):
            J = int(slot_agg[I, s])
            if J >= 0:
                assert int(slot_agg[J, 26 - s]) == I
    # every node block (i, p) once, in the group (agg(i), slot) whose slot
    # holds agg(j), ascending within the group
    pair_off = a["pair_off"].long()
    pn, pp = a["pair_node"].long(), a["pair_pos"].long()
    assert int(pair_off[-1]) == adj.numel()
    seen = torch.zeros(adj.numel(), dtype=torch.int64)
    for g in range(n_agg * 27):
        I, s = divmod(g, 27)
        lo, hi = int(pair_off[g]), int(pair_off[g + 1])
        keys = []
        for q in range(lo, hi):
            i, p = int(pn[q]), int(pp[q])
            assert 0 <= p < int(ptr[i + 1] - ptr[i])
            j = int(adj[int(ptr[i]) + p])
            assert int(agg[i]) == I and int(slot_agg[I, s]) == int(agg[j])
            seen[int(ptr[i]) + p] += 1
            keys.append((i, p))
        assert keys == sorted(keys)
    assert bool((seen == 1).all())
    if size is not None:
        # the cells are cubic, so the count rounds up, by at most 2x per axis
        assert n_agg * a["nm"] <= 8 * size


def test_cells_are_wider_than_elements():
    """A coarse size asking for cells smaller than an element gets cells of
    at least 1.01 element extents, so every element stays within two cells
    per axis (the 27-slot stencil)."""
    X, el = O.rect_quad(11, 3, 10.0, 1.0)
    ptr, adj = adjacency(X.shape[0], el)
    a = aggregates(X.to(DT), el, ptr, adj, X.shape[0] * 2, 2, size=10**6)
    # cells of at least 1.01 h: at most 10 along x for 10 elements
    assert int(a["agg_ptr"].numel()) - 1 <= 10


def test_preconditioner_resolution():
    """preconditioner=None takes the default, "auto": the two-level
    preconditioner from AUTO_TWO_LEVEL_DOFS unknowns, block-Jacobi below;
    explicit names pass through; set_default_preconditioner changes the
    default; unknown names raise ValueError."""
    from types import SimpleNamespace

    from torchfem import sparse
    from torchfem.base import FEM

    small = SimpleNamespace(n_dofs=sparse.AUTO_TWO_LEVEL_DOFS - 1)
    big = SimpleNamespace(n_dofs=sparse.AUTO_TWO_LEVEL_DOFS)
    assert sparse.default_preconditioner() == "auto"
    assert FEM._resolve_pc(small, None) == "block_jacobi"
    assert FEM._resolve_pc(big, None) == "two_level"
    assert FEM._resolve_pc(big, "auto") == "two_level"
    assert FEM._resolve_pc(big, "jacobi") == "jacobi"
    assert FEM._resolve_pc(small, "two_level") == "two_level"
    with pytest.raises(ValueError):
        FEM._resolve_pc(big, "amg")
    sparse.set_default_preconditioner("block_jacobi")
    try:
        assert FEM._resolve_pc(big, None) == "block_jacobi"
    finally:
        sparse.set_default_preconditioner("auto")
    with pytest.raises(ValueError):
        sparse.set_default_preconditioner("multigrid")
