"""Coarse space of the two-level preconditioner (TFM_PC_TWO_LEVEL), in the
place of the aggregation AMG the reference reaches through AmgX
(sparse.py:422-442, amgx.py:71-98).

Nodes are grouped into aggregates by a grid of cubic cells over the mesh's
bounding box. The cell is at least 1.01 times the largest element extent, so
K couples an aggregate only to the 27 cells around it. The cells are sized
for about one coarse unknown per 128 fine ones and at most COARSE_SIZE
(256 aggregates with 6 rigid-body modes each in 3-D, 512 with 3 in 2-D,
1,536 with the constant for a scalar field). The index work below runs once
per pattern with torch on the device. The coarse operator P^T K P is formed
per system by a HIP kernel (tfm_coarse_operator), and its dense inverse by
the library's block Gauss-Jordan (tfm_spd_inverse) on the device.
"""

from __future__ import annotations

import ctypes as C
import math
import weakref

import torch
from torch import Tensor

from . import _lib as L

COARSE_SIZE = 1536  # most coarse unknowns (timed on the 64^3 cube: 768, 1536, 3072)


def spd_inverse(A: Tensor, stream=None) -> Tensor | None:
    """A^{-1} of a dense SPD FP64 matrix on the device, in place
    (tfm_spd_inverse: block Gauss-Jordan); None if a pivot is not positive."""
    n = A.shape[0]
    count = L.I64()
    L.check(L.lib().tfm_spd_inverse_work(n, C.byref(count)))
    work = torch.empty(int(count.value), dtype=torch.float64, device=A.device)
    status = torch.zeros(1, dtype=torch.int32, device=A.device)
    if stream is None:
        stream = L.stream_of(A.device)
    L.check(L.lib().tfm_spd_inverse(stream, n, L.ptr(A), L.ptr(work), L.ptr(status)))
    return A if int(status) == 0 else None


def aggregates(X: Tensor, elements: Tensor, adj_ptr: Tensor, adj: Tensor, n_dofs: int,
               ndof: int, size: int | None = None) -> dict:
    """The index arrays of the coarse space, on X's device: node coordinates
    X [n_nod, dim], connectivity, and the node adjacency (adj_ptr [n_nod+1],
    adj: the neighbour nodes of every node, in the pattern's order). Returns
    agg [n_nod] (the aggregate of each node), agg_ptr / agg_nodes (the nodes
    of each aggregate, ascending), cen [n_agg, dim] (centroids), slot_agg
    [n_agg, 27] (the aggregate in each neighbour cell, -1 for none),
    pair_off [n_agg*27+1] / pair_node / pair_pos (every node block (i, p),
    node i and position p in its adjacency, grouped by the aggregate of i
    and the slot of the aggregate of its column, stable in (i, p)), and nm,
    the modes per aggregate."""
    dev = X.device
    D = X.shape[1]
    n_nod = X.shape[0]
    nm = {3: 6, 2: 3, 1: 1}[ndof]
    if size is None:
        # about one coarse unknown per 128 fine ones, at most COARSE_SIZE
        size = min(COARSE_SIZE, max(nm, n_dofs // 128))
    target = max(1, size // nm)
    lo = X.min(0).values
    ext = X.max(0).values - lo
    if elements.shape[0]:
        Xe = X[elements.long()]
        emax = float((Xe.max(1).values - Xe.min(1).values).max())
    else:
        emax = float(ext.max()) if n_nod else 1.0
    emax = max(emax, 1e-300)
    ext_eff = torch.clamp(ext, min=emax)
    vol = float(torch.prod(ext_eff))
    cell = max((vol / target) ** (1.0 / D), 1.01 * emax)
    nc = [max(1, int(math.ceil(float(e) / cell))) for e in ext_eff]
    nc_t = torch.tensor(nc, device=dev)
    cidx = torch.minimum(torch.floor((X - lo) / cell).long().clamp(min=0), nc_t - 1)
    if D == 2:
        cidx = torch.cat([cidx, torch.zeros_like(cidx[:, :1])], 1)
        nc = nc + [1]
    lin = (cidx[:, 0] * nc[1] + cidx[:, 1]) * nc[2] + cidx[:, 2]
    uniq, agg = torch.unique(lin, return_inverse=True)
    n_agg = int(uniq.numel())
    # nodes of each aggregate, ascending
    agg_nodes = torch.sort(agg, stable=True).indices
    counts = torch.bincount(agg, minlength=n_agg)
    agg_ptr = torch.zeros(n_agg + 1, dtype=torch.int64, device=dev)
    agg_ptr[1:] = torch.cumsum(counts, 0)
    # centroids: segment sums over the aggregate-sorted nodes (an ordered
    # scan, repeatable bit for bit, where index_add_ adds with atomics)
    csum = torch.cat([torch.zeros(D, 1, dtype=torch.float64, device=dev),
                      torch.cumsum(X[agg_nodes].double().T.contiguous(), 1)], 1)
    cen = ((csum[:, agg_ptr[1:]] - csum[:, agg_ptr[:-1]]) / counts.double()).T.contiguous()
    acell = torch.stack([uniq // (nc[1] * nc[2]), (uniq // nc[2]) % nc[1], uniq % nc[2]], 1)
    # cell -> aggregate, then the aggregate of each of the 27 slots
    cmap = torch.full((nc[0] * nc[1] * nc[2],), -1, dtype=torch.int64, device=dev)
    cmap[uniq] = torch.arange(n_agg, device=dev)
    off = torch.stack(torch.meshgrid(*[torch.arange(-1, 2, device=dev)] * 3, indexing="ij"),
                      -1).reshape(27, 3)
    nb = acell[:, None, :] + off[None, :, :]
    ok = ((nb >= 0) & (nb < torch.tensor(nc, device=dev))).all(-1)
    nbl = (nb[..., 0] * nc[1] + nb[..., 1]) * nc[2] + nb[..., 2]
    slot_agg = torch.where(ok, cmap[nbl.clamp(min=0, max=cmap.numel() - 1)],
                           torch.full_like(nbl, -1))
    # node blocks (i, p) grouped by (aggregate of i, slot of the aggregate of j)
    deg = (adj_ptr[1:] - adj_ptr[:-1]).long()
    rows = torch.repeat_interleave(torch.arange(n_nod, device=dev), deg)
    pos = torch.arange(rows.numel(), device=dev) - adj_ptr[:-1].long()[rows]
    cols = adj.long()
    d = acell[agg[cols]] - acell[agg[rows]] + 1
    if rows.numel() and (int(d.min()) < 0 or int(d.max()) > 2):
        raise RuntimeError("an element spans more than two aggregate cells")
    slot = (d[:, 0] * 3 + d[:, 1]) * 3 + d[:, 2]
    key = agg[rows] * 27 + slot
    order = torch.sort(key, stable=True).indices
    kc = torch.bincount(key, minlength=n_agg * 27)
    pair_off = torch.zeros(n_agg * 27 + 1, dtype=torch.int64, device=dev)
    pair_off[1:] = torch.cumsum(kc, 0)
    return dict(
        nm=nm,
        agg=agg.to(torch.int32).contiguous(),
        agg_ptr=agg_ptr.to(torch.int32).contiguous(),
        agg_nodes=agg_nodes.to(torch.int32).contiguous(),
        cen=cen,
        slot_agg=slot_agg.to(torch.int32).contiguous(),
        pair_off=pair_off.to(torch.int32).contiguous(),
        pair_node=rows[order].to(torch.int32).contiguous(),
        pair_pos=pos[order].to(torch.int32).contiguous(),
    )


class CoarseSpace:
    """Aggregates of one node-block pattern and their C-ABI handle."""

    def __init__(self, pat, size: int | None = None):
        a = aggregates(pat.nodes, pat.elements, pat.adj_ptr, pat.adj, pat.n_dofs, pat.ndof, size)
        self.nm = a["nm"]
        n_agg = int(a["agg_ptr"].numel()) - 1
        self.n_agg = n_agg
        self.agg = a["agg"]
        self._arrays = a
        h = L.P()
        with torch.cuda.device(pat.device):
            L.check(L.lib().tfm_coarse_create(
                C.byref(h), C.byref(pat.mesh), n_agg, L.ptr(a["agg"]), L.ptr(a["agg_ptr"]),
                L.ptr(a["agg_nodes"]), L.ptr(a["cen"]), L.ptr(a["slot_agg"]),
                L.ptr(a["pair_off"]), int(a["pair_node"].numel()), L.ptr(a["pair_node"]),
                L.ptr(a["pair_pos"])))
        self.h = h
        self._release = weakref.finalize(self, L.lib().tfm_coarse_destroy, C.c_void_p(h.value))
        nc_out = L.I32()
        L.check(L.lib().tfm_coarse_size(h, C.byref(nc_out)))
        self.n_c = int(nc_out.value)
        self.pat = pat

    def inverse(self, system) -> tuple[Tensor, Tensor]:
        """(P^T K P)^{-1} of an assembled system, FP64 and FP32, symmetric
        positive definite. The operator is scaled to a unit diagonal first
        (aggregates of very different stiffness, as in density designs with
        void elements, and the translation / rotation scales then meet at one
        scale); coarse modes that vanish (aggregates whose DOFs are all
        constrained) get a unit diagonal, and a 1e-12 shift keeps the factor
        defined when an aggregate's modes are dependent (their prolongation is
        zero either way). The inverse is the library's block Gauss-Jordan
        (tfm_spd_inverse); if one of its pivots is not positive, the
        eigenvalues below 1e-12 of the largest are dropped instead."""
        pat = self.pat
        Ac = torch.empty(self.n_c, self.n_c, dtype=torch.float64, device=pat.device)
        L.check(L.lib().tfm_coarse_operator(pat.stream(), self.h, C.byref(system.sys), L.ptr(Ac)))
        Ac = 0.5 * (Ac + Ac.T)
        dg = Ac.diagonal().clone()
        top = float(dg.abs().max()) if self.n_c else 1.0
        top = top if top > 0.0 else 1.0
        dead = dg <= 1e-14 * top
        dg = torch.where(dead, torch.ones_like(dg), dg)
        sc = 1.0 / torch.sqrt(dg)
        A = sc[:, None] * Ac * sc[None, :]
        A.diagonal().copy_(torch.where(dead, torch.ones_like(dg), A.diagonal() + 1e-12))
        inv = spd_inverse(A.clone(), pat.stream())
        if inv is None:
            w, V = torch.linalg.eigh(A)
            w = torch.where(w > 1e-12 * float(w.abs().max()), 1.0 / w, torch.zeros_like(w))
            inv = (V * w) @ V.T
        inv = sc[:, None] * (0.5 * (inv + inv.T)) * sc[None, :]
        inv = inv.contiguous()
        return inv, inv.float().contiguous()
