"""Node-block sparsity pattern of a mesh, built once on the device by
libtfm_hip.so (tfm_pattern_device_count/fill) and kept there. The host
builder (tfm_pattern_create) produces the same arrays on the CPU; the tests
use it where no GPU is present.

This replaces the pattern block of FEM.__init__ (base.py:72-122): instead of
sorting the packed (row << 32 | col) keys of every element DOF pair, the host
library forms each node's sorted adjacency; the CSR values of a node's rows
are contiguous and in the reference's row-major order, so the reference's
index arrays (glob_idx, k_map, diag_map) are exact expansions of it
(`expand`).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from . import _lib as L


def build_host_pattern(elements: Tensor, n_nod: int, ndof: int) -> dict[str, np.ndarray]:
    """Runs the host pattern builder; returns numpy arrays (no GPU needed)."""
    lib = L.lib()
    el = np.ascontiguousarray(elements.detach().cpu().numpy().astype(np.int64))
    n_elem, npe = el.shape if el.ndim == 2 else (0, 0)
    h = C.c_void_p()
    L.check(lib.tfm_pattern_create(C.byref(h), el.ctypes.data, n_elem, npe, n_nod, ndof))
    try:
        nnz, n_adj = L.I64(), L.I64()
        max_deg, max_inc = L.I32(), L.I32()
        L.check(lib.tfm_pattern_info(h, C.byref(nnz), C.byref(n_adj), C.byref(max_deg),
                                     C.byref(max_inc)))
        out = {
            "adj_ptr": np.empty(n_nod + 1, np.int32),
            "adj": np.empty(n_adj.value, np.int32),
            "val_ptr": np.empty(n_nod + 1, np.int64),
            "ne_ptr": np.empty(n_nod + 1, np.int32),
            "ne_elem": np.empty(n_elem * npe, np.int32),
            "ne_loc": np.empty(n_elem * npe, np.int32),
            "epos": np.empty(n_elem * npe * npe, np.uint16),
        }
        L.check(lib.tfm_pattern_export(h, *(out[k].ctypes.data for k in
                                             ["adj_ptr", "adj", "val_ptr", "ne_ptr", "ne_elem",
                                              "ne_loc", "epos"])))
        out["nnz"] = int(nnz.value)
        out["max_deg"] = int(max_deg.value)
        out["max_inc"] = int(max_inc.value)
        out["npe"] = int(npe)
        return out
    finally:
        lib.tfm_pattern_destroy(h)


def build_device_pattern(elements: Tensor, n_nod: int, ndof: int, device) -> dict:
    """Runs the device pattern builder on int32 device connectivity; returns
    device tensors (uint16 epos in int16 storage) and the sizes."""
    lib = L.lib()
    el = elements
    n_elem, npe = el.shape
    i32 = dict(dtype=torch.int32, device=device)
    out = {
        "ne_ptr": torch.empty(n_nod + 1, **i32),
        "ne_elem": torch.empty(max(n_elem * npe, 1), **i32),
        "ne_loc": torch.empty(max(n_elem * npe, 1), **i32),
        "adj_ptr": torch.empty(n_nod + 1, **i32),
        "val_ptr": torch.empty(n_nod + 1, dtype=torch.int64, device=device),
    }
    work = torch.empty(3 * n_nod + 16 + 4 * (n_nod // 1024 + 2), **i32)
    info = (L.I64 * 4)()
    stream = L.stream_of(device)
    L.check(lib.tfm_pattern_device_count(stream, L.ptr(el), n_elem, npe, n_nod, ndof,
                                         L.ptr(out["ne_ptr"]), L.ptr(out["ne_elem"]),
                                         L.ptr(out["ne_loc"]), L.ptr(out["adj_ptr"]),
                                         L.ptr(out["val_ptr"]), L.ptr(work), info))
    n_adj = int(info[0])
    out["adj"] = torch.empty(max(n_adj, 1), **i32)
    out["epos"] = torch.empty(max(n_elem * npe * npe, 1), dtype=torch.int16, device=device)
    adj16 = torch.empty(max(n_adj, 1), dtype=torch.int16, device=device)
    fits = L.I32()
    L.check(lib.tfm_pattern_device_fill(stream, L.ptr(el), n_elem, npe, n_nod,
                                        L.ptr(out["ne_ptr"]), L.ptr(out["ne_elem"]),
                                        L.ptr(out["adj_ptr"]), L.ptr(out["adj"]),
                                        L.ptr(out["epos"]), L.ptr(adj16), L.ptr(work),
                                        C.byref(fits)))
    out["adj16"] = adj16 if (fits.value and n_adj > 0) else None
    out["nnz"] = int(info[1])
    out["max_deg"] = int(info[2])
    out["max_inc"] = int(info[3])
    out["npe"] = int(npe)
    return out


class DevicePattern:
    """Mesh geometry and pattern on the device, with the tfm_mesh_t view. The
    pattern is built on the device (tfm_pattern_device_count/fill)."""

    def __init__(self, nodes: Tensor, elements: Tensor, ndof: int, etype: str,
                 device: torch.device):
        self.device = device
        self.n_nod, self.dim = nodes.shape
        self.n_elem = elements.shape[0]
        self.ndof = ndof
        self.etype = etype
        self.nodes = nodes.detach().to(device=device, dtype=torch.float64).contiguous()
        self.elements = elements.detach().to(device=device, dtype=torch.int32).contiguous()
        dev = build_device_pattern(self.elements, self.n_nod, ndof, device)
        self.nnz = dev["nnz"]
        self.max_deg = dev["max_deg"]
        self.max_inc = dev["max_inc"]
        self.npe = dev["npe"] if self.n_elem else elements.shape[1]
        self.adj_ptr = dev["adj_ptr"]
        self.adj = dev["adj"]
        self.val_ptr = dev["val_ptr"]
        self.ne_ptr = dev["ne_ptr"]
        self.ne_elem = dev["ne_elem"]
        self.ne_loc = dev["ne_loc"]
        self.epos = dev["epos"]
        # 16-bit neighbour offsets (adj[k] - row node) when they all fit
        self.adj16 = dev["adj16"]
        self.mesh = L.Mesh(
            nodes=self.nodes.data_ptr(),
            elements=self.elements.data_ptr(),
            n_nod=self.n_nod,
            n_elem=self.n_elem,
            dim=self.dim,
            npe=self.npe,
            ndof=ndof,
            etype=L.ETYPE[etype],
            max_deg=self.max_deg,
            max_inc=self.max_inc,
            nnz=self.nnz,
            adj_ptr=self.adj_ptr.data_ptr(),
            adj=self.adj.data_ptr(),
            val_ptr=self.val_ptr.data_ptr(),
            ne_ptr=self.ne_ptr.data_ptr(),
            ne_elem=self.ne_elem.data_ptr(),
            ne_loc=self.ne_loc.data_ptr(),
            epos=self.epos.data_ptr(),
            adj16=self.adj16.data_ptr() if self.adj16 is not None else None,
        )

    @property
    def n_dofs(self) -> int:
        return self.n_nod * self.ndof

    def stream(self) -> int:
        return L.stream_of(self.device)

    def expand(self, rows=True, cols=True, rowptr=False, k_map=False, diag_map=False):
        """The reference's index arrays, expanded on the device."""
        dev = self.device
        W = self.npe * self.ndof
        out = {
            "rows": torch.empty(self.nnz, dtype=torch.int64, device=dev) if rows else None,
            "cols": torch.empty(self.nnz, dtype=torch.int64, device=dev) if cols else None,
            "rowptr": torch.empty(self.n_dofs + 1, dtype=torch.int64, device=dev) if rowptr else None,
            "k_map": torch.empty(self.n_elem * W * W, dtype=torch.int32, device=dev) if k_map else None,
            "diag_map": torch.empty(self.n_dofs, dtype=torch.int32, device=dev) if diag_map else None,
        }
        L.check(
            L.lib().tfm_pattern_expand(
                self.stream(), C.byref(self.mesh), L.ptr(out["rows"]), L.ptr(out["cols"]),
                L.ptr(out["rowptr"]), L.ptr(out["k_map"]), L.ptr(out["diag_map"]),
            )
        )
        return out


def locality_order(nodes: Tensor, n_elem: int) -> Tensor:
    """A node order with neighbours close together: nodes sorted by the bins
    of a grid whose spacing is the mean element size, the axis with the most
    bins varying slowest (the thinnest slabs), ties kept in the mesh's order.
    On a structured mesh this is the lexicographic order with mid-side nodes
    beside their corners. Returns perm (new index -> mesh index), on the
    nodes' device."""
    x = nodes - nodes.min(dim=0).values
    ext = x.max(dim=0).values
    d = nodes.shape[1]
    vol = float(torch.prod(torch.clamp(ext, min=1e-300)))
    h = (vol / max(n_elem, 1)) ** (1.0 / d)
    bins = torch.floor(x / h).to(torch.int64)
    nb = bins.max(dim=0).values + 1
    axes = torch.argsort(nb, descending=True).tolist()
    key = torch.zeros(nodes.shape[0], dtype=torch.int64, device=nodes.device)
    for a in axes:
        key = key * int(nb[a]) + bins[:, a]
    return torch.argsort(key, stable=True)


@dataclass
class SolverView:
    """A pattern in the solver's node order and the map solver DOF -> model DOF."""
    pattern: DevicePattern
    perm_dof: Tensor
