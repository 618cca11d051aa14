"""Node-block sparsity pattern of a mesh, built once on the host by
libtfm_hip.so (tfm_pattern_create) and kept on the device.

This replaces the pattern block of FEM.__init__ (base.py:72-122): instead of
sorting the packed (row << 32 | col) keys of every element DOF pair, the host
library forms each node's sorted adjacency; the CSR values of a node's rows
are contiguous and in the reference's row-major order, so the reference's
index arrays (glob_idx, k_map, diag_map) are exact expansions of it
(`expand`).
"""

from __future__ import annotations

import ctypes as C

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


class DevicePattern:
    """Mesh geometry and pattern on the device, with the tfm_mesh_t view."""

    def __init__(self, nodes: Tensor, elements: Tensor, ndof: int, etype: str,
                 device: torch.device):
        self.device = device
        self.n_nod, self.dim = nodes.shape
        self.n_elem = elements.shape[0]
        self.ndof = ndof
        self.etype = etype
        host = build_host_pattern(elements, self.n_nod, ndof)
        self.nnz = host["nnz"]
        self.max_deg = host["max_deg"]
        self.max_inc = host["max_inc"]
        self.npe = host["npe"] if self.n_elem else elements.shape[1]
        self.nodes = nodes.detach().to(device=device, dtype=torch.float64).contiguous()
        self.elements = elements.detach().to(device=device, dtype=torch.int32).contiguous()
        up = lambda a: torch.from_numpy(a).to(device)  # noqa: E731
        self.adj_ptr = up(host["adj_ptr"])
        self.adj = up(host["adj"])
        self.val_ptr = up(host["val_ptr"])
        self.ne_ptr = up(host["ne_ptr"])
        self.ne_elem = up(host["ne_elem"])
        self.ne_loc = up(host["ne_loc"])
        # uint16 travels as int16 storage (same bits)
        self.epos = torch.from_numpy(host["epos"].view(np.int16)).to(device)
        # 16-bit neighbour offsets (adj[k] - row node) when they all fit
        rows = np.repeat(np.arange(self.n_nod, dtype=np.int64), np.diff(host["adj_ptr"]))
        delta = host["adj"].astype(np.int64) - rows
        fits = delta.size > 0 and int(np.abs(delta).max()) <= np.iinfo(np.int16).max
        self.adj16 = up(delta.astype(np.int16)) if fits else None
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
