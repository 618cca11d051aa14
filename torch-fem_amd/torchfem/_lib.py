"""ctypes binding of libtfm_hip.so, the gfx950 assemble-and-solve library.

The binding follows the conventions of the reference's only native binding
(src/torchfem/amgx.py:100-201): a signature table applied at load time, int
status codes turned into exceptions with the library's own message, opaque
handles released by an explicit call, and raw device pointers taken from torch
tensors. The library is loaded from the package directory (or TFM_HIP_LIB);
there is no fallback: if it cannot be loaded, every compute call raises.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Any

import torch

LIB_NAME = "libtfm_hip.so"

TFM_OK = 0
TFM_ERR_ARG = 1
TFM_ERR_HIP = 2
TFM_ERR_NOCONV = 3
TFM_ERR_JACOBIAN = 4

ETYPE = {"Quad1": 1, "Tetra1": 2, "Hexa1": 3, "Hexa2": 4}
MAT_ISO, MAT_GENERAL, MAT_HEAT_ISO, MAT_HEAT_GENERAL = 1, 2, 3, 4
PREC_FP64, PREC_MIXED = 0, 1
PC_JACOBI, PC_BLOCK_JACOBI, PC_TWO_LEVEL = 0, 1, 2
FP32_PLAIN, FP32_DIFF = 0, 1

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F64 = C.c_double


class Mesh(C.Structure):
    """tfm_mesh_t (include/tfm.h)."""

    _fields_ = [
        ("nodes", P),
        ("elements", P),
        ("n_nod", I64),
        ("n_elem", I64),
        ("dim", I32),
        ("npe", I32),
        ("ndof", I32),
        ("etype", I32),
        ("max_deg", I32),
        ("max_inc", I32),
        ("nnz", I64),
        ("adj_ptr", P),
        ("adj", P),
        ("val_ptr", P),
        ("ne_ptr", P),
        ("ne_elem", P),
        ("ne_loc", P),
        ("epos", P),
        ("adj16", P),
    ]


class Material(C.Structure):
    """tfm_material_t (include/tfm.h)."""

    _fields_ = [("kind", I32), ("_pad", I32), ("p0", P), ("p1", P), ("scale", P)]


class System(C.Structure):
    """tfm_system_t (include/tfm.h)."""

    _fields_ = [
        ("mesh", C.POINTER(Mesh)),
        ("mat", C.POINTER(Material)),
        ("con", P),
        ("val64", P),
        ("val32", P),
        ("diag_blocks", P),
        ("val32_form", I32),
        ("_pad", I32),
    ]


_SIGNATURES: dict[str, tuple[list[Any], Any]] = {
    "tfm_last_error": ([], C.c_char_p),
    "tfm_version": ([], C.c_int),
    "tfm_pattern_create": ([C.POINTER(P), P, I64, I32, I64, I32], C.c_int),
    "tfm_pattern_info": (
        [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I32), C.POINTER(I32)],
        C.c_int,
    ),
    "tfm_pattern_export": ([P, P, P, P, P, P, P, P], C.c_int),
    "tfm_pattern_destroy": ([P], C.c_int),
    "tfm_pattern_device_count": (
        [P, P, I64, I32, I64, I32, P, P, P, P, P, P, C.POINTER(I64)],
        C.c_int,
    ),
    "tfm_pattern_device_fill": (
        [P, P, I64, I32, I64, P, P, P, P, P, P, P, C.POINTER(I32)],
        C.c_int,
    ),
    "tfm_pattern_expand": ([P, C.POINTER(Mesh), P, P, P, P, P], C.c_int),
    "tfm_element_matrices": ([P, C.POINTER(Mesh), C.POINTER(Material), P, P], C.c_int),
    "tfm_element_matrices_gp": ([P, C.POINTER(Mesh), P, P, P, P], C.c_int),
    "tfm_element_mass": ([P, C.POINTER(Mesh), P, P, P, P], C.c_int),
    "tfm_assemble": (
        [P, C.POINTER(Mesh), C.POINTER(Material), P, P, P, P, P, I32],
        C.c_int,
    ),
    "tfm_assemble_elements": ([P, C.POINTER(Mesh), P, P, P, P, P, I32], C.c_int),
    "tfm_internal_force": (
        [P, C.POINTER(Mesh), C.POINTER(Material), P, P, P, P, P, P, P, P],
        C.c_int,
    ),
    "tfm_internal_force_static": (
        [P, C.POINTER(Mesh), C.POINTER(Material), P, P, P, P, P, P, P],
        C.c_int,
    ),
    "tfm_gather_nodes": ([P, C.POINTER(Mesh), P, P], C.c_int),
    "tfm_sensitivity": (
        [P, C.POINTER(Mesh), C.POINTER(Material), P, P, I32, F64, P],
        C.c_int,
    ),
    "tfm_spmv": ([P, C.POINTER(Mesh), I32, P, P, P], C.c_int),
    "tfm_csr_spmv": ([P, I64, P, P, P, P, P], C.c_int),
    "tfm_pcg_create": ([C.POINTER(P), I64, I32, I32], C.c_int),
    "tfm_pcg_destroy": ([P], C.c_int),
    "tfm_pcg_solve": (
        [
            P,
            P,
            C.POINTER(System),
            I32,
            I32,
            P,
            P,
            I32,
            F64,
            F64,
            F64,
            C.POINTER(I32),
            C.POINTER(F64),
            C.POINTER(I32),
        ],
        C.c_int,
    ),
    "tfm_pcg_timing": ([P, I32, C.POINTER(F64), C.POINTER(I64)], C.c_int),
    "tfm_pcg_fp64_fallback": ([P, C.POINTER(I32)], C.c_int),
    "tfm_coarse_create": ([C.POINTER(P), P, I32, P, P, P, P, P, P, I64, P, P], C.c_int),
    "tfm_coarse_destroy": ([P], C.c_int),
    "tfm_coarse_size": ([P, C.POINTER(I32)], C.c_int),
    "tfm_coarse_operator": ([P, P, P, P], C.c_int),
    "tfm_coarse_set_inverse": ([P, P, P], C.c_int),
    "tfm_spd_inverse_work": ([I32, C.POINTER(I64)], C.c_int),
    "tfm_spd_inverse": ([P, I32, P, P, P], C.c_int),
    "tfm_pcg_set_coarse": ([P, P], C.c_int),
    "tfm_csr_pcg_solve": (
        [P, P, I64, P, P, P, P, P, F64, C.POINTER(I32), C.POINTER(F64)],
        C.c_int,
    ),
    "tfm_krylov_create": ([C.POINTER(P), I64, I32], C.c_int),
    "tfm_krylov_destroy": ([P], C.c_int),
    "tfm_csr_minres_solve": (
        [P, P, I64, P, P, P, P, P, F64, I32, C.POINTER(I32), C.POINTER(F64)],
        C.c_int,
    ),
    "tfm_csr_gmres_solve": (
        [P, P, I64, P, P, P, P, P, F64, I32, C.POINTER(I32), C.POINTER(F64)],
        C.c_int,
    ),
    "tfm_facet_count": ([C.POINTER(Mesh), C.POINTER(I32), C.POINTER(I32)], C.c_int),
    "tfm_boundary_facets": (
        [P, C.POINTER(Mesh), P, P, P, P, P, P, I64, C.POINTER(I64)],
        C.c_int,
    ),
    "tfm_facet_nodes": ([P, C.POINTER(Mesh), P, I64, P], C.c_int),
    "tfm_facet_load": ([P, C.POINTER(Mesh), P, P, P, I64, P, I32, I32, P, P], C.c_int),
    "tfm_body_load": ([P, C.POINTER(Mesh), P, P, I32, I32, P, P, P], C.c_int),
    "tfm_shape_gradients": ([P, C.POINTER(Mesh), P, I32, P, P, P], C.c_int),
}


def library_path() -> Path:
    env = os.environ.get("TFM_HIP_LIB")
    return Path(env) if env else Path(__file__).resolve().parent / LIB_NAME


def _bind_library() -> C.CDLL:
    path = library_path()
    try:
        lib = C.CDLL(str(path))
    except OSError as err:
        raise ImportError(
            f"Could not load the HIP library {str(path)!r}. Build it with "
            "`python torch-fem_amd/build.py` (or __graft_entry__.build())."
        ) from err
    for fname, (argtypes, restype) in _SIGNATURES.items():
        func = getattr(lib, fname)
        func.argtypes = argtypes
        func.restype = restype
    return lib


_lib: C.CDLL | None = None
_load_error: Exception | None = None


def lib() -> C.CDLL:
    """The loaded library; raises ImportError if it is missing."""
    global _lib, _load_error
    if _lib is None:
        if _load_error is not None:
            raise ImportError(str(_load_error)) from _load_error
        try:
            _lib = _bind_library()
        except ImportError as err:
            _load_error = err
            raise
    return _lib


class HIPError(RuntimeError):
    """A failure reported by libtfm_hip.so."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libtfm_hip error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    """Raise with the library's own message unless `rc` is TFM_OK."""
    if rc == TFM_OK:
        return
    msg = lib().tfm_last_error()
    raise HIPError(rc, msg.decode(errors="replace") if msg else "")


def ptr(t: torch.Tensor | None) -> int | None:
    """Raw data pointer of a contiguous tensor (None for None)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor passed to libtfm_hip must be contiguous")
    return t.data_ptr()


def stream_of(device: torch.device) -> int:
    """The current HIP stream of `device` as a raw handle."""
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu(device: torch.device | str | None = None) -> torch.device:
    """The ROCm device the path runs on; raises if there is none.

    The assemble-and-solve path has no CPU implementation: without a GPU and
    the HIP library every call fails loudly instead of falling back.
    """
    lib()
    if not torch.cuda.is_available():
        raise RuntimeError(
            "torchfem (MI355X build) needs a ROCm GPU: the assemble-and-solve path "
            "runs only as HIP kernels and has no CPU fallback."
        )
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    device = torch.device(device)
    if device.type != "cuda":
        return torch.device("cuda", torch.cuda.current_device())
    if device.index is None:
        return torch.device("cuda", torch.cuda.current_device())
    return device
