"""ctypes binding of ``libcvfes_mi355x.so`` (C ABI declared in ``include/cvfes_mi355x.h``).

``torch`` is imported before the library is opened so that the library's ``libamdhip64.so.7``
dependency resolves to the HIP runtime torch already loaded (same SONAME): device pointers and
streams handed over from torch tensors are then valid inside the library.  There is no fallback:
if the library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import torch  # noqa: F401  (load torch's HIP runtime first)

LIB_PATH = Path(__file__).resolve().parent / "libcvfes_mi355x.so"
ABI_VERSION = 12

METHOD_STORED = 0
METHOD_MATFREE = 1
METHODS = {"stored": METHOD_STORED, "bcsr": METHOD_STORED, "matfree": METHOD_MATFREE,
           "ebe": METHOD_MATFREE}

VARIANT_AUTO, VARIANT_GENERIC, VARIANT_ROWLOAD, VARIANT_ROWLOAD_NT = 0, 1, 2, 3
VARIANT_ROWLOAD1, VARIANT_ROWLOAD1_NT = 4, 5
VARIANT_LDS, VARIANT_LDS2 = 6, 7
VARIANT_SYM, VARIANT_SYM_NT, VARIANT_SYM_NTL = 8, 9, 10
VARIANT_CACHED, VARIANT_CACHED1 = 11, 12
VARIANT_EBE = 13
VARIANTS = {"auto": VARIANT_AUTO, "generic": VARIANT_GENERIC, "rowload": VARIANT_ROWLOAD,
            "rowload_nt": VARIANT_ROWLOAD_NT, "rowload1": VARIANT_ROWLOAD1,
            "rowload1_nt": VARIANT_ROWLOAD1_NT, "lds": VARIANT_LDS, "lds2": VARIANT_LDS2,
            "sym": VARIANT_SYM, "sym_nt": VARIANT_SYM_NT, "sym_ntl": VARIANT_SYM_NTL,
            "cached": VARIANT_CACHED, "cached1": VARIANT_CACHED1, "ebe": VARIANT_EBE}

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_int = ctypes.c_int


class CvfesError(RuntimeError):
    """A negative return code from the native library."""


class CvfesSystem(ctypes.Structure):
    """Mirror of ``cvfes_system_t``."""

    _fields_ = [
        ("n_nodes", _i64), ("n_smp", _i64), ("nnzb", _i64), ("n_elms", _i64),
        ("indptr", _vp), ("indices", _vp), ("values", _vp),
        ("cptr", _vp), ("celem", _vp), ("khat", _vp), ("coef", _vp),
        ("lm", _vp), ("lm_sstride", _i64), ("f1", _vp), ("fixed", _vp), ("bdy_u", _f64),
        ("max_row_blocks", _i64), ("max_offdiag_contrib", _i64), ("max_diag_contrib", _i64),
        ("variant", _i64), ("nptr", _vp), ("nent", _vp), ("cslot", _vp), ("max_node_elems", _i64),
        ("uvalues", _vp), ("vblk", _vp), ("nnzu", _i64), ("row_order", _vp),
        ("color_ptr", _vp), ("color_elems", _vp), ("elems", _vp), ("cpos", _vp), ("n_colors", _i64),
        ("max_color_elems", _i64), ("waves_per_simd", _i64), ("row_records", _vp),
        ("record_width", _i64), ("minv", _vp),
    ]


class CvfesLoad(ctypes.Structure):
    """Mirror of ``cvfes_load_t``."""

    _fields_ = [("p", _vp), ("ga", _vp), ("qa", _vp), ("gb", _vp), ("qb", _vp)]


class CvfesGeometry(ctypes.Structure):
    """Mirror of ``cvfes_geometry_t``."""

    _fields_ = [("nodes", _vp), ("elems", _vp), ("nu", _f64), ("kshear", _f64)]


class CvfesUpdate(ctypes.Structure):
    """Mirror of ``cvfes_update_t``."""

    _fields_ = [("interval", _i64), ("counter", _i64), ("has_geometry", _i64), ("ug", _vp)]


class CvfesBatchMesh(ctypes.Structure):
    """Mirror of ``cvfes_batch_mesh_t``."""

    _fields_ = [("sys", CvfesSystem), ("method", _i64), ("state", _vp * 3), ("load", CvfesLoad),
                ("dt", _f64), ("beta", _f64), ("damp", _f64), ("task0", _i64), ("lanes", _i64),
                ("vw", _i64), ("packed", _i64), ("n_mtasks", _i64), ("slice", _i64), ("xtask0", _i64)]


_SIGNATURES = {
    "cvfes_abi_version": (_int, []),
    "cvfes_last_error": (ctypes.c_char_p, []),
    "cvfes_stream_sync": (_int, [_vp]),
    "cvfes_struct_size": (_i64, [_int]),
    "cvfes_stream_blocks": (_i64, [_i64]),
    "cvfes_stream_read": (_int, [_vp, _vp, _i64, _int, _vp]),
    "cvfes_stream_copy": (_int, [_vp, _vp, _i64, _int, _vp]),
    "cvfes_element_matrices": (_int, [_vp, _i64, _vp, _vp, _f64, _f64, _vp, _vp, _vp, _vp]),
    "cvfes_element_coef": (_int, [_vp, _i64, _i64, _vp, _vp, _i64, _vp, _vp, _int, _vp]),
    "cvfes_lumped_mass_load": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                      _f64, _f64, _vp, _vp]),
    "cvfes_assemble_bcsr": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "cvfes_assemble_upper": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cvfes_bcsr_spmv": (_int, [_vp, ctypes.POINTER(CvfesSystem), _int, _vp, _vp]),
    "cvfes_initial_state": (_int, [_vp, ctypes.POINTER(CvfesSystem), _int, _vp, _vp,
                                   ctypes.POINTER(CvfesLoad), _f64, _f64, _f64, _vp, _vp, _vp]),
    "cvfes_initial_state_static": (_int, [_vp, ctypes.POINTER(CvfesSystem), _int, _vp, _vp,
                                          ctypes.POINTER(CvfesLoad), _f64, _f64, _f64, _vp, _vp]),
    "cvfes_step": (_int, [_vp, ctypes.POINTER(CvfesSystem), _int, ctypes.POINTER(_vp),
                          ctypes.POINTER(CvfesLoad), _i64, _f64, _f64, _f64]),
    "cvfes_adjoint_step": (_int, [_vp, ctypes.POINTER(CvfesSystem), _int, _vp, _vp, _vp, _vp, _vp,
                                  _vp, _f64, _f64, _f64, _int]),
    "cvfes_grad_accum": (_int, [_vp, ctypes.POINTER(CvfesSystem), _vp, _vp, _vp, _vp, _vp, _f64,
                                _f64, _f64, _vp]),
    "cvfes_inverse_mass": (_int, [_vp, _i64, _vp, _vp]),
    "cvfes_step_kernel": (_i64, [ctypes.POINTER(CvfesSystem), _int]),
    "cvfes_adjoint_fused_step": (_int, [_vp, ctypes.POINTER(CvfesSystem), _int, ctypes.POINTER(_vp),
                                        ctypes.POINTER(CvfesLoad), _vp, _vp, _vp, _vp, _vp, _vp,
                                        _f64, _f64, _f64]),
    "cvfes_element_stress": (_int, [_vp, _i64, _i64, ctypes.POINTER(CvfesGeometry), _vp, _vp, _vp,
                                    _vp]),
    "cvfes_step_updated": (_int, [_vp, ctypes.POINTER(CvfesSystem), ctypes.POINTER(CvfesGeometry),
                                  ctypes.POINTER(CvfesUpdate), ctypes.POINTER(_vp),
                                  ctypes.POINTER(CvfesLoad), _i64, _f64, _f64, _f64]),
    "cvfes_batch_plan_bytes": (_i64, [_int]),
    "cvfes_batch_plan": (_int, [_vp, ctypes.POINTER(CvfesBatchMesh), _int, _vp,
                                ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "cvfes_batch_step": (_int, [_vp, _vp, _int, _i64, _int, _vp, _i64, _i64, _i64]),
    "cvfes_counter_add": (_int, [_vp, _vp, _i64]),
    "cvfes_batch_graph_create": (_int, [_vp, _int, _i64, _vp, _i64, _int, _i64, _i64,
                                        ctypes.POINTER(_vp)]),
    "cvfes_graph_launch": (_int, [_vp, _vp]),
    "cvfes_graph_destroy": (_int, [_vp]),
}

EXPORTED = tuple(_SIGNATURES)

# Doubles in the static device scratch of cvfes_initial_state_static (CVFES_STATIC_SCRATCH_DOUBLES).
STATIC_SCRATCH_DOUBLES = 8 << 20

_LIB = None


def lib(path: Path | str | None = None):
    """Open (once) and return the native library with argument types declared."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise CvfesError(f"{p} is not built: run `python -m cvfes_amd.build` (hipcc, gfx950)")
    handle = ctypes.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    if handle.cvfes_abi_version() != ABI_VERSION:
        raise CvfesError("libcvfes_mi355x.so ABI version mismatch; rebuild the library")
    if path is None:
        _LIB = handle
    return handle


def check(rc: int) -> None:
    """Raise :class:`CvfesError` with the library's message when ``rc`` is an error code."""
    if rc != 0:
        msg = lib().cvfes_last_error().decode(errors="replace")
        raise CvfesError(f"cvfes error {rc}: {msg}")


def ptr(t) -> int | None:
    """Device (or host) address of a tensor, ``None`` for ``None``."""
    return None if t is None else t.data_ptr()


def stream_handle(device=None) -> int:
    """Raw ``hipStream_t`` of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream
