"""ctypes binding of libcxg.so (include/cxg.h).

The library is loaded from this package directory only; there is no fallback. A
missing or unloadable library raises CxgLibraryError at first use, so no caller can
silently run anything but the HIP path.

Errors from the library surface as CxgError (a RuntimeError, as the reference's
pybind11 module translates std::runtime_error, python/bindings.cpp:1-4) carrying
the library's message, e.g. the reference's "non-positive element Jacobian (bad
connectivity?)" (src/fem/element.cpp:87).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcxg.so")

# cxg_status
OK, ERR_ARG, ERR_HIP, ERR_JACOBIAN, ERR_UNSUPPORTED, ERR_STATE = range(6)
# element types (cxpp::ElementType order, core/element.hpp:13-22)
C3D4, C3D10, C3D8, C3D8R, C3D20, C3D20R, C3D6, C3D15 = range(8)
TYPE_NAMES = {"C3D4": C3D4, "C3D10": C3D10, "C3D8": C3D8, "C3D8R": C3D8R,
              "C3D20": C3D20, "C3D20R": C3D20R, "C3D6": C3D6, "C3D15": C3D15}
HARD_ISOTROPIC, HARD_KINEMATIC, HARD_COMBINED = range(3)
MAX_TABLE = 16
STATE_COMPONENTS = 13


class CxgLibraryError(RuntimeError):
    """libcxg.so is missing or cannot be loaded (build it: python -m calculixpp_amd.build)."""


class CxgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class Block(C.Structure):
    _fields_ = [("type", C.c_int32), ("n_elem", C.c_int64), ("elem_offset", C.c_int64),
                ("gp_offset", C.c_int64), ("conn", C.c_void_p)]


class J2Material(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("hardening", C.c_int32),
                ("n_table", C.c_int32), ("eqplastic", C.c_double * MAX_TABLE),
                ("yield_", C.c_double * MAX_TABLE), ("kinematic_modulus", C.c_double)]


_P = C.c_void_p
_SIGS = {
    "cxg_version": (C.c_char_p, []),
    "cxg_nodes_per_element": (C.c_int32, [C.c_int32]),
    "cxg_gauss_points": (C.c_int32, [C.c_int32]),
    "cxg_create": (C.c_int, [C.c_int, _P, C.POINTER(_P)]),
    "cxg_destroy": (None, [_P]),
    "cxg_set_stream": (C.c_int, [_P, _P]),
    "cxg_last_error": (C.c_char_p, [_P]),
    "cxg_set_option": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "cxg_mesh_create": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(Block), C.POINTER(_P)]),
    "cxg_mesh_destroy": (None, [_P]),
    "cxg_mesh_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32)]),
    "cxg_pattern_build": (C.c_int, [_P, _P, _P, C.c_int32, C.POINTER(_P)]),
    "cxg_pattern_destroy": (None, [_P]),
    "cxg_pattern_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "cxg_pattern_row_ptr": (_P, [_P]),
    "cxg_pattern_col_idx": (_P, [_P]),
    "cxg_pattern_copy": (C.c_int, [_P, _P, _P, _P]),
    "cxg_pattern_compact": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, C.POINTER(C.c_int64)]),
    "cxg_assemble_linear": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "cxg_assemble_linear_cases": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_int32, _P, _P, _P]),
    "cxg_internal_force": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "cxg_internal_force_noalloc": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "cxg_reduce_free": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P]),
    "cxg_assemble_j2": (C.c_int, [_P, _P, _P, _P, _P, _P, C.POINTER(J2Material), C.c_int32,
                                  _P, _P, _P, _P, _P, _P]),
    "cxg_commit": (C.c_int, [_P, C.c_int64, _P, _P]),
    "cxg_material_evaluate": (C.c_int, [_P, C.c_int64, _P, C.POINTER(J2Material), C.c_int32,
                                        _P, _P, _P, _P, _P, _P, _P]),
    "cxg_vjp": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "cxg_assemble_mass": (C.c_int, [_P, _P, _P, _P, C.c_int32, _P, _P]),
    "cxg_assemble_geometric": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "cxg_pattern_same_direction": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_int64)]),
    "cxg_assemble_mass_sd": (C.c_int, [_P, _P, _P, _P, C.c_int32, _P, _P]),
    "cxg_assemble_geometric_sd": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "cxg_compact_csr": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P,
                                  C.POINTER(C.c_int64)]),
    "cxg_recover_fields": (C.c_int, [_P, _P, _P, _P, _P, _P, C.POINTER(J2Material), C.c_int32,
                                     _P, _P, _P, _P, _P, _P, _P, _P]),
    "cxg_graph_capture_linear": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "cxg_graph_capture_j2": (C.c_int, [_P, _P, _P, _P, _P, _P, C.POINTER(J2Material), C.c_int32,
                                       _P, _P, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "cxg_graph_capture_linear_cases": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_int32, _P, _P,
                                                 C.POINTER(_P)]),
    "cxg_pcg": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, C.c_double, C.c_int32,
                          C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "cxg_graph_launch": (C.c_int, [_P, _P]),
    "cxg_graph_info": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "cxg_graph_destroy": (None, [_P]),
    "cxg_check": (C.c_int, [_P]),
    "cxg_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "cxg_free": (C.c_int, [_P, _P]),
    "cxg_copy_to_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cxg_copy_to_host": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "cxg_profile_enable": (C.c_int, [_P, C.c_int]),
    "cxg_profile_reset": (C.c_int, [_P]),
    "cxg_profile_count": (C.c_int32, [_P]),
    "cxg_profile_get": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded libcxg.so (loaded once). Import torch first so the library binds
    to the HIP runtime torch already loaded (same SONAME)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CxgLibraryError(
                f"{LIB_PATH} not found; build it with `python -m calculixpp_amd.build`")
        try:
            import torch  # noqa: F401  (loads the HIP runtime first)
        except ImportError:
            pass
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - depends on the host
            raise CxgLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def check(ctx, status: int) -> None:
    if status != OK:
        msg = lib().cxg_last_error(ctx)
        raise CxgError(status, msg.decode() if msg else f"cxg status {status}")
