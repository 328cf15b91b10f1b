"""ctypes binding of ``libradis_ldm.so`` (declared in ``include/radis_ldm.h``).

The library is built in-tree by ``radis_amd.build.build_library()``.  There
is no fallback: if the library is missing or fails to load, every entry
point of the package raises :class:`NativeLibraryError`.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint8, c_void_p

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libradis_ldm.so")
MAX_ISO = 16
N_STAGES = 8
STAGE_NAMES = ("axis", "finalize", "tile_bounds", "fill", "fft_a", "fft_b", "fft_c", "wrap")

# status codes (radis_ldm.h)
OK, EINVAL, EHIP, ENOMEM, ERANGE, EDEVICE = 0, -1, -2, -3, -4, -5
_CODE_NAMES = {EINVAL: "EINVAL", EHIP: "EHIP", ENOMEM: "ENOMEM", ERANGE: "ERANGE", EDEVICE: "EDEVICE"}
# device status bits (radis_ldm_request_info.status)
ST_AXIS, ST_CAP, ST_FOLD, ST_WIDTH = 1, 2, 4, 8


class NativeLibraryError(RuntimeError):
    """The HIP library is missing or could not be loaded."""


class LdmError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code, message):
        super().__init__(f"{_CODE_NAMES.get(code, code)}: {message}")
        self.code = code


class BudgetExceeded(LdmError):
    """Workspace or line list does not fit the device byte budget."""


class Lines(ctypes.Structure):
    _fields_ = [
        ("nu0", POINTER(c_double)),
        ("S_ref", POINTER(c_float)),
        ("El", POINTER(c_float)),
        ("g_air", POINTER(c_float)),
        ("g_self", POINTER(c_float)),
        ("n_air", POINTER(c_float)),
        ("n_self", POINTER(c_float)),
        ("delta", POINTER(c_float)),
        ("iso", POINTER(c_uint8)),
        ("n", c_int64),
        ("molar_mass", POINTER(c_double)),
        ("n_iso_slots", c_int32),
        ("Evib_l", POINTER(c_float)),
        ("Evib_u", POINTER(c_float)),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("wmin", c_double),
        ("wmax", c_double),
        ("wstep", c_double),
        ("Tgas", c_double),
        ("pressure_bar", c_double),
        ("mole_fraction", c_double),
        ("Tref", c_double),
        ("dxG", c_double),
        ("dxL", c_double),
        ("folding_thresh", c_double),
        ("q_ratio", c_double * MAX_ISO),
        ("fft_len", c_int64),
        ("n_v", c_int64),
        ("line_lo", c_double),
        ("line_hi", c_double),
        ("flags", c_int32),
        ("fft_n1", c_int32),
        ("tile_w", c_int32),
        ("Tvib", c_double),
        ("Trot", c_double),
    ]


FP64_WIDTHS = 1
GENERIC_FFT = 2
MIN_RMS = 4
NONEQ = 8
SCAN_AXIS = 16
DEVICE_Q = 32
PERIOD_L = 64
ABI_VERSION = 6
NOALLOC_BYTES = 256 << 20


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("n_v", c_int64),
        ("fft_len", c_int64),
        ("n1", c_int32),
        ("n2", c_int32),
        ("line_begin", c_int64),
        ("line_end", c_int64),
        ("m_cap", c_int32),
        ("tile_w", c_int32),
        ("scale_log2", c_int32),
        ("n_g", c_int32),
        ("n_l", c_int32),
        ("status", c_int32),
        ("compiled_fft", c_int32),
        ("wg_min", c_double),
        ("wg_max", c_double),
        ("wl_min", c_double),
        ("wl_max", c_double),
        ("axis_scanned", c_int64),
        ("axis_candidates", c_int64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("n_g", c_int32),
        ("n_l", c_int32),
        ("status", c_int32),
        ("pad_", c_int32),
        ("wg_min", c_double),
        ("wg_max", c_double),
        ("wl_min", c_double),
        ("wl_max", c_double),
    ]


class Derived(ctypes.Structure):
    _fields_ = [
        ("absorbance", c_void_p),
        ("transmittance", c_void_p),
        ("emissivity", c_void_p),
        ("radiance", c_void_p),
        ("path_length", c_double),
    ]


_PROTOTYPES = {
    "radis_ldm_last_error": (c_char_p, []),
    "radis_ldm_abi_version": (c_int, []),
    "radis_ldm_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "radis_ldm_ctx_destroy": (None, [c_void_p]),
    "radis_ldm_ctx_set_budget": (c_int, [c_void_p, c_int64]),
    "radis_ldm_ctx_bytes_in_use": (c_int, [c_void_p, POINTER(c_int64)]),
    "radis_ldm_db_upload": (c_int, [c_void_p, POINTER(Lines), POINTER(c_void_p)]),
    "radis_ldm_db_destroy": (None, [c_void_p]),
    "radis_ldm_db_size": (c_int, [c_void_p, POINTER(c_int64)]),
    "radis_ldm_db_set_qspline": (c_int, [c_void_p, c_int, POINTER(c_double), c_int32, POINTER(c_double), c_int32,
                                         c_double, c_double]),
    "radis_ldm_qspline_eval": (c_int, [c_void_p, c_int, c_void_p, c_int64, c_void_p, c_void_p]),
    "radis_ldm_plan_grid": (c_int, [c_double, c_double, c_double, POINTER(c_int64), POINTER(c_int64)]),
    "radis_ldm_request_bytes": (c_int, [c_void_p, POINTER(Request), POINTER(c_int64), POINTER(RequestInfo)]),
    "radis_ldm_batch_create": (c_int, [c_void_p, c_void_p, c_int, POINTER(Request), POINTER(c_void_p)]),
    "radis_ldm_batch_set_request": (c_int, [c_void_p, c_int, POINTER(Request)]),
    "radis_ldm_batch_run": (c_int, [c_void_p, POINTER(c_void_p), c_void_p]),
    "radis_ldm_batch_set_derived": (c_int, [c_void_p, c_int, POINTER(Derived)]),
    "radis_ldm_batch_graph_capture": (c_int, [c_void_p, POINTER(c_void_p), c_void_p]),
    "radis_ldm_batch_graph_replay": (c_int, [c_void_p, c_void_p]),
    "radis_ldm_batch_info": (c_int, [c_void_p, c_int, POINTER(RequestInfo), c_void_p]),
    "radis_ldm_batch_workspace_bytes": (c_int, [c_void_p, POINTER(c_int64)]),
    "radis_ldm_batch_destroy": (None, [c_void_p]),
    "radis_ldm_batch_enable_timing": (c_int, [c_void_p, c_int]),
    "radis_ldm_batch_stage_ms": (c_int, [c_void_p, POINTER(c_float)]),
    "radis_ldm_batch_line_terms": (c_int, [c_void_p, c_int] + [c_void_p] * 7 + [c_void_p]),
    "radis_ldm_batch_copy_ldm": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "radis_ldm_slit_convolve": (c_int, [c_void_p, c_int64, c_void_p, c_int32, c_double, c_void_p, c_void_p]),
    "radis_ldm_run_noalloc": (c_int, [c_void_p, c_void_p, POINTER(Request), c_void_p, c_void_p,
                                      POINTER(RequestInfo), c_void_p]),
}

EXPORTED = tuple(_PROTOTYPES)

_lib = None


def load():
    """Load the library once (after torch, so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (binds libamdhip64 first)
    except ImportError:  # pragma: no cover - torch is part of the image
        pass
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in _PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.radis_ldm_abi_version() != ABI_VERSION:
        raise NativeLibraryError(f"{LIB_PATH} has ABI {lib.radis_ldm_abi_version()}, this binding {ABI_VERSION}: "
                                 "rebuild it")
    _lib = lib
    return lib


def check(rc):
    """Raise on a negative return code, with the library's message."""
    if rc == OK:
        return
    msg = load().radis_ldm_last_error()
    msg = msg.decode() if msg else ""
    if rc == ENOMEM and "budget" in msg:
        raise BudgetExceeded(rc, msg)
    raise LdmError(rc, msg)
