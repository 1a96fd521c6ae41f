"""ctypes binding of ``libspeckle_hip.so`` (C ABI in ``include/speckle_hip.h``).

Loading fails loudly: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_size_t, c_uint32, c_void_p

LIB_NAME = "libspeckle_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# algorithm and semantics codes (speckle_hip.h)
HIO, ER, RAAR, SF, DM, NOOP = 0, 1, 2, 3, 4, 15
ALG_CODES = {"hio": HIO, "er": ER, "raar": RAAR, "sf": SF, "dm": DM}
SEM_NUMPY, SEM_OPENCL = 0, 1
SEM_CODES = {"numpy": SEM_NUMPY, "opencl": SEM_OPENCL}
# spk_plan_bytes option bits
PB_DM, PB_SHRINKWRAP, PB_IPSF, PB_SPECTRUM, PB_TRACE = 1, 2, 4, 8, 16

ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ESTATE", -5: "EUNSUP"}


class SpkStep(ctypes.Structure):
    _fields_ = [("alg", c_int32), ("semantics", c_int32), ("beta", c_float),
                ("sw_sigma", c_float), ("sw_tau", c_float), ("pad_", c_int32)]


class SpkTune(ctypes.Structure):
    _fields_ = [("tile", c_int32), ("ept", c_int32), ("persist", c_int32),
                ("sw_separate", c_int32), ("resident", c_int32), ("exchange", c_int32),
                ("barrier_polls", c_int32), ("reserved", c_int32)]


class SpkError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code, message):
        self.code = code
        super().__init__(f"speckle_hip {ERRORS.get(code, code)}: {message}")


# name -> (restype, argtypes); every symbol declared in include/speckle_hip.h
SIGNATURES = {
    "spk_ctx_create": (c_int, [c_int, c_void_p, POINTER(c_void_p)]),
    "spk_ctx_destroy": (c_int, [c_void_p]),
    "spk_ctx_set_stream": (c_int, [c_void_p, c_void_p]),
    "spk_sync": (c_int, [c_void_p]),
    "spk_last_error": (c_char_p, []),
    "spk_version": (c_char_p, []),
    "spk_supported_sizes": (c_uint32, []),
    "spk_plan_create": (c_int, [c_void_p, c_int, c_int, c_int, POINTER(SpkTune), POINTER(c_void_p)]),
    "spk_plan_destroy": (c_int, [c_void_p]),
    "spk_plan_check": (c_int, [c_void_p]),
    "spk_plan_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "spk_plan_sw_counts": (c_int, [c_void_p, POINTER(ctypes.c_longlong),
                                   POINTER(ctypes.c_longlong)]),
    "spk_set_ipsf": (c_int, [c_void_p, c_void_p]),
    "spk_set_spectrum": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "spk_plan_bytes": (c_int, [c_int, c_int, c_int, c_int, POINTER(c_size_t)]),
    "spk_set_modulus": (c_int, [c_void_p, c_int, c_void_p]),
    "spk_set_support": (c_int, [c_void_p, c_int, c_void_p]),
    "spk_get_support": (c_int, [c_void_p, c_int, c_void_p]),
    "spk_set_psi": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "spk_get_psi": (c_int, [c_void_p, c_int, c_void_p]),
    "spk_iterate": (c_int, [c_void_p, POINTER(SpkStep), c_int, c_int, c_void_p]),
    "spk_shrinkwrap": (c_int, [c_void_p, c_int, c_float, c_float]),
    "spk_copy_roi": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int]),
    "spk_fft2": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "spk_align_global_phase": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int]),
    "spk_align_global_phase_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int]),
    "spk_copy_bytes": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "spk_stack_mean": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "spk_com_average": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "spk_blur2d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_float]),
    "spk_prtf": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "spk_rftf": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "spk_rftf_stack": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                               c_void_p]),
    "spk_coherence_rl": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_int,
                                 c_void_p]),
    "spk_coherence_gauss_prepare": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "spk_coherence_gauss_eval": (c_int, [c_void_p, c_double, c_double, POINTER(c_double)]),
    "spk_time_passes": (c_int, [c_void_p, c_int, c_int, POINTER(SpkStep), POINTER(c_float)]),
    "spk_plan_profile": (c_int, [c_void_p, c_int]),
    "spk_graph_begin": (c_int, [c_void_p]),
    "spk_graph_end": (c_int, [c_void_p, POINTER(c_void_p)]),
    "spk_graph_launch": (c_int, [c_void_p]),
    "spk_graph_destroy": (c_int, [c_void_p]),
    "spk_plan_profile_read": (c_int, [c_void_p, c_int, POINTER(c_int), POINTER(c_double)]),
    "spk_plan_resident_counts": (c_int, [c_void_p, POINTER(ctypes.c_longlong),
                                         POINTER(ctypes.c_longlong)]),
    "spk_group_shape": (c_int, [c_int, POINTER(c_int), POINTER(c_int)]),
    "spk_group_create": (c_int, [c_void_p, POINTER(c_void_p)]),
    "spk_group_destroy": (c_int, [c_void_p]),
    "spk_group_iterate": (c_int, [c_void_p, POINTER(c_void_p), c_int, POINTER(c_void_p), c_int,
                                  POINTER(c_int), POINTER(c_void_p)]),
    "spk_group_counts": (c_int, [c_void_p, POINTER(ctypes.c_longlong),
                                 POINTER(ctypes.c_longlong)]),
}

_lib = None


def load(path: str | None = None):
    """Load the library once; raise if it is missing or incomplete."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{LIB_NAME} is not built at {p}; build it with "
            "`python -m als_speckle_amd.csrc.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if an export is missing
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().spk_last_error()
        raise SpkError(rc, msg.decode() if msg else "")


def last_error() -> str:
    return load().spk_last_error().decode()


# numpy twin of SpkStep: step tables built with array operations (serving
# builds one per launch for every slot)
STEP_DTYPE = np.dtype([("alg", "<i4"), ("semantics", "<i4"), ("beta", "<f4"),
                       ("sw_sigma", "<f4"), ("sw_tau", "<f4"), ("pad_", "<i4")])
assert STEP_DTYPE.itemsize == ctypes.sizeof(SpkStep)


def step_table(entries) -> np.ndarray:
    """Contiguous STEP_DTYPE array from (alg, sem, beta, sw_sigma, sw_tau)
    tuples or from a STEP_DTYPE array (returned as is when contiguous)."""
    if isinstance(entries, np.ndarray) and entries.dtype == STEP_DTYPE:
        return np.ascontiguousarray(entries).reshape(-1)
    entries = list(entries)
    arr = np.zeros(len(entries), STEP_DTYPE)
    if entries:
        cols = list(zip(*entries))
        arr["alg"], arr["semantics"] = cols[0], cols[1]
        arr["beta"], arr["sw_sigma"], arr["sw_tau"] = cols[2], cols[3], cols[4]
    return arr


def make_steps(entries) -> "ctypes.Array":
    """entries: iterable of (alg_code, sem_code, beta, sw_sigma, sw_tau), or a
    STEP_DTYPE array."""
    tab = step_table(entries)
    arr = (SpkStep * max(1, tab.size))()
    if tab.size:
        ctypes.memmove(arr, tab.ctypes.data, tab.nbytes)
    return arr
