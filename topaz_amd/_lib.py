"""Loader for libtopaz_amd.so, the HIP implementation behind the C ABI of
include/topaz_amd.h.  There is no fallback: if the library or a HIP device is
missing, every call raises."""
import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtopaz_amd.so"

_lib = None


class TopazError(RuntimeError):
    pass


def lib():
    """The loaded library (ctypes.CDLL); raises TopazError if it is absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise TopazError("%s is not built (run python -m topaz_amd.build)" % LIB_PATH)
        L = C.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


P8 = C.POINTER(C.c_int8)
PU8 = C.POINTER(C.c_uint8)
P32 = C.POINTER(C.c_int32)
P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
VP = C.c_void_p


def _declare(L):
    L.tpz_init.argtypes = [C.c_int]
    L.tpz_init.restype = C.c_int
    L.tpz_last_error.restype = C.c_char_p
    L.tpz_version.restype = C.c_char_p
    L.tpz_sw_batch_create.argtypes = [C.POINTER(VP), C.c_int32, P8, P64, P32,
                                      C.c_int32, P8, P64, P32, C.c_int32, P32,
                                      P32, P8, C.c_int32, C.c_int32]
    L.tpz_sw_batch_create.restype = C.c_int
    L.tpz_sw_batch_run.argtypes = [VP, C.c_int32]
    L.tpz_sw_batch_run.restype = C.c_int
    L.tpz_sw_batch_results.argtypes = [VP, P32, P32, P32]
    L.tpz_sw_batch_results.restype = C.c_int
    L.tpz_sw_batch_kernel_ms.argtypes = [VP, PD]
    L.tpz_sw_batch_kernel_ms.restype = C.c_int
    L.tpz_sw_batch_destroy.argtypes = [VP]
    L.tpz_sw_batch_destroy.restype = None
    from .sw import SwConfig
    L.tpz_sw_batch_get_config.argtypes = [VP, C.POINTER(SwConfig)]
    L.tpz_sw_batch_get_config.restype = C.c_int
    L.tpz_sw_batch_set_config.argtypes = [VP, C.POINTER(SwConfig)]
    L.tpz_sw_batch_set_config.restype = C.c_int
    L.tpz_sw_batch_autotune.argtypes = [VP, C.POINTER(SwConfig), PD, C.POINTER(C.c_int32)]
    L.tpz_sw_batch_autotune.restype = C.c_int


def check(rc, what="topaz_amd"):
    if rc != 0:
        msg = lib().tpz_last_error().decode(errors="replace")
        raise TopazError("%s: %s" % (what, msg))


def ptr(a, typ):
    return a.ctypes.data_as(typ)
