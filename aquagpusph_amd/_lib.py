"""ctypes binding of libaquagpusph_amd.so (the C ABI in include/aquagpusph_amd.h).

The product path has no CPU fallback: if the HIP library is missing this
module raises at import of any entry point.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libaquagpusph_amd.so")
HEADER = os.path.join(ROOT, "include", "aquagpusph_amd.h")

_LIB = None


class AquaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("aquagpusph_amd error %d: %s" % (code, msg))
        self.code = code


class Defs(ctypes.Structure):
    """aq_defs: preset definitions passed by value."""
    _fields_ = [("H", ctypes.c_float), ("CONW", ctypes.c_float),
                ("CONF", ctypes.c_float), ("DIMS", ctypes.c_float),
                ("lap_formulation", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libaquagpusph_amd.so is not built (%s); run "
                "`python -m aquagpusph_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.aq_last_error.restype = ctypes.c_char_p
        L.aq_strerror.restype = ctypes.c_char_p
        L.aq_version.restype = ctypes.c_char_p
        for name in ("aq_linklist_workspace", "aq_radix_sort_workspace",
                     "aq_reduce_workspace"):
            getattr(L, name).restype = ctypes.c_size_t
        L.aq_radix_key_bits.restype = ctypes.c_uint
        L.aq_radix_key_bits.argtypes = [ctypes.c_uint64]
        L.aq_permute.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_size_t, ctypes.c_void_p,
                                 ctypes.c_uint32, ctypes.c_void_p]
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise AquaError(rc, lib().aq_last_error().decode())
    return rc


def call(name, *args):
    """Call an int-returning entry point and raise on failure."""
    return check(getattr(lib(), name)(*args))


def header_symbols():
    """Names of every entry point declared in include/aquagpusph_amd.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"AQ_API\s+[\w\s\*]+?\b(aq_\w+)\s*\(", txt)))
