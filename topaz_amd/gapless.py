"""Batched gapless HSP extension on the GPU (TOPAZ's sa_hsp, search.c:99-190).

Host-side mirror of the sa_hsp calls made by query_find_suffixes
(query.c:117,125) and suffix_increment (suffix.c:47), for many seeds at once.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P32, P64, check, ptr

SEED_DTYPE = np.dtype([("query", "<i4"), ("offset", "<i4"), ("position", "<i8")])


def _declare(L):
    if getattr(L, "_gapless_declared", False):
        return
    VP = C.c_void_p
    L.tpz_text_create.argtypes = [C.POINTER(VP), C.c_char_p, C.c_uint64, P64]
    L.tpz_text_create.restype = C.c_int
    L.tpz_text_destroy.argtypes = [VP]
    L.tpz_text_destroy.restype = None
    L.tpz_gapless_batch_create.argtypes = [C.POINTER(VP), VP, C.c_int32, C.c_char_p, P64, P32,
                                           C.c_int64, VP, C.POINTER(C.c_int8)]
    L.tpz_gapless_batch_create.restype = C.c_int
    L.tpz_gapless_batch_run.argtypes = [VP]
    L.tpz_gapless_batch_run.restype = C.c_int
    L.tpz_gapless_batch_results.argtypes = [VP, P32, P32]
    L.tpz_gapless_batch_results.restype = C.c_int
    L.tpz_gapless_batch_kernel_ms.argtypes = [VP, C.POINTER(C.c_double)]
    L.tpz_gapless_batch_kernel_ms.restype = C.c_int
    L.tpz_gapless_batch_destroy.argtypes = [VP]
    L.tpz_gapless_batch_destroy.restype = None
    L._gapless_declared = True


class Text:
    """Device-resident '\\n'-separated protein text, optionally with its
    suffix array (int64, one entry per text byte)."""

    def __init__(self, text, sa=None):
        L = _lib.lib()
        _declare(L)
        self._text = bytes(text)
        self._sa = None if sa is None else np.ascontiguousarray(sa, np.int64)
        h = C.c_void_p()
        check(L.tpz_text_create(C.byref(h), self._text, len(self._text),
                                None if self._sa is None else ptr(self._sa, P64)),
              "tpz_text_create")
        self._h = h
        self.size = len(self._text)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tpz_text_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GaplessBatch:
    """Seeds (SEED_DTYPE records) of queries (bytes, internal alphabet)
    against a Text."""

    def __init__(self, text, queries, seeds, matrix):
        L = _lib.lib()
        _declare(L)
        qs = [bytes(q) for q in queries]
        ql = np.array([len(q) for q in qs], np.int32)
        qo = np.zeros(len(qs), np.int64)
        if len(qs):
            qo[1:] = np.cumsum(ql[:-1])
        qc = b"".join(qs)
        self._seeds = np.ascontiguousarray(seeds, dtype=SEED_DTYPE)
        mat = np.ascontiguousarray(matrix, np.int8).reshape(-1)
        self.n = len(self._seeds)
        self._text = text
        h = C.c_void_p()
        check(L.tpz_gapless_batch_create(C.byref(h), text._h, len(qs), qc, ptr(qo, P64),
                                         ptr(ql, P32), self.n, self._seeds.ctypes.data,
                                         mat.ctypes.data_as(C.POINTER(C.c_int8))),
              "tpz_gapless_batch_create")
        self._h = h

    def run(self):
        check(_lib.lib().tpz_gapless_batch_run(self._h), "tpz_gapless_batch_run")
        return self

    def results(self):
        sc = np.zeros(self.n, np.int32)
        ln = np.zeros(self.n, np.int32)
        check(_lib.lib().tpz_gapless_batch_results(self._h, ptr(sc, P32), ptr(ln, P32)),
              "tpz_gapless_batch_results")
        return sc, ln

    def kernel_ms(self):
        a = C.c_double()
        check(_lib.lib().tpz_gapless_batch_kernel_ms(self._h, C.byref(a)))
        return a.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tpz_gapless_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
