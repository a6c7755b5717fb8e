"""Batched traceback statistics of reported hits on the GPU (TOPAZ's
ssw_align_stats, ssw.c:900-978, as called by stats_using_ssw_profile,
ssw_wrapper.c:132-155)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P8, P32, P64, check, ptr
from .sw import _concat


def _declare(L):
    if getattr(L, "_stats_declared", False):
        return
    VP = C.c_void_p
    L.tpz_stats_batch_create.argtypes = [C.POINTER(VP), C.c_int32, P8, P64, P32, C.c_int32, P8,
                                         P64, P32, C.c_int32, P32, P32, P32, P32, P32, P8,
                                         C.c_int32, C.c_int32]
    L.tpz_stats_batch_create.restype = C.c_int
    L.tpz_stats_batch_run.argtypes = [VP]
    L.tpz_stats_batch_run.restype = C.c_int
    L.tpz_stats_batch_results.argtypes = [VP, P32, P32, P32, P32, P32, P32,
                                          C.POINTER(C.c_double)]
    L.tpz_stats_batch_results.restype = C.c_int
    L.tpz_stats_batch_destroy.argtypes = [VP]
    L.tpz_stats_batch_destroy.restype = None
    L._stats_declared = True


class StatsBatch:
    """One batch of hits: inputs copied at creation, run() computes the
    statistics (and may be repeated), results() returns them per hit as
    rc, read_begin1, ref_begin1, length, gapopen, mismatch, identity."""

    def __init__(self, queries, subjects, hit_query, hit_subject, read_end, ref_end, score,
                 matrix, gap_open=11, gap_extend=1):
        L = self.L = _lib.lib()
        _declare(L)
        q = queries if isinstance(queries, tuple) else _concat(queries)
        s = subjects if isinstance(subjects, tuple) else _concat(subjects)
        qr, qo, ql = [np.ascontiguousarray(x) for x in q]
        sr, so, sl = [np.ascontiguousarray(x) for x in s]
        hq = np.ascontiguousarray(hit_query, np.int32)
        hs = np.ascontiguousarray(hit_subject, np.int32)
        re_ = np.ascontiguousarray(read_end, np.int32)
        fe = np.ascontiguousarray(ref_end, np.int32)
        sc = np.ascontiguousarray(score, np.int32)
        mat = np.ascontiguousarray(matrix, np.int8).reshape(-1)
        self.n = len(hq)
        self.h = C.c_void_p()
        check(L.tpz_stats_batch_create(C.byref(self.h), len(ql), ptr(qr.astype(np.int8), P8),
                                       ptr(qo.astype(np.int64), P64), ptr(ql.astype(np.int32), P32),
                                       len(sl), ptr(sr.astype(np.int8), P8),
                                       ptr(so.astype(np.int64), P64),
                                       ptr(sl.astype(np.int32), P32), self.n, ptr(hq, P32),
                                       ptr(hs, P32), ptr(re_, P32), ptr(fe, P32), ptr(sc, P32),
                                       ptr(mat, P8), gap_open, gap_extend),
              "tpz_stats_batch_create")

    def run(self):
        check(self.L.tpz_stats_batch_run(self.h), "tpz_stats_batch_run")
        return self

    def results(self):
        out = [np.zeros(self.n, np.int32) for _ in range(6)]
        ident = np.zeros(self.n, np.float64)
        check(self.L.tpz_stats_batch_results(self.h, *[ptr(x, P32) for x in out],
                                             ident.ctypes.data_as(C.POINTER(C.c_double))),
              "tpz_stats_batch_results")
        return tuple(out) + (ident,)

    def close(self):
        if self.h:
            self.L.tpz_stats_batch_destroy(self.h)
            self.h = C.c_void_p()


def align_stats(queries, subjects, hit_query, hit_subject, read_end, ref_end, score,
                matrix, gap_open=11, gap_extend=1):
    """Per hit: rc, read_begin1, ref_begin1, length, gapopen, mismatch, identity."""
    b = StatsBatch(queries, subjects, hit_query, hit_subject, read_end, ref_end, score, matrix,
                   gap_open, gap_extend)
    try:
        return b.run().results()
    finally:
        b.close()
