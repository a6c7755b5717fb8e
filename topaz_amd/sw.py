"""Batched striped Smith-Waterman on the GPU (forward pass: score + ends).

Host-side mirror of the per-pair call sequence
    get_ssw_profile -> align_using_ssw_profile -> ssw_align
(ssw_wrapper.c:105-130, 157-317; ssw.c:835-897) for many pairs at once.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P8, P32, P64, check, ptr

SCHED_PLAIN = 0
SCHED_PERSISTENT = 1
OVERFLOW = 32767      # ssw_align returns NULL; align_using_ssw_profile stores 32767


class SwConfig(C.Structure):
    """tpz_sw_config: launch configuration of a batch (results never depend
    on it)."""
    _fields_ = [("rows_per_chunk", C.c_int32), ("groups_per_cu", C.c_int32),
                ("queries_per_task", C.c_int32), ("use_graph", C.c_int32),
                ("big_fraction", C.c_double), ("waves_per_group", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _concat(seqs):
    lens = np.array([len(s) for s in seqs], dtype=np.int32)
    off = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs):
        off[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    flat = np.concatenate([np.asarray(s, dtype=np.int8) for s in seqs]) if len(seqs) \
        else np.zeros(0, np.int8)
    return np.ascontiguousarray(flat, dtype=np.int8), off, lens


class SwBatch:
    """A device-resident batch of (query, subject) pairs.

    queries/subjects: lists of int8 arrays of matrix indices (0..24), or the
    flat (residues, offsets, lengths) triples.  pairs: (pair_query,
    pair_subject) int32 arrays.
    """

    def __init__(self, queries, subjects, pair_query, pair_subject, matrix,
                 gap_open=11, gap_extend=1):
        L = _lib.lib()
        q = queries if isinstance(queries, tuple) else _concat(queries)
        s = subjects if isinstance(subjects, tuple) else _concat(subjects)
        self._keep = [np.ascontiguousarray(x) for x in (*q, *s)]
        qr, qo, ql = q
        sr, so, sl = s
        pq = np.ascontiguousarray(pair_query, dtype=np.int32)
        ps = np.ascontiguousarray(pair_subject, dtype=np.int32)
        mat = np.ascontiguousarray(matrix, dtype=np.int8).reshape(-1)
        if mat.size != 625:
            raise ValueError("matrix must have 625 entries")
        self.n_pairs = len(pq)
        h = C.c_void_p()
        check(L.tpz_sw_batch_create(
            C.byref(h), len(ql), ptr(np.ascontiguousarray(qr, np.int8), P8),
            ptr(np.ascontiguousarray(qo, np.int64), P64),
            ptr(np.ascontiguousarray(ql, np.int32), P32),
            len(sl), ptr(np.ascontiguousarray(sr, np.int8), P8),
            ptr(np.ascontiguousarray(so, np.int64), P64),
            ptr(np.ascontiguousarray(sl, np.int32), P32),
            self.n_pairs, ptr(pq, P32), ptr(ps, P32), ptr(mat, P8),
            gap_open, gap_extend), "tpz_sw_batch_create")
        self._h = h

    def run(self, schedule=SCHED_PERSISTENT):
        check(_lib.lib().tpz_sw_batch_run(self._h, schedule), "tpz_sw_batch_run")
        return self

    def results(self):
        n = self.n_pairs
        sc = np.zeros(n, np.int32)
        re = np.zeros(n, np.int32)
        rd = np.zeros(n, np.int32)
        check(_lib.lib().tpz_sw_batch_results(self._h, ptr(sc, P32), ptr(re, P32),
                                              ptr(rd, P32)), "tpz_sw_batch_results")
        return sc, re, rd

    def kernel_ms(self):
        a = C.c_double()
        check(_lib.lib().tpz_sw_batch_kernel_ms(self._h, C.byref(a)))
        return a.value

    def config(self):
        c = SwConfig()
        check(_lib.lib().tpz_sw_batch_get_config(self._h, C.byref(c)), "tpz_sw_batch_get_config")
        return c

    def set_config(self, **kw):
        """Change fields of the configuration (rows_per_chunk, groups_per_cu,
        queries_per_task, use_graph, big_fraction) and re-plan the batch."""
        c = self.config()
        for k, v in kw.items():
            if k not in dict(SwConfig._fields_):
                raise KeyError(k)
            setattr(c, k, v)
        check(_lib.lib().tpz_sw_batch_set_config(self._h, C.byref(c)), "tpz_sw_batch_set_config")
        return self

    def autotune(self):
        """Time every candidate configuration on this batch and keep the
        fastest; returns (config, kernel ms, candidates tried)."""
        c, ms, n = SwConfig(), C.c_double(), C.c_int32()
        check(_lib.lib().tpz_sw_batch_autotune(self._h, C.byref(c), C.byref(ms), C.byref(n)),
              "tpz_sw_batch_autotune")
        return c, ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tpz_sw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def align_batch(queries, subjects, pair_query, pair_subject, matrix,
                gap_open=11, gap_extend=1, schedule=SCHED_PERSISTENT):
    """One-shot helper: (score, ref_end, read_end) arrays for every pair."""
    b = SwBatch(queries, subjects, pair_query, pair_subject, matrix, gap_open,
                gap_extend)
    try:
        return b.run(schedule).results()
    finally:
        b.close()
