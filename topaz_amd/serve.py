"""Serving loop for stage S: continuous batching over a resident database.

Requests (queries plus their candidate subjects) are submitted from any
thread and answered as their launch finishes; requests that arrive while
others run are merged into the next launch, within a device-memory budget.
The per-pair results are those of SwBatch (ssw_align, ssw.c:835-897) whatever
else shares the launch.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P8, P32, P64, check, ptr
from .sw import _concat


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("max_inflight", C.c_int32),
                ("max_launch_pairs", C.c_int32), ("waves_per_group", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("launches", C.c_uint64),
                ("max_requests_per_launch", C.c_uint64), ("pairs_done", C.c_uint64),
                ("peak_bytes", C.c_uint64), ("workspace_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _declare(L):
    if getattr(L, "_server_declared", False):
        return
    VP = C.c_void_p
    L.tpz_server_create.argtypes = [C.POINTER(VP), C.c_int32, P8, P64, P32, P8, C.c_int32,
                                    C.c_int32, C.POINTER(ServerConfig)]
    L.tpz_server_create.restype = C.c_int
    L.tpz_server_submit.argtypes = [VP, C.c_int32, P8, P64, P32, C.c_int32, P32, P32,
                                    C.POINTER(C.c_uint64)]
    L.tpz_server_submit.restype = C.c_int
    L.tpz_server_wait.argtypes = [VP, C.c_uint64, P32, P32, P32, C.POINTER(C.c_double)]
    L.tpz_server_wait.restype = C.c_int
    L.tpz_server_get_stats.argtypes = [VP, C.POINTER(ServerStats)]
    L.tpz_server_get_stats.restype = C.c_int
    L.tpz_server_destroy.argtypes = [VP]
    L.tpz_server_destroy.restype = None
    L._server_declared = True


class Server:
    """subjects: list of int8 residue arrays or a (residues, offsets,
    lengths) triple, kept resident on the device."""

    def __init__(self, subjects, matrix, gap_open=11, gap_extend=1, memory_budget=0,
                 max_inflight=0, max_launch_pairs=0, waves_per_group=0):
        L = _lib.lib()
        _declare(L)
        s = subjects if isinstance(subjects, tuple) else _concat(subjects)
        sr, so, sl = (np.ascontiguousarray(s[0], np.int8), np.ascontiguousarray(s[1], np.int64),
                      np.ascontiguousarray(s[2], np.int32))
        mat = np.ascontiguousarray(matrix, np.int8).reshape(-1)
        cfg = ServerConfig(memory_budget, max_inflight, max_launch_pairs, waves_per_group)
        h = C.c_void_p()
        check(L.tpz_server_create(C.byref(h), len(sl), ptr(sr, P8), ptr(so, P64), ptr(sl, P32),
                                  ptr(mat, P8), gap_open, gap_extend, C.byref(cfg)),
              "tpz_server_create")
        self._h = h
        self._n = {}

    def submit(self, queries, pair_query, pair_subject):
        """Queue one request; returns its ticket at once."""
        q = queries if isinstance(queries, tuple) else _concat(queries)
        qr, qo, ql = (np.ascontiguousarray(q[0], np.int8), np.ascontiguousarray(q[1], np.int64),
                      np.ascontiguousarray(q[2], np.int32))
        pq = np.ascontiguousarray(pair_query, np.int32)
        ps = np.ascontiguousarray(pair_subject, np.int32)
        t = C.c_uint64()
        check(_lib.lib().tpz_server_submit(self._h, len(ql), ptr(qr, P8), ptr(qo, P64),
                                           ptr(ql, P32), len(pq), ptr(pq, P32), ptr(ps, P32),
                                           C.byref(t)), "tpz_server_submit")
        self._n[t.value] = len(pq)
        return t.value

    def wait(self, ticket):
        """(score, ref_end, read_end, latency in ms) of a submitted request."""
        n = self._n.pop(ticket)
        sc, re, rd = (np.zeros(n, np.int32) for _ in range(3))
        lat = C.c_double()
        check(_lib.lib().tpz_server_wait(self._h, ticket, ptr(sc, P32), ptr(re, P32),
                                         ptr(rd, P32), C.byref(lat)), "tpz_server_wait")
        return sc, re, rd, lat.value

    def stats(self):
        st = ServerStats()
        check(_lib.lib().tpz_server_get_stats(self._h, C.byref(st)), "tpz_server_get_stats")
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tpz_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
