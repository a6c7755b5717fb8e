"""The whole search of a batch of queries on the GPU (tpz_search): TOPAZ's
query_find_suffixes -> query_fetch_alignment_batch -> query_perform_alignments
-> query_perform_statistics -> output_top_hits (query.c:94-299,
engine.c:193-248), for many queries per stage.

A Database holds TOPAZ's index (the .psq text, the .SA suffix array and the
.ind protein table, as db_load maps them, db.c:36-131) resident on the
device; search() returns, per query, the hits TOPAZ prints, in printed
order.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import check

SCHED_PLAIN = 0
SCHED_PERSISTENT = 1

IND_DTYPE = np.dtype([("header_start", "<u8"), ("protein_start", "<u8"), ("header_len", "<i4"),
                      ("protein_len", "<i4")])        # position_t, db.h:10-17


class SearchOptions(C.Structure):
    _fields_ = [("num_hits", C.c_int32), ("evalue", C.c_double), ("lambda_", C.c_double),
                ("k", C.c_double), ("H", C.c_double), ("db_size", C.c_uint64),
                ("db_proteins", C.c_uint64), ("min_hit_length", C.c_int32),
                ("number_of_alignments", C.c_int32), ("number_of_seeds", C.c_int32),
                ("superfast", C.c_int32), ("first_chunk", C.c_int32),
                ("max_stream_elements", C.c_int64)]


class Hit(C.Structure):
    _fields_ = [("db_index", C.c_uint64), ("gapless_score", C.c_int32), ("rawscore", C.c_int32),
                ("qstart", C.c_int32), ("qend", C.c_int32), ("sstart", C.c_int32),
                ("send", C.c_int32), ("length", C.c_int32), ("mismatch", C.c_int32),
                ("gapopen", C.c_int32), ("identity", C.c_double), ("evalue", C.c_double),
                ("bitscore", C.c_double)]


class SearchStats(C.Structure):
    _fields_ = [("offsets", C.c_int64), ("streams", C.c_int64), ("elements", C.c_int64),
                ("rounds", C.c_int64), ("pops", C.c_int64), ("runs", C.c_int64),
                ("refills", C.c_int64), ("hits", C.c_int64), ("pairs", C.c_int64),
                ("cells", C.c_int64), ("gpu_s", C.c_double), ("seed_s", C.c_double),
                ("walk_s", C.c_double), ("align_s", C.c_double), ("stats_s", C.c_double),
                ("total_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


HIT_FIELDS = [f for f, _ in Hit._fields_]


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("lanes", C.c_int32),
                ("max_wave_queries", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("waves", C.c_uint64),
                ("max_requests_per_wave", C.c_uint64), ("queries_done", C.c_uint64),
                ("peak_bytes", C.c_uint64), ("fixed_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _declare(L):
    if getattr(L, "_search_declared", False):
        return
    VP = C.c_void_p
    L.tpz_search_server_create.argtypes = [C.POINTER(VP), VP, C.POINTER(SearchOptions),
                                           C.POINTER(ServerConfig)]
    L.tpz_search_server_create.restype = C.c_int
    L.tpz_search_server_submit.argtypes = [VP, C.c_int32, VP, VP, VP, C.POINTER(C.c_uint64)]
    L.tpz_search_server_submit.restype = C.c_int
    L.tpz_search_server_wait.argtypes = [VP, C.c_uint64, C.POINTER(VP), C.POINTER(C.c_double)]
    L.tpz_search_server_wait.restype = C.c_int
    L.tpz_search_server_get_stats.argtypes = [VP, C.POINTER(ServerStats)]
    L.tpz_search_server_get_stats.restype = C.c_int
    L.tpz_search_server_destroy.argtypes = [VP]
    L.tpz_search_server_destroy.restype = None
    L.tpz_db_create.argtypes = [C.POINTER(VP), VP, C.c_uint64, VP, C.c_uint64, VP, VP, VP,
                                C.c_int32, C.c_int32]
    L.tpz_db_create.restype = C.c_int
    L.tpz_db_destroy.argtypes = [VP]
    L.tpz_db_destroy.restype = None
    L.tpz_search_defaults.argtypes = [C.POINTER(SearchOptions)]
    L.tpz_search_defaults.restype = None
    L.tpz_search.argtypes = [VP, C.POINTER(SearchOptions), C.c_int32, VP, VP, VP, C.c_int32,
                             C.POINTER(VP)]
    L.tpz_search.restype = C.c_int
    L.tpz_search_result_count.argtypes = [VP, C.c_int32]
    L.tpz_search_result_count.restype = C.c_int32
    L.tpz_search_result_hits.argtypes = [VP, C.c_int32]
    L.tpz_search_result_hits.restype = C.POINTER(Hit)
    L.tpz_search_result_stats.argtypes = [VP, C.POINTER(SearchStats)]
    L.tpz_search_result_stats.restype = C.c_int
    L.tpz_search_result_destroy.argtypes = [VP]
    L.tpz_search_result_destroy.restype = None
    L._search_declared = True


def options(**kw):
    """TOPAZ's defaults (options.h:6-23) with the given fields changed."""
    L = _lib.lib()
    _declare(L)
    o = SearchOptions()
    L.tpz_search_defaults(C.byref(o))
    for k, v in kw.items():
        setattr(o, "lambda_" if k == "lambda" else k, v)
    return o


def load_index(prefix):
    """(text bytes, suffix array, protein table) of a TOPAZ index, the files
    memory-mapped (.psq, .SA, .ind)."""
    p = str(prefix)
    text = np.memmap(p + ".psq", np.uint8, "r")
    sa = np.memmap(p + ".SA", np.int64, "r")
    ind = np.fromfile(p + ".ind", IND_DTYPE)
    if len(sa) != len(text):
        raise ValueError("%s.SA holds %d entries for %d text bytes" % (p, len(sa), len(text)))
    return text, sa, ind


class Database:
    """A TOPAZ index resident on the device, with its scoring (matrix and
    gaps).  text and sa must stay alive (they are read on the host too);
    the object keeps references to them."""

    def __init__(self, text, sa, protein_start, protein_len, matrix, gap_open=11, gap_extend=1):
        L = _lib.lib()
        _declare(L)
        self._text = np.ascontiguousarray(np.asarray(text, np.uint8))
        self._sa = np.ascontiguousarray(np.asarray(sa, np.int64))
        self._start = np.ascontiguousarray(protein_start, np.uint64)
        self._len = np.ascontiguousarray(protein_len, np.int32)
        self._mat = np.ascontiguousarray(matrix, np.int8).reshape(-1)
        self.t_size = len(self._text)
        self.n_proteins = len(self._start)
        h = C.c_void_p()
        check(L.tpz_db_create(C.byref(h), self._text.ctypes.data, self.t_size,
                              self._sa.ctypes.data, self.n_proteins, self._start.ctypes.data,
                              self._len.ctypes.data, self._mat.ctypes.data, gap_open, gap_extend),
              "tpz_db_create")
        self._h = h

    @classmethod
    def from_index(cls, prefix, matrix, gap_open=11, gap_extend=1):
        text, sa, ind = load_index(prefix)
        return cls(text, sa, ind["protein_start"], ind["protein_len"], matrix, gap_open,
                   gap_extend)

    def db_size(self):
        """(db_size, db_proteins) as engine.c:52-55 sets them."""
        return self.t_size - self.n_proteins, self.n_proteins

    def default_options(self, **kw):
        ds, dp = self.db_size()
        return options(db_size=ds, db_proteins=dp, **kw)

    def search(self, queries, opt=None, schedule=SCHED_PERSISTENT):
        """queries: list of bytes (internal alphabet, after SEG).  Returns
        (list per query of hit dicts in printed order, stats dict)."""
        L = _lib.lib()
        if opt is None:
            ds, dp = self.db_size()
            opt = options(db_size=ds, db_proteins=dp)
        buf, qo, ql = _pack(queries)
        r = C.c_void_p()
        check(L.tpz_search(self._h, C.byref(opt), len(ql), buf, qo.ctypes.data, ql.ctypes.data,
                           schedule, C.byref(r)), "tpz_search")
        return _take_result(L, r, len(ql))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tpz_db_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _pack(queries):
    qs = [bytes(q) for q in queries]
    ql = np.array([len(q) for q in qs], np.int32)
    qo = np.zeros(len(qs), np.int64)
    if len(qs):
        qo[1:] = np.cumsum(ql[:-1])
    return b"".join(qs) + b"\0", qo, ql


def _take_result(L, r, n_queries):
    """(hits per query as dicts, stats dict) of a tpz_search_result, which
    is destroyed."""
    try:
        out = []
        for k in range(n_queries):
            n = L.tpz_search_result_count(r, k)
            hp = L.tpz_search_result_hits(r, k)
            out.append([{f: getattr(hp[i], f) for f in HIT_FIELDS} for i in range(n)])
        st = SearchStats()
        check(L.tpz_search_result_stats(r, C.byref(st)))
        return out, st.as_dict()
    finally:
        L.tpz_search_result_destroy(r)


class SearchServer:
    """tpz_search_server: searches requests (lists of queries) arriving from
    any thread in merged batches over a resident Database."""

    def __init__(self, db, opt=None, memory_budget=0, lanes=0, max_wave_queries=0):
        L = _lib.lib()
        _declare(L)
        self._db = db
        self.opt = opt if opt is not None else db.default_options()
        cfg = ServerConfig(memory_budget, lanes, max_wave_queries)
        h = C.c_void_p()
        check(L.tpz_search_server_create(C.byref(h), db._h, C.byref(self.opt), C.byref(cfg)),
              "tpz_search_server_create")
        self._h = h

    def submit(self, queries):
        buf, qo, ql = _pack(queries)
        t = C.c_uint64()
        check(_lib.lib().tpz_search_server_submit(self._h, len(ql), buf, qo.ctypes.data,
                                                  ql.ctypes.data, C.byref(t)),
              "tpz_search_server_submit")
        return (t.value, len(ql))

    def wait(self, ticket):
        """(hits per query, stats of the batch it ran in, latency in ms)."""
        L = _lib.lib()
        r = C.c_void_p()
        lat = C.c_double()
        check(L.tpz_search_server_wait(self._h, ticket[0], C.byref(r), C.byref(lat)),
              "tpz_search_server_wait")
        hits, st = _take_result(L, r, ticket[1])
        return hits, st, lat.value

    def stats(self):
        st = ServerStats()
        check(_lib.lib().tpz_search_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tpz_search_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# upper_aa (lexicographical.h:35-58): A..Y, '*' and '\n' kept, a..y upper-cased,
# anything else 'X'
_UPPER_AA = bytes(c if (65 <= c <= 89 or c in (42, 10)) else c - 32 if 97 <= c <= 121 else 88
                  for c in range(256))
# aa2internal (lexicographical.h:135-190) of upper-case letters; anything else 'V'
_AA2INTERNAL = bytearray(b"V" * 256)
for _a, _i in zip(b"ACMLJIVTSKRQZEDBNHYFWXP*G\n", b"ABCDEFGHIJKLMNOPQRSTUVWXY\n"):
    _AA2INTERNAL[_a] = _i
_AA2INTERNAL = bytes(_AA2INTERNAL)


def seg(letters):
    """SEG as TOPAZ runs it on a query (run_seg, seg_wrapper.c:15-49):
    low-complexity stretches of `letters` (bytes) replaced by 'X'."""
    L = _lib.lib()
    if not getattr(L, "_seg_declared", False):
        L.tpz_seg.argtypes = [C.c_char_p, C.c_int32]
        L.tpz_seg.restype = C.c_int32
        L._seg_declared = True
    buf = C.create_string_buffer(bytes(letters), len(letters) + 1)
    if L.tpz_seg(buf, len(letters)) < 0:
        raise _lib.TopazError("tpz_seg failed")
    return buf.raw[:len(letters)]


def prepare_query(letters, use_seg=True):
    """A query's FASTA letters (newlines removed) as TOPAZ searches them:
    upper_aa (seq_seq, seq.c:80-104), SEG unless -M (engine.c:254-263), then
    the internal alphabet (seq_2internal, seq.c:106-117)."""
    s = bytes(letters).translate(_UPPER_AA)
    if use_seg:
        s = seg(s)
    return s.translate(_AA2INTERNAL)


def read_queries(path, use_seg=True):
    """The queries of a FASTA file as TOPAZ reads them (fasta.c:60-76): a
    list of (id, internal-alphabet bytes); the id is the header up to its
    end of line, without the '>'."""
    out = []
    data = Path(path).read_bytes()
    for rec in data.split(b"\n>"):
        rec = rec[1:] if rec.startswith(b">") else rec
        if not rec.strip():
            continue
        head, _, body = rec.partition(b"\n")
        out.append((head.decode(), prepare_query(body.replace(b"\n", b""), use_seg)))
    return out


def format_hits(qid, hits, subject_ids, superfast=True):
    """The lines hit_print (engine.c:139-167) writes for one query."""
    lines = []
    for h in hits:
        sid = subject_ids[h["db_index"]]
        if superfast:
            lines.append("%s\t%s\t%.3e\t%d" % (qid, sid, h["evalue"], int(h["bitscore"])))
        else:
            lines.append("%s\t%s\t%.2f\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%.3e\t%d" % (
                qid, sid, h["identity"] * 100, h["length"], h["mismatch"], h["gapopen"],
                h["qstart"] + 1, h["qend"] + 1, h["sstart"] + 1, h["send"] + 1, h["evalue"],
                int(h["bitscore"])))
    return lines


def header_ids(prefix, ind):
    """Protein ids from the .phr file (db_proteinid, db.c:207-221)."""
    phr = Path(str(prefix) + ".phr").read_bytes()
    return [phr[s:s + n].decode() for s, n in zip(ind["header_start"].tolist(),
                                                   ind["header_len"].tolist())]
