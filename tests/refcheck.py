"""ctypes bindings to the two CPU checkers used by the tests.

  Oracle     oracle/liboracle.so       clean-room C restatement (orc.h)
  Reference  oracle/_ref/libtopazref.so  the reference's own sources, compiled
             by oracle/Makefile (present only where it was built)

Both are test infrastructure: the product never imports this module.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libtopazref.so"

_mats = json.loads((ROOT / "topaz_amd" / "data" / "matrices.json").read_text())
MATRIX_ORDER = _mats["order"]


def matrix(name="BLOSUM62"):
    return np.array(_mats["matrices"][name], dtype=np.int8)


def _ptr8(a):
    return a.ctypes.data_as(C.POINTER(C.c_int8))


class OrcEnd(C.Structure):
    _fields_ = [("score", C.c_int32), ("ref", C.c_int32), ("read", C.c_int32)]


class Oracle:
    def __init__(self, path=ORACLE_SO):
        L = C.CDLL(str(path))
        P8 = C.POINTER(C.c_int8)
        I32 = C.c_int32
        PI = C.POINTER(C.c_int32)
        L.orc_sw_scan.restype = OrcEnd
        L.orc_sw_scan.argtypes = [P8, I32, P8, I32, C.c_int, P8, I32, C.c_int,
                                  I32, I32, I32, I32]
        L.orc_sw_scan_s1.restype = OrcEnd
        L.orc_sw_scan_s1.argtypes = L.orc_sw_scan.argtypes
        L.orc_sw_textbook.restype = OrcEnd
        L.orc_sw_textbook.argtypes = [P8, I32, P8, I32, P8, I32, C.c_int, I32, I32, I32]
        L.orc_ssw_align.restype = C.c_int
        L.orc_ssw_align.argtypes = [P8, I32, P8, I32, P8, I32, I32, I32,
                                    C.POINTER(OrcEnd)]
        L.orc_ssw_align_stats.restype = C.c_int
        L.orc_ssw_align_stats.argtypes = [P8, I32, P8, I32, P8, I32, I32, I32,
                                          I32, I32, I32, PI, PI, PI, PI, PI,
                                          C.POINTER(C.c_double)]
        L.orc_banded_sw.restype = C.c_int
        L.orc_banded_sw.argtypes = [P8, P8, I32, I32, I32, I32, I32, I32, P8,
                                    I32, PI, PI, PI, C.POINTER(C.c_double)]
        L.orc_sa_hsp.restype = C.c_int
        L.orc_sa_hsp.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_char_p,
                                 I32, I32, P8, PI]
        P64 = C.POINTER(C.c_int64)
        L.orc_suffix_array.restype = C.c_int
        L.orc_suffix_array.argtypes = [C.c_char_p, C.c_int64, P64]
        L.orc_sa_search.restype = C.c_int
        L.orc_sa_search.argtypes = [C.c_char_p, C.c_int64, P64, C.c_char_p, C.c_int64,
                                    P64, P64, P64]
        L.orc_sa_gapless.restype = C.c_int
        L.orc_sa_gapless.argtypes = L.orc_sa_hsp.argtypes
        L.orc_sa_hsp_batch.restype = None
        L.orc_sa_hsp_batch.argtypes = [C.c_char_p, C.c_int64, C.c_char_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        L.orc_evalue.restype = None
        L.orc_evalue.argtypes = [I32, I32, C.c_uint64, C.c_uint64, C.c_double,
                                 C.c_double, C.c_double, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), PI]
        L.orc_aa_table2.restype = C.c_int8
        L.orc_aa_table2.argtypes = [C.c_ubyte]
        self.L = L

    def sw_scan(self, q, s, P, go=11, ge=1, mat=None, dir=0, terminate=-1,
                closed_form=False):
        mat = matrix() if mat is None else mat
        bias = -min(0, int(mat.min()))
        f = self.L.orc_sw_scan_s1 if closed_form else self.L.orc_sw_scan
        r = f(_ptr8(q), len(q), _ptr8(s), len(s), dir, _ptr8(mat), 25, P, go, ge,
              bias, terminate)
        return r.score, r.ref, r.read

    def sw_textbook(self, q, s, P, go=11, ge=1, mat=None):
        mat = matrix() if mat is None else mat
        bias = -min(0, int(mat.min()))
        r = self.L.orc_sw_textbook(_ptr8(q), len(q), _ptr8(s), len(s), _ptr8(mat), 25, P,
                                   go, ge, bias)
        return r.score, r.ref, r.read

    def ssw_align(self, q, s, go=11, ge=1, mat=None):
        """(score, ref_end, read_end) or None where the reference gives NULL."""
        mat = matrix() if mat is None else mat
        out = OrcEnd()
        if self.L.orc_ssw_align(_ptr8(q), len(q), _ptr8(s), len(s), _ptr8(mat),
                                25, go, ge, C.byref(out)):
            return None
        return out.score, out.ref, out.read

    def ssw_align_stats(self, q, s, read_end, ref_end, score, go=11, ge=1,
                        mat=None):
        mat = matrix() if mat is None else mat
        v = [C.c_int32() for _ in range(5)]
        ident = C.c_double()
        rc = self.L.orc_ssw_align_stats(
            _ptr8(q), len(q), _ptr8(s), len(s), _ptr8(mat), 25, go, ge,
            read_end, ref_end, score, *[C.byref(x) for x in v], C.byref(ident))
        return (rc,) + tuple(x.value for x in v) + (ident.value,)

    def suffix_array(self, T):
        sa = np.zeros(max(len(T), 1), np.int64)
        assert self.L.orc_suffix_array(T, len(T), sa.ctypes.data_as(C.POINTER(C.c_int64))) == 0
        return sa[:len(T)]

    def sa_search(self, T, SA, P):
        """(return, idx, num_matches, last_position) for the suffix P."""
        v = [C.c_int64(-9) for _ in range(3)]
        rc = self.L.orc_sa_search(T, len(T), SA.ctypes.data_as(C.POINTER(C.c_int64)), P, len(P),
                                  *[C.byref(x) for x in v])
        return (rc,) + tuple(x.value for x in v)

    def sa_hsp(self, T, idx, Q, o, mat=None):
        mat = matrix() if mat is None else mat
        sc = C.c_int32(0)
        ln = self.L.orc_sa_hsp(T, len(T), idx, Q, len(Q), o, _ptr8(mat),
                               C.byref(sc))
        return (sc.value if ln != -1 else None), ln

    def sa_hsp_batch(self, T, queries, seeds, mat=None):
        """(score, length) of every seed record (query, offset, position);
        score is 0 where length is -1.  T: bytes; queries: list of bytes."""
        mat = matrix() if mat is None else mat
        ql = np.array([len(q) for q in queries], np.int32)
        qo = np.zeros(len(ql), np.int64)
        qo[1:] = np.cumsum(ql[:-1])
        seeds = np.ascontiguousarray(seeds)
        sc = np.zeros(len(seeds), np.int32)
        ln = np.zeros(len(seeds), np.int32)
        self.L.orc_sa_hsp_batch(T, len(T), b"".join(queries), qo.ctypes.data, ql.ctypes.data,
                                seeds.ctypes.data, len(seeds), _ptr8(mat), sc.ctypes.data,
                                ln.ctypes.data)
        return sc, ln

    def evalue(self, raw, qlen, db_size, db_proteins, lam=0.267, k=0.041,
               H=0.140):
        b, e, el = C.c_double(), C.c_double(), C.c_int32()
        self.L.orc_evalue(raw, qlen, db_size, db_proteins, lam, k, H,
                          C.byref(b), C.byref(e), C.byref(el))
        return b.value, e.value, el.value


class SAlign(C.Structure):
    _fields_ = [("score1", C.c_uint16), ("score2", C.c_uint16),
                ("ref_begin1", C.c_int32), ("ref_end1", C.c_int32),
                ("read_begin1", C.c_int32), ("read_end1", C.c_int32),
                ("ref_end2", C.c_int32), ("cigar", C.c_void_p),
                ("cigarLen", C.c_int32)]


class SProfile(C.Structure):
    _fields_ = [("profile_byte", C.c_void_p), ("profile_word", C.c_void_p),
                ("read", C.c_void_p), ("mat", C.c_void_p),
                ("readLen", C.c_int32), ("n", C.c_int32), ("bias", C.c_uint8)]


class Reference:
    """The reference's compiled ssw.c / search.c, driven through its own API."""

    def __init__(self, path=REF_SO):
        L = C.CDLL(str(path))
        P8 = C.POINTER(C.c_int8)
        L.ssw_init.restype = C.POINTER(SProfile)
        L.ssw_init.argtypes = [P8, C.c_int32, P8, C.c_int32, C.c_int8]
        L.ssw_align.restype = C.POINTER(SAlign)
        L.ssw_align.argtypes = [C.POINTER(SProfile), P8, C.c_int32, C.c_uint8,
                                C.c_uint8, C.c_uint8, C.c_uint16, C.c_int32,
                                C.c_int32]
        L.align_destroy.argtypes = [C.POINTER(SAlign)]
        L.init_destroy.argtypes = [C.POINTER(SProfile)]
        PI = C.POINTER(C.c_int32)
        L.ssw_align_stats.restype = C.c_int
        L.ssw_align_stats.argtypes = [C.POINTER(SProfile), P8, C.c_int32,
                                      C.c_uint8, C.c_uint8, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_uint16, PI, PI,
                                      PI, PI, PI, C.POINTER(C.c_double)]
        self.L = L

    def profile(self, q, mat=None):
        mat = matrix() if mat is None else mat
        # keep buffers alive for the profile's lifetime
        qb = np.ascontiguousarray(q, dtype=np.int8)
        mb = np.ascontiguousarray(mat, dtype=np.int8)
        p = self.L.ssw_init(_ptr8(qb), len(qb), _ptr8(mb), 25, 2)
        return p, (qb, mb)

    def ssw_align(self, q, s, go=11, ge=1, mat=None, prof=None):
        own = prof is None
        if own:
            prof = self.profile(q, mat)
        sb = np.ascontiguousarray(s, dtype=np.int8)
        r = self.L.ssw_align(prof[0], _ptr8(sb), len(sb), go, ge, 0, 0, 0, 15)
        out = None
        if r:
            a = r.contents
            out = (a.score1, a.ref_end1, a.read_end1)
            self.L.align_destroy(r)
        if own:
            self.L.init_destroy(prof[0])
        return out

    def ssw_align_stats(self, q, s, read_end, ref_end, score, go=11, ge=1,
                        mat=None):
        prof = self.profile(q, mat)
        sb = np.ascontiguousarray(s, dtype=np.int8)
        v = [C.c_int32(0) for _ in range(5)]
        ident = C.c_double()
        rc = self.L.ssw_align_stats(prof[0], _ptr8(sb), len(sb), go, ge, 15,
                                    read_end, ref_end, score,
                                    *[C.byref(x) for x in v], C.byref(ident))
        self.L.init_destroy(prof[0])
        return (rc,) + tuple(x.value for x in v) + (ident.value,)


# ---- the reference's structs (seq.h, db.h, options.h, hit.h) -------------
class BufferT(C.Structure):
    _fields_ = [("buf", C.c_char_p), ("len", C.c_int)]


class SeqT(C.Structure):
    _fields_ = [("id_buf", BufferT), ("seq_buf", BufferT), ("id_len", C.c_int),
                ("seq_len", C.c_int)]


class DbT(C.Structure):
    _fields_ = [("prefix", C.c_char_p), ("headers", C.c_char_p), ("Hsize", C.c_uint64),
                ("index", C.c_void_p), ("Psize", C.c_uint64), ("T", C.c_char_p),
                ("Tsize", C.c_uint64), ("SA", C.POINTER(C.c_int64)), ("SAsize", C.c_uint64)]


class OptionsT(C.Structure):
    _fields_ = [("num_threads", C.c_int), ("num_hits", C.c_int), ("evalue", C.c_double),
                ("smithwaterman", C.c_int), ("gap_open", C.c_int), ("gap_extend", C.c_int),
                ("sub_name", C.c_char_p), ("substitution", C.c_void_p),
                ("db_size", C.c_uint64), ("db_proteins", C.c_uint64),
                ("lambda_", C.c_double), ("k", C.c_double), ("H", C.c_double),
                ("min_hit_length", C.c_int), ("min_hit_score", C.c_int),
                ("number_of_alignments", C.c_int), ("number_of_seeds", C.c_int),
                ("seg_enabled", C.c_int), ("superfast", C.c_int)]


class HitT(C.Structure):
    _fields_ = [("db_index", C.c_uint64), ("is_aligned", C.c_int),
                ("alignment_rank", C.c_int), ("rawscore", C.c_int), ("mismatch", C.c_int),
                ("length", C.c_int), ("gapopen", C.c_int), ("identity", C.c_double),
                ("bitscore", C.c_double), ("evalue", C.c_double), ("qstart", C.c_int),
                ("qend", C.c_int), ("sstart", C.c_int), ("send", C.c_int),
                ("effective_length", C.c_int), ("gapless_score", C.c_int),
                ("gapless_length", C.c_int), ("suffix_counter", C.c_int),
                ("suffix_offset", C.c_int), ("suffix_multiple", C.c_int),
                ("time_stats", C.c_double), ("time_align", C.c_double),
                ("priority_node", C.c_void_p)]


