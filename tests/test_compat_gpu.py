"""The drop-in ABI (include/topaz_ssw_compat.h) called the way TOPAZ calls
it, against the reference's own functions (compiled checker) or the oracle."""
import ctypes as C

import numpy as np
import pytest

from tests.refcheck import (REF_SO, HitT, OptionsT, Oracle, Reference, SAlign, SeqT, SProfile,
                            matrix)
from topaz_amd import _lib, synth
from topaz_amd.matrices import INTERNAL_ORDER, MATRIX_ORDER

pytestmark = pytest.mark.gpu


def lib():
    L = C.CDLL(str(_lib.LIB_PATH))
    P8 = C.POINTER(C.c_int8)
    L.ssw_init.restype = C.POINTER(SProfile)
    L.ssw_init.argtypes = [P8, C.c_int32, P8, C.c_int32, C.c_int8]
    L.ssw_align.restype = C.POINTER(SAlign)
    L.ssw_align.argtypes = [C.POINTER(SProfile), P8, C.c_int32, C.c_uint8, C.c_uint8, C.c_uint8,
                            C.c_uint16, C.c_int32, C.c_int32]
    L.align_destroy.argtypes = [C.POINTER(SAlign)]
    L.init_destroy.argtypes = [C.POINTER(SProfile)]
    PI = C.POINTER(C.c_int32)
    L.ssw_align_stats.restype = C.c_int
    L.ssw_align_stats.argtypes = [C.POINTER(SProfile), P8, C.c_int32, C.c_uint8, C.c_uint8,
                                  C.c_int32, C.c_int32, C.c_int32, C.c_uint16, PI, PI, PI, PI,
                                  PI, C.POINTER(C.c_double)]
    L.align_using_ssw_profile.argtypes = [C.c_void_p, C.POINTER(SeqT), C.POINTER(HitT),
                                          C.POINTER(OptionsT)]
    L.get_ssw_profile.restype = C.c_void_p
    L.get_ssw_profile.argtypes = [C.POINTER(SeqT), C.POINTER(OptionsT)]
    return L


def _p8(a):
    return a.ctypes.data_as(C.POINTER(C.c_int8))


def test_ssw_align_and_stats_match_oracle():
    L = lib()
    o = Oracle()
    mat = matrix()
    rng = np.random.default_rng(3)
    for k in range(40):
        q = synth.random_residues(rng, int(rng.integers(1, 300)))
        if k % 2:
            r_, o_, l_ = synth.mutate(rng, q, np.array([0]), np.array([len(q)]), np.array([0.6]))
            s = r_[:l_[0]]
        else:
            s = synth.random_residues(rng, int(rng.integers(1, 300)))
        prof = L.ssw_init(_p8(q), len(q), _p8(mat), 25, 2)
        a = L.ssw_align(prof, _p8(s), len(s), 11, 1, 0, 0, 0, 15)
        exp = o.ssw_align(q, s)
        got = (a.contents.score1, a.contents.ref_end1, a.contents.read_end1)
        assert got == exp, (k, got, exp)
        assert (a.contents.ref_begin1, a.contents.read_begin1, a.contents.score2) == (-1, -1, 0)
        if exp[0] > 0:
            v = [C.c_int32() for _ in range(5)]
            ident = C.c_double()
            rc = L.ssw_align_stats(prof, _p8(s), len(s), 11, 1, 15, exp[2], exp[1], exp[0],
                                   *[C.byref(x) for x in v], C.byref(ident))
            e2 = o.ssw_align_stats(q, s, exp[2], exp[1], exp[0])
            assert (rc,) + tuple(x.value for x in v) == tuple(e2[:6])
            assert ident.value == e2[6] or (ident.value != ident.value and e2[6] != e2[6])
        L.align_destroy(a)
        L.init_destroy(prof)


def _internal(idx):
    return bytes(ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[i]) for i in idx)


@pytest.mark.skipif(not REF_SO.exists(), reason="reference checker not built here")
def test_align_using_ssw_profile_matches_reference():
    """hit_t fields written by align_using_ssw_profile, ours vs the
    reference's compiled function, on the same seq_t / options_t."""
    L = lib()
    R = Reference().L
    R.get_ssw_profile.restype = C.c_void_p
    R.get_ssw_profile.argtypes = [C.POINTER(SeqT), C.POINTER(OptionsT)]
    R.align_using_ssw_profile.argtypes = L.align_using_ssw_profile.argtypes
    rng = np.random.default_rng(4)
    mat = matrix()
    mat_c = (C.c_int8 * 625)(*mat.tolist())
    opt = OptionsT()
    opt.substitution = C.cast(mat_c, C.c_void_p)
    opt.gap_open, opt.gap_extend = 11, 1
    opt.lambda_, opt.k, opt.H = 0.267, 0.041, 0.140
    opt.db_size, opt.db_proteins = 35_000_000, 100_000
    fields = ["rawscore", "send", "qend", "sstart", "qstart", "bitscore", "evalue",
              "is_aligned", "effective_length", "length", "identity"]
    for k in range(30):
        qi = synth.random_residues(rng, int(rng.integers(5, 300)))
        r_, o_, l_ = synth.mutate(rng, qi, np.array([0]), np.array([len(qi)]), np.array([0.5]))
        si = r_[:l_[0]] if k % 3 else synth.random_residues(rng, 200)
        qb, sb = _internal(qi), _internal(si)
        seqs = []
        for b in (qb, sb):
            sq = SeqT()
            sq.seq_buf.buf = b
            sq.seq_buf.len = len(b)
            sq.seq_len = len(b)
            seqs.append(sq)
        hits = []
        for impl in (L, R):
            prof = impl.get_ssw_profile(C.byref(seqs[0]), C.byref(opt))
            h = HitT()
            h.gapless_score = [10, 39, 40, 100][k % 4]
            impl.align_using_ssw_profile(prof, C.byref(seqs[1]), C.byref(h), C.byref(opt))
            hits.append(tuple(getattr(h, f) for f in fields))
        assert hits[0] == hits[1], (k, hits)


def test_sa_hsp_compat_matches_oracle():
    from tests.refcheck import DbT
    L = lib()
    L.sa_hsp.restype = C.c_int
    L.sa_hsp.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.c_int64,
                         C.POINTER(OptionsT), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rng = np.random.default_rng(5)
    T = b"".join(_internal(synth.random_residues(rng, int(L_))) + b"\n"
                 for L_ in synth.lognormal_lengths(rng, 50, 120, 80, 5, 400))
    SA = np.arange(len(T), dtype=np.int64)
    db = DbT()
    db.T = T
    db.Tsize = len(T)
    db.SA = SA.ctypes.data_as(C.POINTER(C.c_int64))
    db.SAsize = len(T) * 8
    mat = matrix()
    mat_c = (C.c_int8 * 625)(*mat.tolist())
    opt = OptionsT()
    opt.substitution = C.cast(mat_c, C.c_void_p)
    o = Oracle()
    Q = T[100:400].replace(b"\n", b"A")
    sq = SeqT()
    sq.seq_buf.buf = Q
    sq.seq_buf.len = len(Q)
    sq.seq_len = len(Q)
    for _ in range(60):
        off = int(rng.integers(0, len(Q)))
        pos = int(rng.integers(0, len(T)))
        if rng.random() < 0.5:
            pos = 100 + off
        sc, ln = C.c_int(-7), C.c_int(0)
        L.sa_hsp(C.byref(db), C.byref(sq), off, pos, C.byref(opt), C.byref(sc), C.byref(ln))
        es, el = o.sa_hsp(T, pos, Q, off)
        assert ln.value == el
        assert sc.value == (es if el != -1 else -7)
