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
    # a permuted suffix array: positions are SA slots, the text index SA[p]
    SA = rng.permutation(len(T)).astype(np.int64)
    slot_of = np.argsort(SA)
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
        idx = int(rng.integers(0, len(T)))
        if rng.random() < 0.5:
            idx = 100 + off
        sc, ln = C.c_int(-7), C.c_int(0)
        L.sa_hsp(C.byref(db), C.byref(sq), off, int(slot_of[idx]), C.byref(opt), C.byref(sc),
                 C.byref(ln))
        es, el = o.sa_hsp(T, idx, Q, off)
        assert ln.value == el
        assert sc.value == (es if el != -1 else -7)


def _upper_aa(b):
    """upper_aa of lexicographical.h: A-Y, '*' and newline kept, a-y raised,
    anything else 'X'."""
    if 65 <= b <= 89 or b in (42, 10):
        return b
    if 97 <= b <= 121:
        return b - 32
    return 88


def _db_seq(T, start, n):
    """db_seq (db.c:233-247): the protein's bytes without newlines, through
    upper_aa."""
    return bytes(_upper_aa(c) for c in T[start:start + n] if c != 10)


class PositionT(C.Structure):
    _fields_ = [("header_start", C.c_uint64), ("protein_start", C.c_uint64),
                ("header_len", C.c_int), ("protein_len", C.c_int)]


def _seq(b):
    sq = SeqT()
    sq.seq_buf.buf = b
    sq.seq_buf.len = len(b)
    sq.seq_len = len(b)
    return sq


HIT_FIELDS = ["db_index", "is_aligned", "alignment_rank", "rawscore", "mismatch", "length",
              "gapopen", "identity", "bitscore", "evalue", "qstart", "qend", "sstart", "send",
              "effective_length", "gapless_score"]


def _same(a, b):
    return a == b or (a != a and b != b)


def test_batch_entry_points_equal_per_hit_calls():
    """align_batch_using_ssw_profile / stats_batch_using_ssw_profile /
    sa_hsp_batch write exactly what the per-hit calls write, on a database
    with an index (some letters in lower case, one entry spanning a
    newline), and agree with the reference's compiled per-hit functions."""
    from tests.refcheck import DbT
    L = lib()
    PH = C.POINTER(HitT)
    L.align_batch_using_ssw_profile.restype = C.c_int
    L.align_batch_using_ssw_profile.argtypes = [C.c_void_p, C.POINTER(DbT), C.POINTER(PH),
                                                C.c_int, C.POINTER(OptionsT)]
    L.stats_batch_using_ssw_profile.argtypes = [C.c_void_p, C.POINTER(DbT), C.POINTER(PH),
                                                C.c_int, C.POINTER(OptionsT)]
    L.stats_using_ssw_profile.argtypes = L.align_using_ssw_profile.argtypes
    PI = C.POINTER(C.c_int)
    L.sa_hsp.restype = C.c_int
    L.sa_hsp.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.c_int64,
                         C.POINTER(OptionsT), PI, PI]
    L.sa_hsp_batch.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, PI,
                               C.POINTER(C.c_int64), C.POINTER(OptionsT), PI, PI]
    R = None
    if REF_SO.exists():
        R = Reference().L
        R.get_ssw_profile.restype = C.c_void_p
        R.get_ssw_profile.argtypes = L.get_ssw_profile.argtypes
        R.align_using_ssw_profile.argtypes = L.align_using_ssw_profile.argtypes
        R.stats_using_ssw_profile.argtypes = L.align_using_ssw_profile.argtypes

    rng = np.random.default_rng(6)
    query_i = synth.random_residues(rng, 260)
    lens = synth.lognormal_lengths(rng, 120, 250, 150, 5, 900)
    prots = []
    for k, n in enumerate(lens):
        if k % 3 == 0:      # homologs of the query
            r_, o_, l_ = synth.mutate(rng, query_i, np.array([0]), np.array([len(query_i)]),
                                      np.array([0.4 + 0.5 * rng.random()]))
            p = _internal(r_[:l_[0]])
        else:
            p = _internal(synth.random_residues(rng, int(n)))
        if k % 7 == 0:     # a few letters in lower case (the text stays
            p = p.translate(bytes.maketrans(b"ALK", b"alk"))   # within 31 letters)
        prots.append(p)
    T = b"".join(p + b"\n" for p in prots)
    starts = np.cumsum([0] + [len(p) + 1 for p in prots])[:-1]
    index = (PositionT * (len(prots) + 1))()
    for k, p in enumerate(prots):
        index[k].protein_start = int(starts[k])
        index[k].protein_len = len(p)
    # one entry spanning the newline between proteins 4 and 5
    index[len(prots)].protein_start = int(starts[4])
    index[len(prots)].protein_len = len(prots[4]) + 1 + len(prots[5])
    SA = rng.permutation(len(T)).astype(np.int64)       # positions are SA slots
    slot_of = np.argsort(SA)
    db = DbT()
    db.T = T
    db.Tsize = len(T)
    db.index = C.cast(index, C.c_void_p)
    db.Psize = len(prots) + 1
    db.SA = SA.ctypes.data_as(C.POINTER(C.c_int64))
    db.SAsize = len(T) * 8

    mat = matrix()
    mat_c = (C.c_int8 * 625)(*mat.tolist())
    opt = OptionsT()
    opt.substitution = C.cast(mat_c, C.c_void_p)
    opt.gap_open, opt.gap_extend = 11, 1
    opt.lambda_, opt.k, opt.H = 0.267, 0.041, 0.140
    opt.db_size, opt.db_proteins = 35_000_000, 100_000
    opt.evalue = 10.0
    q = _seq(_internal(query_i))

    n = 150
    db_index = rng.integers(0, len(prots) + 1, n)
    gscore = rng.choice([10, 39, 40, 100], n)

    def fresh():
        hs = [HitT() for _ in range(n)]
        for h, d, g in zip(hs, db_index, gscore):
            h.db_index, h.gapless_score, h.rawscore = int(d), int(g), 0
        return hs

    # per hit, as query.c:199-228 and :276-291 drive it
    def per_hit(impl):
        prof = impl.get_ssw_profile(C.byref(q), C.byref(opt))
        hs = fresh()
        count = 0
        for h in hs:
            pos = index[h.db_index]
            s = _seq(_db_seq(T, pos.protein_start, pos.protein_len))
            impl.align_using_ssw_profile(prof, C.byref(s), C.byref(h), C.byref(opt))
            count += h.evalue < opt.evalue
        for h in hs:
            if h.evalue <= opt.evalue and h.rawscore != 32767:
                pos = index[h.db_index]
                s = _seq(_db_seq(T, pos.protein_start, pos.protein_len))
                impl.stats_using_ssw_profile(prof, C.byref(s), C.byref(h), C.byref(opt))
        return hs, count

    mine, count = per_hit(L)
    assert 10 < count < n
    prof = L.get_ssw_profile(C.byref(q), C.byref(opt))
    batch = fresh()
    got = L.align_batch_using_ssw_profile(prof, C.byref(db), (PH * n)(*[C.pointer(h) for h in batch]),
                                          n, C.byref(opt))
    assert got == count
    todo = [h for h in batch if h.evalue <= opt.evalue]
    L.stats_batch_using_ssw_profile(prof, C.byref(db),
                                    (PH * len(todo))(*[C.pointer(h) for h in todo]),
                                    len(todo), C.byref(opt))
    for k in range(n):
        for f in HIT_FIELDS:
            assert _same(getattr(batch[k], f), getattr(mine[k], f)), (k, f)
    assert any(h.length > 0 for h in batch)
    if R is not None:
        ref, ref_count = per_hit(R)
        assert ref_count == count
        for k in range(n):
            for f in HIT_FIELDS:
                assert _same(getattr(batch[k], f), getattr(ref[k], f)), (k, f)

    # gapless: n calls of sa_hsp against one call of sa_hsp_batch
    m = 400
    offs = rng.integers(0, q.seq_len, m).astype(np.int32)
    pos = rng.integers(0, len(T), m).astype(np.int64)
    pos[::2] = slot_of[(starts[(np.arange(0, m, 2) * 3) % len(prots)] + offs[::2]).clip(0, len(T) - 1)]
    sc = np.full(m, -7, np.int32)
    ln = np.zeros(m, np.int32)
    L.sa_hsp_batch(C.byref(db), C.byref(q), m, offs.ctypes.data_as(PI),
                   pos.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(opt),
                   sc.ctypes.data_as(PI), ln.ctypes.data_as(PI))
    assert (ln != -1).sum() > 20
    o = Oracle()
    Q = _internal(query_i)
    for k in range(m):
        s1, l1 = C.c_int(-7), C.c_int(0)
        L.sa_hsp(C.byref(db), C.byref(q), int(offs[k]), int(pos[k]), C.byref(opt),
                 C.byref(s1), C.byref(l1))
        assert (sc[k], ln[k]) == (s1.value, l1.value), k
        es, el = o.sa_hsp(T, int(SA[pos[k]]), Q, int(offs[k]))
        assert (ln[k], sc[k]) == (el, es if el != -1 else -7), k


def test_sa_search_batch_equals_reference():
    """sa_search_batch, and the library's single sa_search with and without
    a pattern length, give what the reference's own sa_search
    (search.c:237-361, compiled from its sources) writes: the interval
    start, the number of matches and the last match length, for suffixes
    that occur (query parts copied from the text), suffixes that do not,
    repeats, and an offset past the query's end (return -1)."""
    from tests.refcheck import DbT
    if not REF_SO.exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    L, R = lib(), Reference().L
    rng = np.random.default_rng(61)
    prots = [_internal(synth.random_residues(rng, int(n)))
             for n in synth.lognormal_lengths(rng, 300, 200, 120, 5, 900)]
    prots[7] = b"A" * 40 + prots[7]                   # a run the queries repeat
    T = b"".join(p + b"\n" for p in prots)
    SA = np.zeros(len(T), np.int64)
    R.divsufsort64.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
    R.divsufsort64.restype = C.c_int32
    assert R.divsufsort64(T, SA.ctypes.data_as(C.POINTER(C.c_int64)), len(T)) == 0
    db = DbT()
    db.T = T
    db.Tsize = len(T)
    db.SA = SA.ctypes.data_as(C.POINTER(C.c_int64))
    db.SAsize = len(T) * 8
    PI64 = C.POINTER(C.c_int64)
    R.sa_search.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int64, C.c_int64, PI64, PI64, PI64]
    R.sa_search.restype = C.c_int
    L.sa_search_batch.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.POINTER(C.c_int),
                                  PI64, PI64, PI64]
    L.sa_search_batch.restype = C.c_int
    L.sa_search.argtypes = R.sa_search.argtypes
    L.sa_search.restype = C.c_int
    queries = [prots[3][10:200] + _internal(synth.random_residues(rng, 60)) + prots[11][:50],
               _internal(synth.random_residues(rng, 300)),
               b"A" * 45 + prots[7][40:90],
               prots[20]]
    checked = 0
    for qb in queries:
        q = _seq(qb)
        offs = np.arange(len(qb), dtype=np.int32)
        n = len(offs)
        idx = np.full(n, -9, np.int64)
        nm = np.full(n, -9, np.int64)
        lp = np.full(n, -9, np.int64)
        assert L.sa_search_batch(C.byref(db), C.byref(q), n, offs.ctypes.data_as(C.POINTER(C.c_int)),
                                 idx.ctypes.data_as(PI64), nm.ctypes.data_as(PI64),
                                 lp.ctypes.data_as(PI64)) == 0
        for k in range(n):
            ri, rn, rl = C.c_int64(-9), C.c_int64(-9), C.c_int64(-9)
            assert R.sa_search(C.byref(db), C.byref(q), int(offs[k]), 0, C.byref(ri), C.byref(rn),
                               C.byref(rl)) == 0
            assert (idx[k], nm[k], lp[k]) == (ri.value, rn.value, rl.value), (k, qb[k:k + 20])
            checked += 1
        # the single call, with the pattern's length given (search.c:268-270)
        # or taken to the query's end
        for k in [int(x) for x in rng.integers(0, n, 25)]:
            for length in [x for x in (0, 1, 3, 12, len(qb) - k) if x <= len(qb) - k]:
                got, exp = [], []
                for fn, out in ((L.sa_search, got), (R.sa_search, exp)):
                    vi, vn, vl = C.c_int64(-9), C.c_int64(-9), C.c_int64(-9)
                    rc = fn(C.byref(db), C.byref(q), k, length, C.byref(vi), C.byref(vn),
                            C.byref(vl))
                    out.append((rc, vi.value, vn.value, vl.value))
                assert got == exp, (k, length, got, exp)
                checked += 1
        for fn in (L.sa_search, R.sa_search):
            vi, vn = C.c_int64(-9), C.c_int64(-9)
            assert fn(C.byref(db), C.byref(q), len(qb), 0, C.byref(vi), C.byref(vn), None) == -1
            assert (vi.value, vn.value) == (-1, -9)
        # past the end: both return -1 with idx -1
        end = np.array([len(qb)], np.int32)
        one = np.full(1, -9, np.int64)
        assert L.sa_search_batch(C.byref(db), C.byref(q), 1, end.ctypes.data_as(C.POINTER(C.c_int)),
                                 one.ctypes.data_as(PI64), None, None) == -1 and one[0] == -1
    assert checked > 500


def test_statistics_from_threads_equal_one_at_a_time():
    """ssw_align_stats called from 8 threads at once, with two gap settings
    mixed, gives what the same calls give one at a time: the library's
    statistics worker merges concurrent calls into shared runs, and keeps
    calls of different settings apart."""
    from concurrent.futures import ThreadPoolExecutor
    L = lib()
    mat = matrix()
    rng = np.random.default_rng(71)
    jobs = []
    keep = []
    for k in range(48):
        q = synth.random_residues(rng, int(rng.integers(80, 700)))
        r_, o_, l_ = synth.mutate(rng, q, np.array([0]), np.array([len(q)]),
                                  np.array([rng.uniform(0.4, 0.9)]))
        s = r_[:l_[0]].copy()
        go, ge = (11, 1) if k % 2 else (10, 2)
        prof = L.ssw_init(_p8(q), len(q), _p8(mat), 25, 2)
        a = L.ssw_align(prof, _p8(s), len(s), go, ge, 0, 0, 0, 15)
        sc, fe, re_ = a.contents.score1, a.contents.ref_end1, a.contents.read_end1
        L.align_destroy(a)
        keep.append((q, s, prof))
        if sc > 0:
            jobs.append((prof, s, go, ge, re_, fe, sc))

    def stats(job):
        prof, s, go, ge, re_, fe, sc = job
        v = [C.c_int32() for _ in range(5)]
        ident = C.c_double()
        rc = L.ssw_align_stats(prof, _p8(s), len(s), go, ge, 15, re_, fe, sc,
                               *[C.byref(x) for x in v], C.byref(ident))
        return (rc,) + tuple(x.value for x in v) + (repr(ident.value),)

    alone = [stats(j) for j in jobs]
    with ThreadPoolExecutor(8) as ex:
        together = list(ex.map(stats, jobs * 3))
    assert len(jobs) > 40
    assert together == alone * 3
    for _, _, prof in keep:
        L.init_destroy(prof)
