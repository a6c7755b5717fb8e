"""The reference's exports that TOPAZ never calls (sa_lcp, sa_gapless,
ungapped_rawscore, ungapped_length, the Gotoh aligner and its helpers):
libtopaz_amd.so against the reference's own compiled functions
(search.c:32-97, alignment.c:27-396), call for call, on the same structs and
arrays.  Host code on both sides, so these run without a GPU."""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest

from tests.refcheck import REF_SO, DbT, HitT, OptionsT, SeqT, matrix
from topaz_amd import _lib, synth

pytestmark = pytest.mark.skipif(not REF_SO.exists(),
                                reason="reference checker not built (make -C oracle ref)")

PI = C.POINTER(C.c_int)
PU16 = C.POINTER(C.c_uint16)
P8 = C.POINTER(C.c_int8)

LETTERS = b"ACDEFGHIKLMNPQRSTVWYBZXJUO*acdefghiklmnpqrstvwy"


def _bind(L):
    L.sa_lcp.restype = C.c_int
    L.sa_lcp.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.c_int64]
    L.sa_gapless.restype = C.c_int
    L.sa_gapless.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.c_int64,
                             C.POINTER(OptionsT), PI, PI]
    L.ungapped_rawscore.restype = C.c_int
    L.ungapped_rawscore.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(OptionsT)]
    L.ungapped_length.restype = C.c_int
    L.ungapped_length.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(OptionsT)]
    L._initialise.restype = None
    L._initialise.argtypes = [C.POINTER(SeqT), C.POINTER(SeqT), PI, PI, PI, PU16, PU16, PU16,
                              C.POINTER(OptionsT)]
    L._forward_pass.restype = C.c_int
    L._forward_pass.argtypes = [P8, P8, C.c_size_t, C.c_size_t, PI, PU16, PI, PU16, PI, PU16,
                                C.POINTER(HitT), C.POINTER(OptionsT)]
    L._backtrack.restype = None
    L._backtrack.argtypes = [PU16, PU16, PU16, C.c_int, C.c_int, C.POINTER(HitT)]
    L.gotoh.restype = C.c_int
    L.gotoh.argtypes = [C.POINTER(SeqT), C.POINTER(SeqT), C.POINTER(HitT), C.POINTER(OptionsT)]
    L.printdp32.restype = None
    L.printdp32.argtypes = [PI, C.c_size_t, C.c_size_t]
    L.printdp8.restype = None
    L.printdp8.argtypes = [C.POINTER(C.c_uint8), C.c_size_t, C.c_size_t]
    return L


@pytest.fixture(scope="module")
def libs():
    return _bind(C.CDLL(str(_lib.LIB_PATH))), _bind(C.CDLL(str(REF_SO)))


@pytest.fixture(scope="module")
def opt():
    mat = matrix()
    mat_c = (C.c_int8 * 625)(*mat.tolist())
    o = OptionsT()
    o.substitution = C.cast(mat_c, C.c_void_p)
    o.gap_open, o.gap_extend = 11, 1
    o.lambda_, o.k, o.H = 0.267, 0.041, 0.140
    o.db_size, o.db_proteins = 35_000_000, 100_000
    o._keep = mat_c
    return o


def _seq(b):
    sq = SeqT()
    sq.seq_buf.buf = b
    sq.seq_buf.len = len(b)
    sq.seq_len = len(b)
    return sq


def _letters(rng, n, alphabet=LETTERS):
    return bytes(alphabet[i] for i in rng.integers(0, len(alphabet), n))


def _db(rng):
    """Upper-case proteins, '\\n'-terminated, with a suffix array (the
    reference's divsufsort64)."""
    R = C.CDLL(str(REF_SO))
    prots = [_letters(rng, int(n), b"ACDEFGHIKLMNPQRSTVWY")
             for n in synth.lognormal_lengths(rng, 120, 150, 100, 2, 600)]
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
    rank = np.empty(len(T), np.int64)
    rank[SA] = np.arange(len(T))
    return db, T, SA, rank, (T, SA)


def test_sa_lcp_and_sa_gapless(libs, opt):
    """Seeds that share their first letter (query windows copied from the
    text, then mutated, lower-cased or given other bytes) and seeds that do
    not; query offsets up to the end, where fewer than 3 letters remain."""
    L, R = libs
    rng = np.random.default_rng(5)
    db, T, SA, rank, keep = _db(rng)
    checked = gapless = 0
    for trial in range(40):
        start = int(rng.integers(0, len(T) - 400))
        q = bytearray(T[start:start + int(rng.integers(3, 300))].replace(b"\n", b"K"))
        for _ in range(int(rng.integers(0, 12))):
            q[int(rng.integers(0, len(q)))] = LETTERS[int(rng.integers(0, len(LETTERS)))]
        if trial % 5 == 0:
            cut = int(rng.integers(0, len(q)))
            q[cut:] = q[cut:].lower()
        qb = bytes(q)
        sq = _seq(qb)
        for off in list(range(max(0, len(qb) - 4), len(qb))) + \
                [int(x) for x in rng.integers(0, len(qb), 10)]:
            pos = int(rank[start + off]) if rng.random() < 0.6 else int(rng.integers(0, len(T)))
            args = (C.byref(db), C.byref(sq), off, pos)
            assert L.sa_lcp(*args) == R.sa_lcp(*args), (trial, off, pos)
            out = []
            for lib in (L, R):
                sc, ln = C.c_int(-7), C.c_int(-7)
                rc = lib.sa_gapless(*args, C.byref(opt), C.byref(sc), C.byref(ln))
                out.append((rc, sc.value, ln.value))
            assert out[0] == out[1], (trial, off, pos, out)
            gapless += out[1][2] >= 0
            checked += 1
    assert checked > 500 and gapless > 100


def test_ungapped_rawscore_and_length(libs, opt):
    L, R = libs
    rng = np.random.default_rng(9)
    for n in [0, 1, 2, 7, 40, 333]:
        a = _letters(rng, n, bytes(range(1, 128)))
        b = _letters(rng, n)
        for k in range(n + 1):
            assert L.ungapped_rawscore(a, b, k, C.byref(opt)) == \
                R.ungapped_rawscore(a, b, k, C.byref(opt))
            assert L.ungapped_length(a, b, k, C.byref(opt)) == \
                R.ungapped_length(a, b, k, C.byref(opt))
    assert L.ungapped_length(b"A", b"A", 4_000_000_000, C.byref(opt)) == \
        R.ungapped_length(b"A", b"A", 4_000_000_000, C.byref(opt))


def _matrices(n1, n2):
    cells = (n1 + 1) * (n2 + 1)
    return ([np.full(cells, -12345, np.int32) for _ in range(3)],
            [np.full(cells, 999, np.uint16) for _ in range(3)])


def _ptrs(ints, trs):
    return ([a.ctypes.data_as(PI) for a in ints], [t.ctypes.data_as(PU16) for t in trs])


HIT_FIELDS = ["rawscore", "send", "qend", "sstart", "qstart", "mismatch", "gapopen", "length",
              "identity", "bitscore", "evalue", "is_aligned", "effective_length"]


def test_gotoh_matrices_trace_and_hit(libs, opt):
    """_initialise and _forward_pass fill the same E, F, H and trace
    matrices; _backtrack and gotoh write the same hit fields, for related
    and unrelated pairs under several gap settings (gap_open below
    gap_extend and a zero extension among them)."""
    L, R = libs
    rng = np.random.default_rng(13)
    from topaz_amd.matrices import INTERNAL_ORDER, MATRIX_ORDER
    done = 0
    for trial in range(60):
        go, ge = [(11, 1), (10, 2), (3, 3), (2, 5), (6, 0)][trial % 5]
        opt.gap_open, opt.gap_extend = go, ge
        q = synth.random_residues(rng, int(rng.integers(1, 160)))
        if trial % 3:
            r_, o_, l_ = synth.mutate(rng, q, np.array([0]), np.array([len(q)]),
                                      np.array([rng.uniform(0.3, 0.9)]))
            s = r_[:l_[0]]
        else:
            s = synth.random_residues(rng, int(rng.integers(1, 160)))
        s1, s2 = [bytes(ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[i]) for i in x)
                  for x in (q, s)]
        if trial % 4 == 1:
            s1 = s1.lower()
        q1, q2 = _seq(s1), _seq(s2)
        n1, n2 = len(s1), len(s2)
        got = []
        for lib in (L, R):
            ints, trs = _matrices(n1, n2)
            pi, pt = _ptrs(ints, trs)
            lib._initialise(C.byref(q1), C.byref(q2), *pi, *pt, C.byref(opt))
            r1 = np.array([matrix_row(c) for c in s1], np.int8)
            r2 = np.array([matrix_row(c) for c in s2], np.int8)
            hit = HitT()
            end = lib._forward_pass(r1.ctypes.data_as(P8), r2.ctypes.data_as(P8), n1, n2,
                                    pi[0], pt[0], pi[1], pt[1], pi[2], pt[2], C.byref(hit),
                                    C.byref(opt))
            got.append((end, [a.copy() for a in ints + trs], hit.rawscore, hit.send, hit.qend))
        assert got[0][0] == got[1][0] and got[0][2:] == got[1][2:], trial
        for a, b in zip(got[0][1], got[1][1]):
            assert np.array_equal(a, b), trial
        if got[1][2] == 0:
            continue                      # a trace from a zero cell aborts (tested below)
        hits = []
        for lib in (L, R):
            ints = [a.copy() for a in got[1][1][:3]]
            trs = [t.copy() for t in got[1][1][3:]]
            pi, pt = _ptrs(ints, trs)
            hit = HitT()
            lib._backtrack(*pt, got[1][0], n1 + 1, C.byref(hit))
            h2 = HitT()
            assert lib.gotoh(C.byref(q1), C.byref(q2), C.byref(h2), C.byref(opt)) == 0
            hits.append(([getattr(hit, f) for f in ("sstart", "qstart", "mismatch", "gapopen",
                                                    "length", "identity")],
                         [getattr(h2, f) for f in HIT_FIELDS]))
        assert hits[0] == hits[1], (trial, hits)
        done += 1
    opt.gap_open, opt.gap_extend = 11, 1
    assert done > 40


def matrix_row(c):
    """aa_table2 (lexicographical.h:19-30) of letter byte c."""
    from topaz_amd.matrices import INTERNAL_ORDER, MATRIX_ORDER
    u = c - 32 if 97 <= c <= 121 else c
    if 65 <= u <= 89:
        return MATRIX_ORDER.index(INTERNAL_ORDER[u - 65])
    return MATRIX_ORDER.index("X")


_ABORT = r"""
import ctypes as C, sys
sys.path.insert(0, {root!r})
from tests.refcheck import HitT, OptionsT, SeqT, matrix
L = C.CDLL({lib!r})
mat = matrix()
m = (C.c_int8 * 625)(*mat.tolist())
o = OptionsT(); o.substitution = C.cast(m, C.c_void_p)
o.gap_open, o.gap_extend = 11, 1
o.lambda_, o.k, o.H = 0.267, 0.041, 0.140
o.db_size, o.db_proteins = 1000, 10
def seq(b):
    s = SeqT(); s.seq_buf.buf = b; s.seq_buf.len = len(b); s.seq_len = len(b); return s
a, b = seq(b"WWWW"), seq(b"PPPP")
L.gotoh(C.byref(a), C.byref(b), C.byref(HitT()), C.byref(o))
print("returned")
"""


def test_gotoh_without_a_positive_cell_aborts_as_the_reference_does(libs):
    """No cell scores above 0: the reference's backtrack meets code 0 at its
    first step and aborts with its error line; so does the library."""
    from pathlib import Path
    root = str(Path(__file__).resolve().parents[1])
    res = []
    for lib in (str(_lib.LIB_PATH), str(REF_SO)):
        p = subprocess.run([sys.executable, "-c", _ABORT.format(root=root, lib=lib)],
                           capture_output=True, text=True, timeout=120)
        res.append((p.returncode, "backtrack failed, illegal condition" in p.stderr,
                    "returned" in p.stdout))
    assert res[0] == res[1] == (-6, True, False), res


def test_printdp(libs, capfd):
    L, R = libs
    m32 = np.arange(-5, 7, dtype=np.int32)
    m8 = np.arange(250, 256, dtype=np.uint8)
    outs = []
    for lib in (L, R):
        lib.printdp32(m32.ctypes.data_as(PI), 4, 3)
        lib.printdp8(m8.ctypes.data_as(C.POINTER(C.c_uint8)), 3, 2)
        C.CDLL(None).fflush(None)
        outs.append(capfd.readouterr().out)
    assert outs[0] == outs[1] and outs[0].count("\n") == 5
