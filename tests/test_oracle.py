"""The oracle, pinned: against the reference's own outputs (golden fixtures
written by tests/golden/gen_golden.py from the compiled reference), against
brute force on tiny inputs, and against closed forms.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from tests.refcheck import REF_SO, Oracle, Reference, matrix
from topaz_amd import synth
from topaz_amd.matrices import NAMES, aa_table2

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _seq(res, off, lens, k):
    return res[off[k]:off[k] + lens[k]]


def test_alphabet_map_matches_table():
    o = Oracle()
    t = aa_table2()
    for c in range(128):
        assert o.L.orc_aa_table2(c) == t[c], chr(c)
    # spot checks of lexicographical.h:26-32: 'A'->0, 'X'->24 ('*'), 'V'->23 ('X')
    assert t[ord("A")] == 0 and t[ord("X")] == 24 and t[ord("V")] == 23 and t[ord("Z")] == 23


def test_sw_golden_scores_and_ends(orc):
    d = np.load(GOLD / "sw_pairs.npz")
    bad = []
    for k in range(len(d["q_len"])):
        q = _seq(d["q_res"], d["q_off"], d["q_len"], k)
        s = _seq(d["s_res"], d["s_off"], d["s_len"], k)
        mat = matrix(NAMES[d["matrix"][k]])
        r = orc.ssw_align(q, s, int(d["gap_open"][k]), int(d["gap_extend"][k]), mat)
        if d["is_null"][k]:
            ok = r is None
        else:
            ok = r == (d["score"][k], d["ref_end"][k], d["read_end"][k])
        if not ok:
            bad.append((k, r))
    assert not bad, bad[:5]
    assert d["is_null"].sum() >= 1 and (d["score"] > 251).sum() > 100


def test_sw_golden_traceback_stats(orc):
    d = np.load(GOLD / "sw_pairs.npz")
    bad = []
    for k in np.nonzero(d["has_stats"])[0]:
        q = _seq(d["q_res"], d["q_off"], d["q_len"], k)
        s = _seq(d["s_res"], d["s_off"], d["s_len"], k)
        mat = matrix(NAMES[d["matrix"][k]])
        st = orc.ssw_align_stats(q, s, int(d["read_end"][k]), int(d["ref_end"][k]),
                                 int(d["score"][k]), int(d["gap_open"][k]),
                                 int(d["gap_extend"][k]), mat)
        exp = (d["stats_rc"][k], d["read_begin"][k], d["ref_begin"][k], d["length"][k],
               d["gapopen"][k], d["mismatch"][k])
        same_ident = st[6] == d["identity"][k] or (math.isnan(st[6]) and math.isnan(d["identity"][k]))
        if tuple(st[:6]) != tuple(int(x) for x in exp) or not same_ident:
            bad.append((k, st, exp, d["identity"][k]))
    assert not bad, bad[:3]


def test_gapless_golden(orc):
    d = np.load(GOLD / "gapless.npz")
    T = d["T"].tobytes()
    bad = []
    for k in range(len(d["seed_pos"])):
        qi = d["seed_query"][k]
        Q = d["q"][d["q_off"][qi]:d["q_off"][qi] + d["q_len"][qi]].tobytes()
        sc, ln = orc.sa_hsp(T, int(d["seed_pos"][k]), Q, int(d["seed_offset"][k]))
        exp_sc = None if d["length"][k] == -1 else int(d["score"][k])
        if (sc, ln) != (exp_sc, int(d["length"][k])):
            bad.append((k, (sc, ln), (exp_sc, int(d["length"][k]))))
    assert not bad, bad[:5]
    assert (d["length"] != -1).sum() > 500


def test_evalue_golden(orc):
    d = np.load(GOLD / "evalue.npz")
    n_aligned = 0
    for k in range(len(d["qlen"])):
        if d["gapless"][k] < 40:            # ssw_wrapper.c:171-176: no alignment
            assert d["bitscore"][k] == 0 and d["evalue"][k] == 1000
            continue
        n_aligned += 1
        raw = int(d["rawscore"][k])
        if raw == 32767:
            continue
        b, e, el = orc.evalue(raw, int(d["qlen"][k]), int(d["db_size"][k]),
                              int(d["db_proteins"][k]))
        assert b == d["bitscore"][k] and e == d["evalue"][k] and el == d["effective_length"][k], k
        q = _seq(d["q_res"], d["q_off"], d["q_len"], k)
        s = _seq(d["s_res"], d["s_off"], d["s_len"], k)
        assert orc.ssw_align(q, s) == (raw, d["send"][k], d["qend"][k])
    assert n_aligned > 50


def test_evalue_closed_form(orc):
    # bits = (lambda*S - ln K)/ln 2 ; E = m' n' e^(-(lambda*S - ln K))
    lam, K, H = 0.267, 0.041, 0.140
    for raw, m, D, N in [(50, 300, 35_000_000, 100_000), (500, 1000, 10 ** 9, 10 ** 6),
                         (12, 10, 1000, 10)]:
        b, e, el = orc.evalue(raw, m, D, N)
        corr = math.log(K * m * D) / H
        cq, cdb = max(m - corr, 1.0), max(D - corr * N, 1.0)
        nats = lam * raw - math.log(K)
        assert b == pytest.approx(nats / math.log(2), rel=1e-15)
        ev = cq * cdb * math.exp(-nats)
        assert e == pytest.approx(ev if ev > 1e-180 else 0.0, rel=1e-12)


def gotoh_textbook(q, s, mat, go, ge):
    """Plain O(mn) local alignment, gap of length k costing go + (k-1) ge."""
    m, n = len(q), len(s)
    W = mat.reshape(25, 25)
    H = np.zeros((n + 1, m + 1), np.int64)
    E = np.zeros((n + 1, m + 1), np.int64)
    F = np.zeros((n + 1, m + 1), np.int64)
    best = 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            E[i, j] = max(E[i - 1, j] - ge, H[i - 1, j] - go)
            F[i, j] = max(F[i, j - 1] - ge, H[i, j - 1] - go)
            H[i, j] = max(0, H[i - 1, j - 1] + W[s[i - 1], q[j - 1]], E[i, j], F[i, j])
            best = max(best, H[i, j])
    return best


def test_sw_against_textbook_gotoh(orc):
    """Brute-force textbook Gotoh (numpy, written independently of the C
    code) has the same best score as the reference's lane recurrence."""
    rng = np.random.default_rng(11)
    mat = matrix()
    for _ in range(150):
        m, n = int(rng.integers(1, 40)), int(rng.integers(1, 40))
        q = rng.integers(0, 25, m).astype(np.int8)
        s = rng.integers(0, 25, n).astype(np.int8) if rng.random() < 0.5 else \
            np.concatenate([q[: n // 2], rng.integers(0, 25, 3).astype(np.int8), q[n // 2:]])
        go, ge = [(11, 1), (3, 1), (2, 0), (5, 2)][int(rng.integers(0, 4))]
        assert orc.ssw_align(q, s, go, ge)[0] == gotoh_textbook(q, s, mat, go, ge)


def test_lane_recurrence_is_textbook_when_gapO_gt_gapE(orc):
    """For gapO > gapE the reference's lane recurrence (lazy-F included, E
    taken before lazy-F) gives the textbook H in every cell: a dropped
    vertical-then-horizontal gap corner always has an equal-cost
    horizontal-then-vertical twin.  Score and both end positions agree on
    random, homologous and gap-rich pairs, for both lane counts."""
    from topaz_amd import synth
    rng = np.random.default_rng(21)
    for _ in range(400):
        name = NAMES[int(rng.integers(0, len(NAMES)))]
        mat = matrix(name)
        go = int(rng.integers(1, 16))
        ge = int(rng.integers(0, go))
        m = int(rng.integers(1, 300))
        q = synth.random_residues(rng, m)
        if rng.random() < 0.6:
            r_, o_, l_ = synth.mutate(rng, q, np.array([0]), np.array([m]),
                                      np.array([rng.uniform(0.2, 1.0)]), rng.uniform(0, 0.15))
            s = r_[:l_[0]]
        else:
            s = synth.random_residues(rng, int(rng.integers(1, 300)))
        for P in (16, 8):
            assert orc.sw_scan(q, s, P, go, ge, mat) == orc.sw_textbook(q, s, P, go, ge, mat)


def test_lane_recurrence_differs_when_gapO_le_gapE(orc):
    """With gapO == gapE the word path's lazy-F loop stops while a lane is
    still raising H, so it differs from textbook on some pairs: the kernels'
    precondition gapO > gapE is necessary, and a mistake that took the
    textbook shortcut everywhere would be caught."""
    from topaz_amd import synth
    rng = np.random.default_rng(22)
    diffs = 0
    for _ in range(300):
        m = int(rng.integers(50, 400))
        q = synth.random_residues(rng, m)
        r_, o_, l_ = synth.mutate(rng, q, np.array([0]), np.array([m]),
                                  np.array([rng.uniform(0.6, 1.0)]), 0.1)
        s = r_[:l_[0]]
        if orc.sw_scan(q, s, 8, 2, 2) != orc.sw_textbook(q, s, 8, 2, 2):
            diffs += 1
    assert diffs > 0


def test_closed_form_equals_lane_emulation_when_gapO_gt_gapE(orc):
    """The closed form the kernels implement equals the reference's lane
    emulation whenever gapO > gapE, for both lane counts."""
    rng = np.random.default_rng(13)
    for _ in range(300):
        m, n = int(rng.integers(1, 200)), int(rng.integers(1, 200))
        q = rng.integers(0, 25, m).astype(np.int8)
        s = q[: n].copy() if rng.random() < 0.5 else rng.integers(0, 25, n).astype(np.int8)
        go = int(rng.integers(2, 16))
        ge = int(rng.integers(0, go))
        for P in (16, 8):
            assert orc.sw_scan(q, s, P, go, ge) == orc.sw_scan(q, s, P, go, ge, closed_form=True)


def test_gapless_brute_force(orc):
    """sa_hsp against a direct prefix-sum formulation on random text."""
    rng = np.random.default_rng(14)
    mat = matrix()
    t2 = aa_table2()
    letters = np.frombuffer(b"ABCDEFGHIJKLMNOPQRSTUVWXY", np.uint8)
    for _ in range(300):
        Tb = bytearray(letters[rng.integers(0, 25, int(rng.integers(1, 300)))].tobytes())
        for k in rng.integers(0, len(Tb), 3):
            Tb[k] = 10                     # protein delimiters
        Qb = bytearray(letters[rng.integers(0, 25, int(rng.integers(1, 200)))].tobytes())
        T, Q = bytes(Tb), bytes(Qb)
        idx = int(rng.integers(0, len(T)))
        o = int(rng.integers(0, len(Q)))
        if rng.random() < 0.7:
            Qb[o] = T[idx] if T[idx] != 10 else Qb[o]
            Q = bytes(Qb)
        sc, ln = orc.sa_hsp(T, idx, Q, o)
        if T[idx] != Q[o]:
            assert ln == -1
            continue

        def side(step, start, limit):
            best, arg, s = 0, 0, 0
            for i in range(start, limit):
                c = T[idx + step * i]
                if c == 10:
                    break
                s += int(mat[int(t2[c]) * 25 + int(t2[Q[o + step * i]])])
                if s >= best:
                    best, arg = s, i
            return best, arg

        br = side(1, 0, min(len(T) - idx, len(Q) - o, 128))
        bl = side(-1, 1, min(idx, o, 128))
        assert (sc, ln) == (br[0] + bl[0], br[1] + bl[1])


@pytest.mark.skipif(not REF_SO.exists(), reason="reference checker not built here")
def test_oracle_vs_compiled_reference_random():
    o, r = Oracle(), Reference()
    rng = np.random.default_rng(99)
    for _ in range(300):
        m, n = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        q = rng.integers(0, 25, m).astype(np.int8)
        s = rng.integers(0, 25, n).astype(np.int8)
        go, ge = [(11, 1), (2, 2), (5, 3), (10, 2)][int(rng.integers(0, 4))]
        assert o.ssw_align(q, s, go, ge) == r.ssw_align(q, s, go, ge)


def test_gapless_batch_and_scan_bytes(orc):
    """The batch loop equals per-seed calls, and the scan-byte accounting the
    bench uses equals the bytes the rule reads, counted by direct walking."""
    from topaz_amd import synth
    rng = np.random.default_rng(21)
    text, queries, seeds, nl = synth.internal_text_and_seeds(rng, 60_000, 40, 3000,
                                                             homolog_frac=0.1)
    seeds = np.concatenate([seeds, synth.matching_seeds(rng, queries, text, 3000, seeds[:300])])
    T = text.tobytes()
    sc, ln = orc.sa_hsp_batch(T, queries, seeds)
    ql = np.array([len(q) for q in queries])
    right, left = synth.gapless_scan_bytes(seeds, ql, text, nl)
    r2, l2 = synth.gapless_scan_bytes(seeds, ql, text, nl, synth.newline_rank(text))
    assert np.array_equal(right, r2) and np.array_equal(left, l2)
    assert (ln != -1).sum() > 3000
    for k in range(len(seeds)):
        q, o, idx = int(seeds["query"][k]), int(seeds["offset"][k]), int(seeds["position"][k])
        s1, l1 = orc.sa_hsp(T, idx, queries[q], o)
        assert (l1, s1 if l1 != -1 else 0) == (ln[k], sc[k])
        if l1 == -1:
            continue
        m = len(queries[q])
        r = 0
        for i in range(min(len(T) - idx, m - o, 128)):
            r += 1
            if T[idx + i] == 10:
                break
        lft = 0
        for i in range(1, min(idx, o, 128)):
            lft += 1
            if T[idx - i] == 10:
                break
        assert (right[k], left[k]) == (r, lft), k


def _gapless_cases(rng, n_text=60, n_q=30):
    letters = np.frombuffer(b"ABCDEFGHIJKLMNOPQRSTUVWXY", np.uint8)
    T = b"".join(letters[rng.integers(0, 25, int(L))].tobytes() + b"\n"
                 for L in rng.integers(1, 200, n_text))
    qs = [letters[rng.integers(0, 25, int(L))].tobytes() for L in rng.integers(1, 150, n_q)]
    cases = []
    Tn = np.frombuffer(T, np.uint8)
    for _ in range(3000):
        q = int(rng.integers(0, n_q))
        o = int(rng.integers(0, len(qs[q])))
        if rng.random() < 0.7:
            cand = np.nonzero(Tn == qs[q][o])[0]
            p = int(cand[int(rng.integers(0, len(cand)))]) if len(cand) else 0
        else:
            p = int(rng.integers(0, len(T)))
        cases.append((q, o, p))
    return T, qs, cases


def test_sa_gapless_brute_force(orc):
    """The dead variant sa_gapless (search.c:48-97): right-only prefix
    maximum, at least 3 positions, first residues equal."""
    rng = np.random.default_rng(61)
    T, qs, cases = _gapless_cases(rng)
    mat = matrix()
    t2 = aa_table2()
    seen = 0
    for q, o, p in cases:
        Q = qs[q]
        sc = C.c_int32(0)
        ln = orc.L.orc_sa_gapless(T, len(T), p, Q, len(Q), o,
                                  mat.ctypes.data_as(C.POINTER(C.c_int8)), C.byref(sc))
        lim = min(len(T) - p, len(Q) - o)
        if lim < 3 or T[p] != Q[o]:
            assert ln == -1
            continue
        best = arg = s = 0
        for i in range(lim):
            if T[p + i] == 10:
                break
            s += int(mat[int(t2[T[p + i]]) * 25 + int(t2[Q[o + i]])])
            if s >= best:
                best, arg = s, i
        assert (ln, sc.value) == (arg, best)
        seen += 1
    assert seen > 1000


@pytest.mark.skipif(not REF_SO.exists(), reason="reference checker not built here")
def test_sa_gapless_vs_compiled_reference(orc):
    from tests.refcheck import DbT, OptionsT, SeqT
    lib = C.CDLL(str(REF_SO))
    lib.sa_gapless.restype = C.c_int
    lib.sa_gapless.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.c_int64,
                               C.POINTER(OptionsT), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rng = np.random.default_rng(62)
    T, qs, cases = _gapless_cases(rng)
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
    for q, o, p in cases:
        seq = SeqT()
        seq.seq_buf.buf = qs[q]
        seq.seq_buf.len = len(qs[q])
        seq.seq_len = len(qs[q])
        rs, rl = C.c_int(-7), C.c_int(0)
        lib.sa_gapless(C.byref(db), C.byref(seq), o, p, C.byref(opt), C.byref(rs), C.byref(rl))
        sc = C.c_int32(-7)
        ln = orc.L.orc_sa_gapless(T, len(T), p, qs[q], len(qs[q]), o,
                                  mat.ctypes.data_as(C.POINTER(C.c_int8)), C.byref(sc))
        assert ln == rl.value and (ln == -1 or sc.value == rs.value), (q, o, p)


@pytest.mark.skipif(not REF_SO.exists(), reason="reference checker not built here")
def test_stats_vs_compiled_reference_homologs(orc):
    """ssw_align_stats (reverse pass + banded_sw with its buffers shared
    across band passes) against the compiled reference on homologous pairs
    with indels, embedded fragments and random pairs."""
    from topaz_amd import synth
    r = Reference()
    rng = np.random.default_rng(63)
    n = 0
    for it in range(3000):
        m = int(rng.integers(5, 300))
        q = synth.random_residues(rng, m).astype(np.int8)
        kind = it % 3
        if kind < 2:
            res, _, _ = synth.mutate(rng, q.astype(np.int64), np.array([0]), np.array([m]),
                                     rng.uniform(0.2, 0.95, 1), indel_rate=0.06)
            s = res.astype(np.int8)
            if kind == 1:
                s = np.concatenate([synth.random_residues(rng, int(rng.integers(0, 200))), s,
                                    synth.random_residues(rng, int(rng.integers(0, 200)))]
                                   ).astype(np.int8)
        else:
            s = synth.random_residues(rng, int(rng.integers(5, 300))).astype(np.int8)
        go, ge = [(11, 1), (10, 2), (5, 3)][it % 3]
        a = r.ssw_align(q, s, go, ge)
        if a is None or a[0] <= 0:
            continue
        sc, re_, rd = a
        x = orc.ssw_align_stats(q, s, rd, re_, sc, go, ge)
        y = r.ssw_align_stats(q, s, rd, re_, sc, go, ge)
        assert x[:6] == y[:6] and (x[6] == y[6] or (math.isnan(x[6]) and math.isnan(y[6]))), it
        n += 1
    assert n > 2500


def _ref_lib():
    if not REF_SO.exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    return Reference().L


def test_suffix_array_against_reference_and_brute_force():
    """orc_suffix_array equals the reference's compiled divsufsort64 on
    random bytes, a two-letter text, one letter x 3,000, a repeated block and
    a protein text with '\\n' separators, and brute-force sorting of the
    suffixes where that is cheap."""
    R = _ref_lib()
    R.divsufsort64.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
    R.divsufsort64.restype = C.c_int32
    o = Oracle()
    rng = np.random.default_rng(91)
    block = rng.integers(0, 256, 700, dtype=np.uint8).tobytes()
    text, _, _, _ = synth.internal_text_and_seeds(rng, 200_000, 4, 4)
    cases = [b"", b"a", b"ba", b"banana", rng.integers(0, 256, 50_000, dtype=np.uint8).tobytes(),
             rng.integers(65, 67, 20_000, dtype=np.uint8).tobytes(), b"Q" * 3000,
             block * 12, text.tobytes()]
    for t in cases:
        got = o.suffix_array(t)
        ref = np.zeros(max(len(t), 1), np.int64)
        assert R.divsufsort64(t, ref.ctypes.data_as(C.POINTER(C.c_int64)), len(t)) == 0
        assert np.array_equal(got, ref[:len(t)]), len(t)
        if len(t) <= 3000:
            assert got.tolist() == sorted(range(len(t)), key=lambda i: t[i:])


def test_sa_search_against_reference():
    """orc_sa_search equals the reference's compiled sa_search (search.c:
    237-361) at every offset of queries that occur in the text, that do not,
    and that run along a repeat, with the index and the last match length
    of absent suffixes."""
    from tests.refcheck import DbT, SeqT
    R = _ref_lib()
    o = Oracle()
    rng = np.random.default_rng(92)
    text, queries, _, _ = synth.internal_text_and_seeds(rng, 100_000, 3, 3)
    t = text.tobytes() + b"A" * 30 + b"\n"
    sa = o.suffix_array(t)
    db = DbT()
    db.T = t
    db.Tsize = len(t)
    db.SA = sa.ctypes.data_as(C.POINTER(C.c_int64))
    P64 = C.POINTER(C.c_int64)
    R.sa_search.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int64, C.c_int64, P64, P64, P64]
    R.sa_search.restype = C.c_int
    for qb in [t[5000:5300], queries[0], b"A" * 40 + t[7000:7100], queries[1][:50] + t[100:200]]:
        q = SeqT()
        q.seq_buf.buf = qb
        q.seq_buf.len = len(qb)
        q.seq_len = len(qb)
        for off in range(len(qb) + 1):
            v = [C.c_int64(-9) for _ in range(3)]
            rc = R.sa_search(C.byref(db), C.byref(q), off, 0, *[C.byref(x) for x in v])
            assert o.sa_search(t, sa, qb[off:]) == (rc,) + tuple(x.value for x in v), off
