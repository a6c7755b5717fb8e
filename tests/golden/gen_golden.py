"""Write the golden fixtures of tests/golden/ from the reference itself.

The reference's sources are compiled by `make -C oracle ref` into
oracle/_ref/libtopazref.so (never committed); this script drives its own C
entry points -- ssw_init/ssw_align/ssw_align_stats (ssw.c:799-978), sa_hsp
(search.c:99-190) and align_using_ssw_profile (ssw_wrapper.c:157-317) -- on
seeded synthetic inputs and stores inputs and outputs as .npz data.

    make -C oracle ref && python tests/golden/gen_golden.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from tests.refcheck import REF_SO, Reference, matrix  # noqa: E402
from topaz_amd import synth  # noqa: E402
from topaz_amd.matrices import INTERNAL_ORDER, MATRIX_ORDER, NAMES  # noqa: E402

OUT = Path(__file__).resolve().parent


def _flat(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    off[1:] = np.cumsum(lens[:-1])
    return np.concatenate(seqs).astype(np.int8), off, lens


def sw_cases(rng):
    """(query, subject, matrix name, gapO, gapE) cases of every kind."""
    cases = []

    def rnd(n, alpha=25):
        return rng.integers(0, alpha, n).astype(np.int8)

    def hom(q, ident, indel=0.03):
        r, o, l = synth.mutate(rng, q, np.array([0]), np.array([len(q)]),
                               np.array([ident]), indel)
        return r[:l[0]]

    for _ in range(300):          # random background pairs, BLOSUM62 11/1
        q = synth.random_residues(rng, int(rng.integers(1, 400)))
        s = synth.random_residues(rng, int(rng.integers(1, 400)))
        cases.append((q, s, "BLOSUM62", 11, 1))
    for _ in range(300):          # homologs, 25-95 % identity
        q = synth.random_residues(rng, int(rng.integers(20, 500)))
        cases.append((q, hom(q, rng.uniform(0.25, 0.95)), "BLOSUM62", 11, 1))
    for _ in range(150):          # tiny: every row a stripe start
        q = rnd(int(rng.integers(1, 17)))
        s = rnd(int(rng.integers(1, 40)))
        cases.append((q, s, "BLOSUM62", 11, 1))
    for _ in range(60):           # near the 250/251 byte -> word switch
        q = synth.random_residues(rng, int(rng.integers(40, 120)))
        cases.append((q, hom(q, rng.uniform(0.55, 0.9), 0.0), "BLOSUM62", 11, 1))
    for _ in range(20):           # W-rich: word path and 32767 overflow
        n = int(rng.integers(2000, 3600))
        q = np.full(n, MATRIX_ORDER.index("W"), np.int8)
        q[rng.random(n) < 0.02] = 0
        s = q.copy() if rng.random() < 0.5 else q[: int(rng.integers(n // 2, n))].copy()
        cases.append((q, s, "BLOSUM62", 11, 1))
    for _ in range(200):          # other matrices and gap penalties
        name = NAMES[int(rng.integers(0, len(NAMES)))]
        go, ge = [(15, 2), (10, 1), (9, 1), (5, 3), (3, 1), (2, 2), (6, 6)][int(rng.integers(0, 7))]
        q = synth.random_residues(rng, int(rng.integers(1, 300)))
        s = hom(q, rng.uniform(0.3, 0.9)) if rng.random() < 0.6 else \
            synth.random_residues(rng, int(rng.integers(1, 300)))
        cases.append((q, s, name, go, ge))
    return cases


def gen_sw(ref):
    rng = np.random.default_rng(20240611)
    cases = sw_cases(rng)
    mats = {n: matrix(n) for n in NAMES}
    res = np.zeros((len(cases), 3), np.int32)
    null = np.zeros(len(cases), bool)
    stats = np.zeros((len(cases), 6), np.int32)
    ident = np.zeros(len(cases), np.float64)
    has_stats = np.zeros(len(cases), bool)
    for k, (q, s, name, go, ge) in enumerate(cases):
        r = ref.ssw_align(q, s, go, ge, mats[name])
        if r is None:
            null[k] = True
            continue
        res[k] = r
        if r[0] > 0 and r[1] >= 0:
            st = ref.ssw_align_stats(q, s, r[2], r[1], r[0], go, ge, mats[name])
            stats[k] = st[:6]
            ident[k] = st[6]
            has_stats[k] = True
    qf = _flat([c[0] for c in cases])
    sf = _flat([c[1] for c in cases])
    np.savez_compressed(
        OUT / "sw_pairs.npz",
        q_res=qf[0], q_off=qf[1], q_len=qf[2], s_res=sf[0], s_off=sf[1], s_len=sf[2],
        matrix=np.array([NAMES.index(c[2]) for c in cases], np.int8),
        gap_open=np.array([c[3] for c in cases], np.int32),
        gap_extend=np.array([c[4] for c in cases], np.int32),
        score=res[:, 0], ref_end=res[:, 1], read_end=res[:, 2], is_null=null,
        stats_rc=stats[:, 0], read_begin=stats[:, 1], ref_begin=stats[:, 2],
        length=stats[:, 3], gapopen=stats[:, 4], mismatch=stats[:, 5],
        identity=ident, has_stats=has_stats)
    print("sw_pairs.npz: %d cases, %d NULL, %d with stats" %
          (len(cases), null.sum(), has_stats.sum()))


from tests.refcheck import BufferT, DbT, HitT, OptionsT, SeqT  # noqa: E402


def internal_text(rng, lengths, lower_frac=0.0):
    """Concatenated internal-alphabet proteins, each followed by '\\n'."""
    parts = []
    for L in lengths:
        idx = synth.random_residues(rng, L)
        s = "".join(chr(ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[i])) for i in idx)
        parts.append(s + "\n")
    return "".join(parts).encode()


def gen_hsp(lib):
    rng = np.random.default_rng(77)
    lengths = synth.lognormal_lengths(rng, 60, 120, 100, 5, 600)
    T = internal_text(rng, lengths)
    Tsize = len(T)
    SA = np.arange(Tsize, dtype=np.int64)       # identity "suffix array": SA[p] = p
    db = DbT()
    db.T = T
    db.Tsize = Tsize
    db.SA = SA.ctypes.data_as(C.POINTER(C.c_int64))
    db.SAsize = Tsize * 8
    mat = matrix("BLOSUM62")
    mat_c = (C.c_int8 * 625)(*mat.tolist())
    opt = OptionsT()
    opt.substitution = C.cast(mat_c, C.c_void_p)
    lib.sa_hsp.restype = C.c_int
    lib.sa_hsp.argtypes = [C.POINTER(DbT), C.POINTER(SeqT), C.c_int, C.c_int64,
                           C.POINTER(OptionsT), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    qlens = synth.lognormal_lengths(rng, 40, 150, 120, 1, 700)
    queries = []
    for L in qlens:
        if rng.random() < 0.5:          # a query copied from T with mutations
            start = int(rng.integers(0, max(1, Tsize - L)))
            q = bytearray(T[start:start + L].replace(b"\n", b"A"))
            for k in range(len(q)):
                if rng.random() < 0.3:
                    q[k] = ord("A") + int(rng.integers(0, 25))
            if rng.random() < 0.2 and len(q) > 3:    # lowercase and odd bytes
                q[int(rng.integers(0, len(q)))] = ord("c")
                q[int(rng.integers(0, len(q)))] = ord("*")
            queries.append(bytes(q))
        else:
            queries.append(internal_text(rng, [L])[:-1])
    rows = []
    for qi, qb in enumerate(queries):
        seq = SeqT()
        seq.seq_buf.buf = qb
        seq.seq_buf.len = len(qb)
        seq.seq_len = len(qb)
        for _ in range(60):
            o = int(rng.integers(0, len(qb)))
            kind = rng.random()
            if kind < 0.5:      # seed on a matching first residue
                cand = np.nonzero(np.frombuffer(T, np.uint8) == qb[o])[0]
                if len(cand) == 0:
                    continue
                p = int(cand[int(rng.integers(0, len(cand)))])
            elif kind < 0.6:    # near the text ends
                p = int(rng.choice([0, 1, 2, Tsize - 1, Tsize - 2]))
            else:
                p = int(rng.integers(0, Tsize))
            sc = C.c_int(-999999)
            ln = C.c_int(0)
            lib.sa_hsp(C.byref(db), C.byref(seq), o, p, C.byref(opt), C.byref(sc), C.byref(ln))
            rows.append((qi, o, p, sc.value if ln.value != -1 else -999999, ln.value))
    qf = np.frombuffer(b"".join(queries), np.uint8)
    qoff = np.cumsum([0] + [len(q) for q in queries[:-1]]).astype(np.int64)
    r = np.array(rows, np.int64)
    np.savez_compressed(OUT / "gapless.npz", T=np.frombuffer(T, np.uint8), q=qf, q_off=qoff,
                        q_len=np.array([len(q) for q in queries], np.int32),
                        seed_query=r[:, 0].astype(np.int32), seed_offset=r[:, 1].astype(np.int32),
                        seed_pos=r[:, 2], score=r[:, 3].astype(np.int32),
                        length=r[:, 4].astype(np.int32))
    print("gapless.npz: %d seeds over Tsize %d, %d hits" % (len(rows), Tsize, (r[:, 4] != -1).sum()))


def gen_evalue(lib, ref):
    """align_using_ssw_profile on pairs: rawscore, bitscore, evalue, ends."""
    rng = np.random.default_rng(5)
    lib.align_using_ssw_profile.argtypes = [C.c_void_p, C.POINTER(SeqT), C.POINTER(HitT),
                                            C.POINTER(OptionsT)]
    lib.align_using_ssw_profile.restype = None
    mat = matrix("BLOSUM62")
    mat_c = (C.c_int8 * 625)(*mat.tolist())
    rows = []
    qs, ss = [], []
    for k in range(200):
        m = int(rng.integers(5, 400))
        q = synth.random_residues(rng, m)
        if k % 2:
            r_, o_, l_ = synth.mutate(rng, q, np.array([0]), np.array([m]),
                                      np.array([rng.uniform(0.3, 0.95)]))
            s = r_[:l_[0]]
        else:
            s = synth.random_residues(rng, int(rng.integers(5, 400)))
        db_size = int(rng.choice([10 ** 6, 35 * 10 ** 6, 3 * 10 ** 9]))
        db_prot = int(db_size // int(rng.integers(200, 500)))
        gl = int(rng.choice([10, 39, 40, 41, 100]))
        opt = OptionsT()
        opt.substitution = C.cast(mat_c, C.c_void_p)
        opt.gap_open, opt.gap_extend = 11, 1
        opt.lambda_, opt.k, opt.H = 0.267, 0.041, 0.140
        opt.db_size, opt.db_proteins = db_size, db_prot
        opt.evalue = 1e-3
        prof, keep = ref.profile(q, mat)
        # subject as internal letters: align_using_ssw_profile maps with aa_table2
        sb = "".join(chr(ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[i])) for i in s).encode()
        seq = SeqT()
        seq.seq_buf.buf = sb
        seq.seq_buf.len = len(sb)
        seq.seq_len = len(sb)
        h = HitT()
        h.gapless_score = gl
        h.rawscore = -1
        lib.align_using_ssw_profile(C.cast(prof, C.c_void_p), C.byref(seq), C.byref(h), C.byref(opt))
        ref.L.init_destroy(prof)
        rows.append((m, db_size, db_prot, gl, h.rawscore, h.send, h.qend, h.bitscore, h.evalue,
                     h.effective_length, h.is_aligned))
        qs.append(q)
        ss.append(s)
    r = rows
    np.savez_compressed(
        OUT / "evalue.npz",
        qlen=np.array([x[0] for x in r], np.int32), db_size=np.array([x[1] for x in r], np.uint64),
        db_proteins=np.array([x[2] for x in r], np.uint64),
        gapless=np.array([x[3] for x in r], np.int32), rawscore=np.array([x[4] for x in r], np.int32),
        send=np.array([x[5] for x in r], np.int32), qend=np.array([x[6] for x in r], np.int32),
        bitscore=np.array([x[7] for x in r]), evalue=np.array([x[8] for x in r]),
        effective_length=np.array([x[9] for x in r], np.int32),
        q_res=_flat(qs)[0], q_off=_flat(qs)[1], q_len=_flat(qs)[2],
        s_res=_flat(ss)[0], s_off=_flat(ss)[1], s_len=_flat(ss)[2])
    print("evalue.npz: %d cases" % len(r))


def main():
    if not REF_SO.exists():
        raise SystemExit("build the reference checker first: make -C oracle ref")
    ref = Reference()
    gen_sw(ref)
    gen_hsp(ref.L)
    gen_evalue(ref.L, ref)


if __name__ == "__main__":
    main()
