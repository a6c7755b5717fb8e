"""Batched gapless extension on the GPU against the reference's sa_hsp
(golden fixture) and the oracle, bit-exact."""
import numpy as np
import pytest

from tests.refcheck import Oracle
from topaz_amd import matrix, synth
from topaz_amd.gapless import SEED_DTYPE, GaplessBatch, Text
from topaz_amd.matrices import INTERNAL_ORDER, MATRIX_ORDER

pytestmark = pytest.mark.gpu
GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def test_golden_fixture():
    d = np.load(GOLD / "gapless.npz")
    T = d["T"].tobytes()
    qs = [d["q"][d["q_off"][k]:d["q_off"][k] + d["q_len"][k]].tobytes()
          for k in range(len(d["q_len"]))]
    seeds = np.zeros(len(d["seed_pos"]), SEED_DTYPE)
    seeds["query"] = d["seed_query"]
    seeds["offset"] = d["seed_offset"]
    seeds["position"] = d["seed_pos"]
    t = Text(T)
    b = GaplessBatch(t, qs, seeds, matrix()).run()
    sc, ln = b.results()
    assert np.array_equal(ln, d["length"])
    hit = d["length"] != -1
    assert np.array_equal(sc[hit], d["score"][hit])
    assert np.all(sc[~hit] == 0)


def _internal(idx):
    return bytes(ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[i]) for i in idx)


@pytest.mark.parametrize("use_sa", [False, True])
def test_random_seeds_against_oracle(use_sa):
    rng = np.random.default_rng(31)
    lens = synth.lognormal_lengths(rng, 400, 200, 150, 1, 900)
    T = b"".join(_internal(synth.random_residues(rng, L)) + b"\n" for L in lens)
    qs = []
    for L in synth.lognormal_lengths(rng, 50, 200, 150, 1, 700):
        if rng.random() < 0.5:
            st = int(rng.integers(0, len(T) - 1))
            q = bytearray(T[st:st + L].replace(b"\n", b"G"))
            for k in np.nonzero(rng.random(len(q)) < 0.2)[0]:
                q[k] = ord("A") + int(rng.integers(0, 25))
            for k in np.nonzero(rng.random(len(q)) < 0.01)[0]:
                q[k] = int(rng.choice([ord("a"), ord("*"), ord("Z"), ord("\n")]))
            qs.append(bytes(q) or b"A")
        else:
            qs.append(_internal(synth.random_residues(rng, L)))
    n = 20000
    seeds = np.zeros(n, SEED_DTYPE)
    seeds["query"] = rng.integers(0, len(qs), n)
    seeds["offset"] = [int(rng.integers(0, len(qs[q]))) for q in seeds["query"]]
    pos = rng.integers(0, len(T), n)
    # half the seeds on a matching first residue, some at the text ends
    Tn = np.frombuffer(T, np.uint8)
    for k in range(0, n, 2):
        c = qs[seeds["query"][k]][seeds["offset"][k]]
        cand = np.nonzero(Tn == c)[0]
        if len(cand):
            pos[k] = cand[int(rng.integers(0, len(cand)))]
    pos[:20] = [0, 1, 2, len(T) - 1, len(T) - 2] * 4
    sa = None
    if use_sa:
        sa = rng.permutation(len(T)).astype(np.int64)
        inv = np.empty_like(sa)
        inv[sa] = np.arange(len(T))
        seeds["position"] = inv[pos]
    else:
        seeds["position"] = pos
    t = Text(T, sa)
    b = GaplessBatch(t, qs, seeds, matrix()).run()
    sc, ln = b.results()
    o = Oracle()
    bad = 0
    for k in range(n):
        es, el = o.sa_hsp(T, int(pos[k]), qs[seeds["query"][k]], int(seeds["offset"][k]))
        if ln[k] != el or (el != -1 and sc[k] != es):
            bad += 1
    assert bad == 0
    assert (ln != -1).sum() > n // 3


def test_out_of_text_positions_are_not_seeds():
    T = b"ABCDE\nFGHIJ\n"
    seeds = np.zeros(4, SEED_DTYPE)
    seeds["position"] = [-5, len(T), len(T) + 100, 2]
    seeds["offset"] = [0, 0, 0, 2]
    b = GaplessBatch(Text(T), [b"ABCDE"], seeds, matrix()).run()
    sc, ln = b.results()
    assert list(ln[:3]) == [-1, -1, -1] and ln[3] != -1


@pytest.mark.parametrize("text_len", [60, 5000])
def test_seeds_at_the_text_start(text_len):
    """Left scans that end at T[1] (i < idx) and Q[1] (i < o): every text
    position below 140 against every query offset below 150, first residues
    matching; a text shorter than one window too."""
    rng = np.random.default_rng(55 + text_len)
    body = _internal(synth.random_residues(rng, text_len))
    T = bytearray(body)
    for k in rng.integers(5, len(T) - 1, len(T) // 300 + 1):
        T[k] = 10
    T = bytes(T) + b"\n"
    qs = [_internal(synth.random_residues(rng, int(L))) for L in (160, 200, 3, 1)]
    recs = []
    for idx in range(min(140, len(T))):
        for q, Q in enumerate(qs):
            for o in range(min(150, len(Q))):
                if Q[o] == T[idx]:
                    recs.append((q, o, idx))
    seeds = np.zeros(len(recs), SEED_DTYPE)
    seeds["query"] = [r[0] for r in recs]
    seeds["offset"] = [r[1] for r in recs]
    seeds["position"] = [r[2] for r in recs]
    assert len(seeds) > 200
    b = GaplessBatch(Text(T), qs, seeds, matrix()).run()
    sc, ln = b.results()
    o = Oracle()
    for k, (q, off, idx) in enumerate(recs):
        es, el = o.sa_hsp(T, idx, qs[q], off)
        assert (int(ln[k]), int(sc[k]) if el != -1 else None) == (el, es), (q, off, idx)
