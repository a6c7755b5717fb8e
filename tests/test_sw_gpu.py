"""Batched Smith-Waterman on the GPU against the oracle, bit-exact.

The oracle (oracle/liboracle.so) is the checker: for every pair the GPU's
(score, ref_end, read_end) must equal what ssw_align returns (ssw.c:835-897),
including the 255 -> word-path switch and the 32767 overflow.
"""
import numpy as np
import pytest

from topaz_amd import matrix, synth
from topaz_amd.sw import OVERFLOW, SCHED_PERSISTENT, SCHED_PLAIN, SwBatch
from tests.refcheck import Oracle

pytestmark = pytest.mark.gpu


def oracle_results(o, w, mat, go, ge):
    q, s = w["q"], w["s"]
    out = []
    for qi, si in zip(w["pair_query"], w["pair_subject"]):
        qq = q[0][q[1][qi]:q[1][qi] + q[2][qi]]
        ss = s[0][s[1][si]:s[1][si] + s[2][si]]
        r = o.ssw_align(qq, ss, go, ge, mat)
        out.append((OVERFLOW, None, None) if r is None else r)
    return out


def compare(w, mat=None, go=11, ge=1, schedule=SCHED_PERSISTENT, config=None):
    mat = matrix() if mat is None else mat
    o = Oracle()
    b = SwBatch(w["q"], w["s"], w["pair_query"], w["pair_subject"], mat, go, ge)
    if config:
        b.set_config(**config)
    sc, re_, rd = b.run(schedule).results()
    exp = oracle_results(o, w, mat, go, ge)
    bad = []
    for p, (es, er, ed) in enumerate(exp):
        if es == OVERFLOW:
            ok = sc[p] == OVERFLOW
        else:
            ok = (sc[p], re_[p], rd[p]) == (es, er, ed)
        if not ok:
            bad.append((p, (int(sc[p]), int(re_[p]), int(rd[p])), (es, er, ed)))
    b.close()
    assert not bad, "%d/%d mismatches, first: %s" % (len(bad), len(exp), bad[:5])
    return sc


def small_workload(seed, nq=12, per_q=40, mean=150, sd=100, hi=600, lo=1):
    rng = np.random.default_rng(seed)
    return synth.pair_workload(rng, nq, per_q, mean=mean, sd=sd, lo=lo, hi=hi)


@pytest.mark.parametrize("schedule", [SCHED_PERSISTENT, SCHED_PLAIN])
def test_random_and_homolog_pairs(schedule):
    sc = compare(small_workload(1), schedule=schedule)
    # the workload reaches both sides of the byte/word switch (255 - bias)
    assert (sc < 251).sum() > 100 and ((sc >= 251) & (sc < 32767)).sum() > 20


@pytest.mark.parametrize("schedule", [SCHED_PERSISTENT, SCHED_PLAIN])
def test_long_and_tiny_queries(schedule):
    rng = np.random.default_rng(2)
    w = synth.pair_workload(rng, 6, 30, mean=400, sd=600, lo=1, hi=1500)
    compare(w, schedule=schedule)


def test_plain_multi_pass_queries():
    """Queries of 2,049 .. 5,000 residues: the plain schedule's wavefront per
    pair takes several passes of 2,048 rows, handing each column's bottom row
    through the workspace."""
    rng = np.random.default_rng(12)
    w = synth.pair_workload(rng, 4, 24, mean=3200, sd=900, lo=2049, hi=5000)
    compare(w, schedule=SCHED_PLAIN)


def test_tiny_sequences_all_corners():
    # m <= 16: every row is a stripe start, all F->E corners are lost
    rng = np.random.default_rng(3)
    qs = [rng.integers(0, 25, int(rng.integers(1, 20))).astype(np.int8) for _ in range(20)]
    ss = [rng.integers(0, 25, int(rng.integers(1, 40))).astype(np.int8) for _ in range(40)]
    pq = np.repeat(np.arange(20), 40).astype(np.int32)
    ps = np.tile(np.arange(40), 20).astype(np.int32)
    from topaz_amd.sw import _concat
    w = dict(q=_concat(qs), s=_concat(ss), pair_query=pq, pair_subject=ps)
    compare(w)


@pytest.mark.parametrize("name,go,ge", [("BLOSUM45", 15, 2), ("PAM30", 9, 1),
                                         ("BLOSUM80", 10, 1), ("PAM250", 5, 3)])
def test_other_matrices_and_gaps(name, go, ge):
    compare(small_workload(4, nq=6, per_q=30), mat=matrix(name), go=go, ge=ge)


@pytest.mark.parametrize("schedule", [SCHED_PERSISTENT, SCHED_PLAIN])
def test_overflow_to_word_and_drop(schedule):
    # W-rich long identical pairs: byte overflows, word may reach 32767
    rng = np.random.default_rng(5)
    q1 = np.full(3100, 17, np.int8)          # W x 3100 -> self score 34100
    q2 = rng.integers(0, 25, 800).astype(np.int8)
    qs = [q1, q2]
    ss = [q1.copy(), q1[:2000].copy(), q2.copy(), q2[100:700].copy(),
          rng.integers(0, 25, 300).astype(np.int8)]
    from topaz_amd.sw import _concat
    pq = np.array([0, 0, 1, 1, 1, 0], np.int32)
    ps = np.array([0, 1, 2, 3, 4, 4], np.int32)
    w = dict(q=_concat(qs), s=_concat(ss), pair_query=pq, pair_subject=ps)
    sc = compare(w, schedule=schedule)
    assert sc[0] == OVERFLOW and sc[1] == 22000


GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("schedule", [SCHED_PERSISTENT, SCHED_PLAIN])
def test_golden_forward_every_gap_setting(schedule):
    """The golden fixture (the reference's own ssw_align, compiled from its
    sources): every (matrix, gap_open, gap_extend) group, gap_open <=
    gap_extend included, where the kernels follow the reference's lanes."""
    from topaz_amd.matrices import NAMES
    d = np.load(GOLD / "sw_pairs.npz")
    q = (d["q_res"], d["q_off"], d["q_len"])
    s = (d["s_res"], d["s_off"], d["s_len"])
    lanes = 0
    for mi, go, ge in sorted(set(zip(d["matrix"], d["gap_open"], d["gap_extend"]))):
        sel = np.nonzero((d["matrix"] == mi) & (d["gap_open"] == go) & (d["gap_extend"] == ge))[0]
        b = SwBatch(q, s, sel.astype(np.int32), sel.astype(np.int32), matrix(NAMES[mi]),
                    int(go), int(ge))
        sc, fe, re_ = b.run(schedule).results()
        b.close()
        null = d["is_null"][sel]
        assert np.array_equal(sc == OVERFLOW, null), (mi, go, ge)
        ok = ~null
        assert np.array_equal(sc[ok], d["score"][sel][ok]), (mi, go, ge)
        assert np.array_equal(fe[ok], d["ref_end"][sel][ok]), (mi, go, ge)
        assert np.array_equal(re_[ok], d["read_end"][sel][ok]), (mi, go, ge)
        lanes += len(sel) if go <= ge else 0
    assert lanes >= 50


@pytest.mark.parametrize("go,ge", [(2, 2), (6, 6), (1, 3), (0, 0), (5, 9), (0, 4)])
@pytest.mark.parametrize("schedule", [SCHED_PERSISTENT, SCHED_PLAIN])
def test_lane_gap_settings_against_oracle(go, ge, schedule):
    """gap_open <= gap_extend: the reference's lanes lose deletion-insertion
    corners and differ from textbook alignment; the lane kernel follows
    them (the oracle's lane emulation, pinned by the golden cases)."""
    from topaz_amd.sw import _concat
    rng = np.random.default_rng(100 + 10 * go + ge)
    w = synth.pair_workload(rng, 6, 25, mean=200, sd=180, lo=1, hi=900)
    compare(w, go=go, ge=ge, schedule=schedule)
    # tiny queries: every row a lane's first row
    qs = [rng.integers(0, 25, int(rng.integers(1, 20))).astype(np.int8) for _ in range(10)]
    ss = [rng.integers(0, 25, int(rng.integers(1, 60))).astype(np.int8) for _ in range(20)]
    w = dict(q=_concat(qs), s=_concat(ss), pair_query=np.repeat(np.arange(10), 20).astype(np.int32),
             pair_subject=np.tile(np.arange(20), 10).astype(np.int32))
    compare(w, go=go, ge=ge, schedule=schedule)


def test_lane_gaps_word_pass_and_drop():
    """gap_open <= gap_extend with scores past the byte pass (word pass) and
    past 32767 (dropped)."""
    from topaz_amd.sw import _concat
    rng = np.random.default_rng(7)
    q1 = np.full(3100, 17, np.int8)
    q2 = rng.integers(0, 25, 800).astype(np.int8)
    ss = [q1.copy(), q1[:2000].copy(), q2.copy(), q2[100:700].copy(),
          rng.integers(0, 25, 300).astype(np.int8)]
    pq = np.array([0, 0, 1, 1, 1, 0], np.int32)
    ps = np.array([0, 1, 2, 3, 4, 4], np.int32)
    w = dict(q=_concat([q1, q2]), s=_concat(ss), pair_query=pq, pair_subject=ps)
    sc = compare(w, go=3, ge=3)
    assert sc[0] == OVERFLOW and 255 < sc[1] < OVERFLOW


def test_empty_batch_and_empty_subjects():
    from topaz_amd.sw import _concat
    qs = [np.array([1, 2, 3], np.int8)]
    ss = [np.zeros(0, np.int8), np.array([1, 2, 3], np.int8)]
    w = dict(q=_concat(qs), s=_concat(ss), pair_query=np.array([0, 0], np.int32),
             pair_subject=np.array([0, 1], np.int32))
    compare(w)
    b = SwBatch(_concat(qs), _concat(ss), np.zeros(0, np.int32), np.zeros(0, np.int32),
                matrix())
    assert len(b.run().results()[0]) == 0
    b.close()


@pytest.mark.parametrize("big_fraction", [0.0, 1e30])
def test_group_and_wave_paths(big_fraction):
    """All tasks worked by whole workgroups (chunks interleaved over four
    wavefronts), then all tasks by single wavefronts: same results."""
    rng = np.random.default_rng(7)
    w = synth.pair_workload(rng, 6, 150, mean=400, sd=350, lo=1, hi=2500)
    compare(w, config=dict(big_fraction=big_fraction))


@pytest.mark.parametrize("schedule", [SCHED_PERSISTENT, SCHED_PLAIN])
def test_packed_tasks_mix_queries(schedule):
    """Wavefronts holding pairs of up to four queries: queries of equal and
    unequal chunk counts (lengths 1..100, some empty), uneven pair counts,
    subjects from empty to long, subjects shared between queries."""
    rng = np.random.default_rng(31)
    nq = 48
    ql = rng.integers(1, 100, nq)
    ql[rng.integers(0, nq, 3)] = 0
    qs = [rng.integers(0, 25, int(n)).astype(np.int8) for n in ql]
    ns = 300
    sl = np.minimum(rng.lognormal(4.5, 1.0, ns).astype(int), 1200)
    sl[:4] = 0
    ss = [rng.integers(0, 25, int(n)).astype(np.int8) for n in sl]
    counts = rng.integers(1, 200, nq)
    pq = np.repeat(np.arange(nq), counts).astype(np.int32)
    ps = rng.integers(0, ns, len(pq)).astype(np.int32)
    from topaz_amd.sw import _concat
    w = dict(q=_concat(qs), s=_concat(ss), pair_query=pq, pair_subject=ps)
    compare(w, schedule=schedule)


CONFIGS = [dict(rows_per_chunk=16, groups_per_cu=2, queries_per_task=4),
           dict(rows_per_chunk=16, groups_per_cu=1, queries_per_task=1, big_fraction=1e30),
           dict(rows_per_chunk=32, groups_per_cu=1, queries_per_task=2, big_fraction=0.01),
           dict(rows_per_chunk=16, groups_per_cu=2, queries_per_task=2, big_fraction=0.01),
           dict(rows_per_chunk=32, groups_per_cu=1, queries_per_task=4, waves_per_group=8),
           dict(rows_per_chunk=32, groups_per_cu=1, queries_per_task=1, waves_per_group=8,
                big_fraction=0.0),
           dict(rows_per_chunk=16, groups_per_cu=1, queries_per_task=4, waves_per_group=8,
                big_fraction=0.01)]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_configurations_match_oracle(cfg):
    """Every tile shape / schedule knob gives the reference's results."""
    w = small_workload(8, nq=20, per_q=50, mean=200, sd=200, hi=900)
    mat = matrix()
    b = SwBatch(w["q"], w["s"], w["pair_query"], w["pair_subject"], mat, 11, 1)
    b.set_config(**cfg)
    got = {s: b.run(s).results() for s in (SCHED_PERSISTENT, SCHED_PLAIN)}
    b.close()
    exp = oracle_results(Oracle(), w, mat, 11, 1)
    for sched, (sc, re_, rd) in got.items():
        bad = [p for p, (es, er, ed) in enumerate(exp)
               if not (sc[p] == OVERFLOW if es == OVERFLOW else (sc[p], re_[p], rd[p]) == (es, er, ed))]
        assert not bad, (cfg, sched, bad[:5])


def test_graph_replays_match_eager():
    """Captured launches (counter cleared by a kernel, no memset node) give
    the eager results on every replay, with eager work in between."""
    w = small_workload(9, nq=16, per_q=60, mean=180, sd=150, hi=800)
    b = SwBatch(w["q"], w["s"], w["pair_query"], w["pair_subject"], matrix(), 11, 1)
    eager = {s: b.run(s).results() for s in (SCHED_PERSISTENT, SCHED_PLAIN)}
    b.set_config(use_graph=1)
    other = SwBatch(w["q"], w["s"], w["pair_query"][::-1].copy(), w["pair_subject"][::-1].copy(),
                    matrix(), 11, 1)
    for rep in range(4):
        for s in (SCHED_PERSISTENT, SCHED_PLAIN):
            got = b.run(s).results()
            for x, y in zip(got, eager[s]):
                assert np.array_equal(x, y), (rep, s)
            other.run(SCHED_PERSISTENT)          # eager work between replays
    assert b.config().use_graph == 1
    b.close()
    other.close()


def test_autotune_keeps_results():
    w = small_workload(10, nq=12, per_q=64, mean=200, sd=150, hi=700)
    b = SwBatch(w["q"], w["s"], w["pair_query"], w["pair_subject"], matrix(), 11, 1)
    before = b.run(SCHED_PERSISTENT).results()
    cfg, ms, n = b.autotune()
    assert n >= 10 and ms > 0
    assert cfg.rows_per_chunk in (16, 32) and 1 <= cfg.queries_per_task <= 4
    assert b.config().as_dict() == cfg.as_dict()
    after = b.run(SCHED_PERSISTENT).results()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    b.close()
