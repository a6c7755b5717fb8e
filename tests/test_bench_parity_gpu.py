"""Parity at the benchmark's own sizes: the bench's 2M-pair Smith-Waterman
batch and its two 50M-seed gapless sets over the 1e9-byte text, run on the
GPU exactly as bench.py launches them, with a sample of the results checked
against the oracle (and not only against other GPU runs).

* Forward SW: every pair with a side longer than 2,500 residues (up to
  5,000: 157 chunks of 32 rows through the workspace rotation) plus 20,000
  pairs drawn at random; (score, ref_end, read_end) equal to orc_ssw_align
  (ssw.c:835-897), 32767 where it gives NULL.
* Gapless: 100,000 seeds of each set, the first and last seeds of the set
  included; (score, length) equal to
  orc_sa_hsp_batch (search.c:99-190), seeds in the last 2 MB of the text
  among them.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bench
from tests.refcheck import Oracle
from topaz_amd import matrix, synth
from topaz_amd.gapless import GaplessBatch, Text
from topaz_amd.sw import OVERFLOW, SwBatch

pytestmark = pytest.mark.gpu


def _workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def test_sw_bench_workload_against_oracle():
    w = bench.make_workload(1.0)
    q, s = w["q"], w["s"]
    pq, ps = w["pair_query"], w["pair_subject"]
    b = SwBatch(q, s, pq, ps, matrix(), 11, 1)
    sc, re_, rd = b.run().results()
    b.close()
    ql = q[2][pq]
    sl = s[2][ps]
    long_pairs = np.nonzero(np.maximum(ql, sl) > 2500)[0]
    rng = np.random.default_rng(5)
    sample = np.union1d(long_pairs, rng.choice(len(pq), 20000, replace=False))
    assert len(long_pairs) > 100 and sl[long_pairs].max() == 5000
    orc = Oracle()
    mat = matrix()

    def one(p):
        qq = q[0][q[1][pq[p]]:q[1][pq[p]] + q[2][pq[p]]]
        ss = s[0][s[1][ps[p]]:s[1][ps[p]] + s[2][ps[p]]]
        return orc.ssw_align(qq, ss, 11, 1, mat)

    with ThreadPoolExecutor(_workers()) as ex:
        want = list(ex.map(one, sample.tolist(), chunksize=64))
    bad = []
    for p, r in zip(sample.tolist(), want):
        got = (int(sc[p]), int(re_[p]), int(rd[p]))
        if r is None:
            if got[0] != OVERFLOW:
                bad.append((p, got, r))
        elif got != tuple(r):
            bad.append((p, got, r))
    assert not bad, bad[:10]
    # the word path (score + bias >= 255) is among the compared pairs
    assert sum(1 for r in want if r is not None and r[0] + 4 >= 255) > 100


def _check_gapless(orc, text_bytes, queries, seeds, sc, ln, pick):
    osc, oln = orc.sa_hsp_batch(text_bytes, queries, seeds[pick], matrix())
    assert np.array_equal(ln[pick], oln)
    assert np.array_equal(sc[pick], osc)


def test_gapless_bench_sets_against_oracle():
    rng = np.random.default_rng(2025)        # bench.py's gapless workload, as generated there
    n_seeds = int(5e7)
    text, queries, uniform, nl = synth.internal_text_and_seeds(rng, int(1e9), 10000, n_seeds,
                                                               homolog_frac=0.05)
    matched = synth.matching_seeds(rng, queries, text, n_seeds, uniform[:int(0.05 * n_seeds)])
    uniform = uniform[np.argsort(uniform["query"], kind="stable")]
    matched = matched[np.argsort(matched["query"], kind="stable")]
    text_bytes = text.tobytes()
    dev_text = Text(text_bytes)
    orc = Oracle()
    pick_rng = np.random.default_rng(9)
    for seeds in (uniform, matched):
        b = GaplessBatch(dev_text, queries, seeds, matrix()).run()
        sc, ln = b.results()
        b.close()
        high = np.nonzero(seeds["position"] > len(text_bytes) - 2000000)[0][:1000]
        pick = np.unique(np.concatenate([
            pick_rng.choice(len(seeds), 100000, replace=False), high,
            np.arange(50), np.arange(len(seeds) - 50, len(seeds))]))
        parts = np.array_split(pick, _workers())
        with ThreadPoolExecutor(len(parts)) as ex:
            list(ex.map(lambda p: _check_gapless(orc, text_bytes, queries, seeds, sc, ln, p),
                        parts))
        assert (ln[pick] >= 5).sum() > 1000
    dev_text.close()
