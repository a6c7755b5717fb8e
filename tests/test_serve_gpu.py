"""Serving loop: each request's results are the same alone and in the mix
(bit-exact), equal to the batch path's (itself checked against the oracle),
and admission keeps the admitted bytes within the memory budget."""
import threading

import numpy as np
import pytest

from topaz_amd import matrix, synth
from topaz_amd.serve import Server
from topaz_amd.sw import OVERFLOW, SwBatch, _concat
from tests.refcheck import Oracle

pytestmark = pytest.mark.gpu


def database(rng, n=400):
    lens = synth.lognormal_lengths(rng, n, 250, 200, 1, 1500)
    return [synth.random_residues(rng, int(k)).astype(np.int8) for k in lens]


def requests(rng, db, n_req, max_q=30, max_pairs=60):
    out = []
    for _ in range(n_req):
        nq = int(rng.integers(1, max_q + 1))
        qs = []
        pq, ps = [], []
        for k in range(nq):
            # half of the queries are mutated copies of database subjects
            if rng.random() < 0.5:
                src = db[int(rng.integers(0, len(db)))]
                res, _, _ = synth.mutate(rng, src.astype(np.int64), np.array([0]),
                                         np.array([len(src)]), rng.uniform(0.3, 0.9, 1))
                qs.append(res.astype(np.int8))
            else:
                qs.append(synth.random_residues(rng, int(rng.integers(1, 700))).astype(np.int8))
            npair = int(rng.integers(0, max_pairs + 1))
            pq += [k] * npair
            ps += list(rng.integers(0, len(db), npair))
        out.append((qs, np.array(pq, np.int32), np.array(ps, np.int32)))
    return out


def batch_results(db, req):
    qs, pq, ps = req
    b = SwBatch(qs, db, pq, ps, matrix(), 11, 1)
    r = b.run().results()
    b.close()
    return r


@pytest.mark.parametrize("waves", [4, 8])
def test_alone_equals_mix_and_batch(waves):
    rng = np.random.default_rng(41)
    db = database(rng)
    reqs = requests(rng, db, 14)
    srv = Server(db, matrix(), 11, 1, max_inflight=2, waves_per_group=waves)
    alone = []
    for q, pq, ps in reqs:
        alone.append(srv.wait(srv.submit(q, pq, ps))[:3])
    st0 = srv.stats()
    assert st0["launches"] == len(reqs)
    # the mix: three client threads submitting at once
    tickets = [None] * len(reqs)

    def client(ids):
        for i in ids:
            tickets[i] = srv.submit(*reqs[i])

    th = [threading.Thread(target=client, args=(list(range(j, len(reqs), 3)),)) for j in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mixed = [srv.wait(t)[:3] for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["requests_done"] == 2 * len(reqs)
    assert st["launches"] - st0["launches"] < len(reqs)       # requests were merged
    assert st["max_requests_per_launch"] > 1
    for i, (a, m) in enumerate(zip(alone, mixed)):
        for x, y in zip(a, m):
            assert np.array_equal(x, y), i
    # equal to the batch path, and spot-checked against the oracle
    o = Oracle()
    mat = matrix()
    for i in (0, 5, 9):
        for x, y in zip(alone[i], batch_results(db, reqs[i])):
            assert np.array_equal(x, y), i
        qs, pq, ps = reqs[i]
        for p in range(0, len(pq), 7):
            r = o.ssw_align(qs[pq[p]], db[ps[p]], 11, 1, mat)
            exp = (OVERFLOW, None, None) if r is None else r
            got = (int(alone[i][0][p]), int(alone[i][1][p]), int(alone[i][2][p]))
            assert got == exp if exp[0] != OVERFLOW else got[0] == OVERFLOW


def test_memory_budget_admission():
    rng = np.random.default_rng(42)
    db = database(rng, 200)
    reqs = requests(rng, db, 10, max_q=20, max_pairs=80)
    sizes = []
    for qs, pq, ps in reqs:
        sizes.append(sum(len(q) for q in qs) + 12 * len(qs) + 72 * len(pq) + 256)
    budget = 2 * max(sizes) + 64
    srv = Server(db, matrix(), 11, 1, memory_budget=budget, max_inflight=3)
    tickets = [srv.submit(*r) for r in reqs]
    got = [srv.wait(t)[:3] for t in tickets]
    st = srv.stats()
    assert st["requests_done"] == len(reqs) and st["peak_bytes"] <= budget
    for i in (1, 4, 8):
        for x, y in zip(got[i], batch_results(db, reqs[i])):
            assert np.array_equal(x, y), i
    # a request larger than the whole budget is refused at submit
    big_q = [synth.random_residues(rng, 300).astype(np.int8) for _ in range(5)]
    n = budget // 40 + 10
    with pytest.raises(Exception, match="budget"):
        srv.submit(big_q, np.zeros(n, np.int32), np.zeros(n, np.int32))
    srv.close()


def test_empty_request():
    rng = np.random.default_rng(43)
    db = database(rng, 20)
    srv = Server(db, matrix(), 11, 1)
    t = srv.submit([np.zeros(0, np.int8)], np.zeros(0, np.int32), np.zeros(0, np.int32))
    sc, re_, rd, lat = srv.wait(t)
    assert len(sc) == 0 and lat >= 0
    srv.close()
