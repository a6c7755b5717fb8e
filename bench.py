"""Benchmark of the candidate-alignment hot path on one MI355X.

Stage S (the headline): 2,000,000 query-subject pairs -- 2,000 queries x
1,000 subjects each, both sides lognormal length (mean 350, sd 250) clipped
to [30, 5000], residues from Robinson-Robinson background frequencies, half
of every query's subjects planted homologs at 25-95 % identity with 2 %
indels; BLOSUM62, gap 11/1.  Inputs are uploaded once; a step is one full
pass of the batched kernel over all pairs (scores and end positions of every
pair).  The plain one-launch-per-query schedule is timed beside it.

Stage G (the "gapless" object): a 1e9-residue newline-separated text
resident in HBM, 10,000 queries and 50,000,000 seeds (query, offset, text
position), twice: drawn uniformly over query x text positions (5 % on the
diagonals of planted homologs), and drawn with agreeing first residues, as
suffix-array seeds are; in both sets the seeds of a query come together, as
the reference generates them query by query.  A step is one pass of the gapless kernel over all
seeds; the rate is the bytes the reference rule reads per seed plus the
seed record and the result, over the kernel time.

The default run (--leg all) also carries the other legs in the same line:
serving alignment requests, traceback statistics, the suffix array, and
end-to-end `topaz search` at 100,000 proteins x 1,000 queries and at
10,000,000 proteins x 100,000 queries (default and -F, the reference CPU
binary on all host cores over a 1,000-query sample, hit lists compared).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--leg all|sw|gapless|...]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "queries/s + SW GCUPS, batched vs plain GPU (bit-exact hits); p50/p99 latency"
VALU_PEAK_TOPS = 78.6      # 256 CU x 64 lanes x 2 (packed int16) x 2.4 GHz
OPS_PER_CELL = 10          # integer ops per DP cell in the reference inner loop


def measured_traffic(key):
    """HBM bytes per launch of a kernel at its bench workload, as measured
    with rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy of
    known size) and recorded in profiles/hbm_traffic.json (the runs are
    described in profiles/hbm_traffic.md); None if absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def make_workload(scale, seed=2024):
    from topaz_amd import synth
    rng = np.random.default_rng(seed)
    nq = max(1, int(round(2000 * scale)))
    return synth.pair_workload(rng, nq, 1000, homolog_frac=0.5, ident_range=(0.25, 0.95),
                               mean=350.0, sd=250.0, lo=30, hi=5000)


def cells_of(w):
    q, s = w["q"], w["s"]
    return int(np.sum(q[2][w["pair_query"]].astype(np.int64) *
                      s[2][w["pair_subject"]].astype(np.int64)))


CPU_BENCH = os.path.join(ROOT, "oracle", "_ref", "cpu_bench")


def _cores():
    return len(os.sched_getaffinity(0))


def cpu_baseline(w, seconds=12.0, n_queries=1000):
    """The reference's own compiled ssw_align (oracle/_ref/cpu_bench, a C
    harness over the reference's objects) on all host cores, one profile per
    query as query.c:194-197 builds it, over the pairs of a random sample of
    the workload's queries, for about `seconds`."""
    import subprocess
    import tempfile
    q, s = w["q"], w["s"]
    pq, ps = w["pair_query"], w["pair_subject"]
    rng = np.random.default_rng(7)
    pick_q = rng.permutation(len(q[2]))[:n_queries]
    order = np.argsort(pq, kind="stable")
    starts = np.searchsorted(pq[order], np.arange(len(q[2]) + 1))
    pairs = np.concatenate([order[starts[k]:starts[k + 1]] for k in pick_q])
    uq = pick_q
    us = np.unique(ps[pairs])
    q_map = {int(k): i for i, k in enumerate(uq)}
    s_map = np.full(len(s[2]), -1, np.int64)
    s_map[us] = np.arange(len(us))
    ql = q[2][uq].astype(np.int32)
    sl = s[2][us].astype(np.int32)
    qo = np.concatenate([[0], np.cumsum(ql[:-1], dtype=np.int64)])
    so = np.concatenate([[0], np.cumsum(sl[:-1], dtype=np.int64)])
    qr = np.concatenate([q[0][q[1][k]:q[1][k] + q[2][k]] for k in uq]).astype(np.int8)
    sr = np.concatenate([s[0][s[1][k]:s[1][k] + s[2][k]] for k in us]).astype(np.int8)
    npq = np.array([q_map[int(k)] for k in pq[pairs]], np.int32)
    nps = s_map[ps[pairs]].astype(np.int32)
    cores = _cores()
    with tempfile.NamedTemporaryFile(suffix=".pairs") as f:
        f.write(np.array([len(ql), len(sl), len(npq), len(qr), len(sr)], np.int64).tobytes())
        for a in (qo, so, ql, sl, npq, nps, qr, sr):
            f.write(np.ascontiguousarray(a).tobytes())
        f.flush()
        out = subprocess.run([CPU_BENCH, "sw", f.name, str(cores), str(seconds)],
                             capture_output=True, text=True, check=True).stdout.split()
    cells, n_pairs, n_done, dt = int(out[0]), int(out[1]), int(out[2]), float(out[3])
    return {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": cores,
            "kind": "reference",
            "sample": "%d pairs of %d queries (%.3g cells) of the same workload, %.1f s: the "
                      "reference's ssw.c (ssw_init once per query, ssw_align per pair) in a C "
                      "pthread harness, %d threads" % (n_pairs, n_done, cells, dt, cores)}


HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
SEED_BYTES, RESULT_BYTES = 16, 8


def gapless_cpu_baseline(text, queries, seeds, per_seed_bytes, seconds=12.0, max_seeds=40_000_000):
    """The reference's own compiled sa_hsp (search.c:99-190, in the C harness
    oracle/_ref/cpu_bench) on all host cores over the first `max_seeds`
    seeds, going round them again, for about `seconds`; the rate counts the
    same algorithmic bytes as the GPU's."""
    import subprocess
    import tempfile
    n = min(len(seeds), max_seeds)
    ql = np.array([len(x) for x in queries], np.int32)
    qo = np.concatenate([[0], np.cumsum(ql[:-1], dtype=np.int64)])
    cores = _cores()
    with tempfile.NamedTemporaryFile(suffix=".seeds") as f:
        f.write(np.array([len(text), len(queries), int(ql.sum()), n], np.int64).tobytes())
        f.write(np.ascontiguousarray(text).tobytes())
        f.write(qo.tobytes())
        f.write(ql.tobytes())
        f.write(b"".join(queries))
        f.write(np.ascontiguousarray(seeds["query"][:n], np.int32).tobytes())
        f.write(np.ascontiguousarray(seeds["offset"][:n], np.int32).tobytes())
        f.write(np.ascontiguousarray(seeds["position"][:n], np.int64).tobytes())
        f.flush()
        out = subprocess.run([CPU_BENCH, "gapless", f.name, str(cores), str(seconds)],
                             capture_output=True, text=True, check=True).stdout.split()
    done, dt = int(out[0]), float(out[1])
    rounds, rest = divmod(done, n)
    total = rounds * int(per_seed_bytes[:n].sum()) + int(per_seed_bytes[:rest].sum())
    return {"value": round(total / dt / 1e9, 3), "unit": "GB/s", "cores": cores,
            "kind": "reference",
            "sample": "%d seed extensions over the first %d seeds of the first-residue-matched "
                      "set, %.1f s: the reference's sa_hsp in a C pthread harness, %d threads"
                      % (done, n, dt, cores)}


def gapless_leg(args):
    from topaz_amd import matrix, synth
    from topaz_amd.gapless import GaplessBatch, Text
    rng = np.random.default_rng(2025)
    t_res = int(1e9 * args.gapless_scale)
    n_seeds = int(5e7 * args.gapless_scale)
    t0 = time.perf_counter()
    text, queries, uniform, nl = synth.internal_text_and_seeds(rng, t_res, 10000, n_seeds,
                                                               homolog_frac=0.05)
    matched = synth.matching_seeds(rng, queries, text, n_seeds, uniform[:int(0.05 * n_seeds)])
    # seeds arrive grouped by query, as query_find_suffixes and the suffix
    # walk produce them one query at a time (query.c:104-131, suffix.c:40-65)
    uniform = uniform[np.argsort(uniform["query"], kind="stable")]
    matched = matched[np.argsort(matched["query"], kind="stable")]
    ql = np.array([len(q) for q in queries], np.int64)
    print("[bench] gapless workload: %d text bytes, %d seeds x 2, generated in %.1f s"
          % (len(text), n_seeds, time.perf_counter() - t0), file=sys.stderr, flush=True)
    text_bytes = text.tobytes()
    dev_text = Text(text_bytes)
    rank = synth.newline_rank(text)
    out = {}
    for name, seeds in (("uniform", uniform), ("matched", matched)):
        b = GaplessBatch(dev_text, queries, seeds, matrix())
        for _ in range(args.warmup):
            b.run()
        times, kms = [], []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            b.run()
            times.append(time.perf_counter() - t1)
            kms.append(b.kernel_ms())
        # results are read after the timed steps (a readback between runs
        # lets the clocks drop), then one more run must reproduce them
        sc, ln = b.results()
        b.run()
        ref = b.results()
        assert np.array_equal(sc, ref[0]) and np.array_equal(ln, ref[1])
        b.close()
        right, left = synth.gapless_scan_bytes(seeds, ql, text, nl, rank)
        per_seed = np.where(ln == -1, 1, right + left) + SEED_BYTES + RESULT_BYTES
        total = int(per_seed.sum())
        kernel_s = float(np.mean(kms)) / 1e3
        gbs = total / kernel_s / 1e9
        out[name] = {"seeds": int(n_seeds), "passing": int((ln != -1).sum()),
                     "hsp_ge_5": int((ln >= 5).sum()),
                     "bytes_per_launch": total, "kernel_ms": round(kernel_s * 1e3, 3),
                     "ms_per_step": round(float(np.mean(times)) * 1e3, 3),
                     "seeds_per_s": round(n_seeds / kernel_s, 1),
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4),
                                  "traffic": measured_traffic("gapless_%s_50M" % name)
                                  if args.gapless_scale == 1.0 else None}}
        print("[bench] gapless %s: kernel %.3f ms, %.1f GB/s"
              % (name, kernel_s * 1e3, gbs), file=sys.stderr, flush=True)
        if name == "matched" and not args.no_cpu:
            out["cpu_baseline"] = gapless_cpu_baseline(text, queries, seeds, per_seed)
    dev_text.close()
    # the memory system's rate for this access pattern: the same number of
    # random window pairs read with no compute (topaz_amd/probe_window_gather)
    import subprocess
    probe = os.path.join(ROOT, "topaz_amd", "probe_window_gather")
    if os.path.exists(probe):
        r = subprocess.run([probe, str(len(text)), str(n_seeds)], capture_output=True, text=True,
                           timeout=300)
        if r.returncode == 0:
            g = json.loads(r.stdout)
            g["matched_kernel_vs_gather"] = round(g["ms_two_windows"] / out["matched"]["kernel_ms"], 3)
            out["window_gather"] = g
            print("[bench] gapless: window-pair gather alone %.3f ms (matched kernel at %.2f of it)"
                  % (g["ms_two_windows"], g["matched_kernel_vs_gather"]), file=sys.stderr, flush=True)
    out["workload"] = ("gapless HSP extension: %d-byte text (lognormal proteins, mean 350, "
                       "clipped [30,5000]), 10000 queries, %d seeds per set grouped by query; "
                       "bytes = text bytes the rule reads + 16 B seed + 8 B result"
                       % (len(text), n_seeds))
    return out


def sw_leg(args):
    from topaz_amd import matrix
    from topaz_amd.sw import SCHED_PERSISTENT, SCHED_PLAIN, SwBatch

    t_gen = time.perf_counter()
    w = make_workload(args.pairs_scale)
    n_pairs = len(w["pair_query"])
    n_queries = len(w["q"][2])
    cells = cells_of(w)
    print("[bench] workload: %d pairs, %d queries, %.4g cells, generated in %.1f s"
          % (n_pairs, n_queries, cells, time.perf_counter() - t_gen), file=sys.stderr, flush=True)

    b = SwBatch(w["q"], w["s"], w["pair_query"], w["pair_subject"], matrix(), 11, 1)
    print("[bench] uploaded", file=sys.stderr, flush=True)
    tuning = None
    if not args.no_tune:
        t0 = time.perf_counter()
        cfg, tuned_ms, n_tried = b.autotune()
        tuning = {"config": cfg.as_dict(), "candidates": n_tried,
                  "best_kernel_ms": round(tuned_ms, 3),
                  "seconds": round(time.perf_counter() - t0, 1)}
        print("[bench] autotune: %s (%d candidates, %.1f s)" % (cfg.as_dict(), n_tried,
              tuning["seconds"]), file=sys.stderr, flush=True)
    b.set_config(use_graph=1)
    for _ in range(args.warmup):
        b.run(SCHED_PERSISTENT)
    times, kms = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        b.run(SCHED_PERSISTENT)
        times.append(time.perf_counter() - t0)
        kms.append(b.kernel_ms())
    # results are read after the timed steps (a readback between runs lets
    # the clocks drop, and the next launch pays for it), then one more run
    # must reproduce them bit for bit
    sc, re_, rd = b.results()
    b.run(SCHED_PERSISTENT)
    ref_sc, ref_re, ref_rd = b.results()
    assert np.array_equal(sc, ref_sc) and np.array_equal(re_, ref_re) and np.array_equal(rd, ref_rd)
    t_step = float(np.sum(times)) / args.steps
    kernel_s = float(np.mean(kms)) / 1e3
    print("[bench] persistent: %.2f ms/step (kernel %.2f ms; per step %s)"
          % (t_step * 1e3, kernel_s * 1e3, " ".join("%.2f" % k for k in kms)),
          file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        b.set_config(use_graph=0)
        b.run(SCHED_PLAIN)
        t0 = time.perf_counter()
        b.run(SCHED_PLAIN)
        t_plain = time.perf_counter() - t0
        psc, pre, prd = b.results()
        assert np.array_equal(psc, sc) and np.array_equal(pre, re_) and np.array_equal(prd, rd)
        p_kernel_s = b.kernel_ms() / 1e3
        plain = {"schedule": "one launch per query, one wavefront per pair (systolic over query "
                             "rows, int32)",
                 "ms_per_step": round(t_plain * 1e3, 3), "gcups": round(cells / t_plain / 1e9, 2),
                 "queries_per_s": round(n_queries / t_plain, 2),
                 "roofline_frac": round(cells * OPS_PER_CELL / p_kernel_s / 1e12 / VALU_PEAK_TOPS,
                                        4)}
        print("[bench] plain: %.2f ms/step" % (t_plain * 1e3), file=sys.stderr, flush=True)
    b.close()

    gcups = cells / t_step / 1e9
    kernel_tops = cells * OPS_PER_CELL / kernel_s / 1e12
    out = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic",
        "config": {"workload": "batched SW kernel: %d query-subject pairs (%d queries x 1000 "
                               "subjects, lognormal mean 350 sd 250 clipped [30,5000], 50%% "
                               "planted homologs at 25-95%% identity), BLOSUM62 11/1"
                               % (n_pairs, n_queries),
                   "pairs": n_pairs, "queries": n_queries, "cells": cells,
                   "schedule": "persistent work queue, length-binned tasks of up to 4 "
                               "queries, autotuned, replayed from a hipGraph"},
        "tuning": tuning,
        "queries_per_s": round(n_queries / t_step, 2),
        # pairs whose byte (16-lane) pass saturates and the reference rescores
        # with 8 word lanes (score + bias >= 255), and those it drops (32767)
        "word_rescored_pairs": int(((sc - int(matrix().min()) >= 255) & (sc < 32767)).sum()),
        "null_pairs": int((sc >= 32767).sum()),
        "plain": plain,
        "speedup_vs_plain": round(plain["ms_per_step"] / (t_step * 1e3), 2) if plain else None,
        "roofline": {"bound": "valu", "achieved": round(kernel_tops, 2), "peak": VALU_PEAK_TOPS,
                     "unit": "Tint16op/s", "frac": round(kernel_tops / VALU_PEAK_TOPS, 4),
                     "traffic": measured_traffic("sw_forward_2Mpairs")
                     if args.pairs_scale == 1.0 else None},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    return out


def serving_leg(args, rates=(10, 50, 100, 200), seconds=3.0):
    """Requests arriving as a Poisson process against a resident database:
    each request is 1-500 queries (uniform), each query with 200 candidate
    subjects (half of the queries carry planted homologs among them).  For
    each arrival rate: p50 / p99 latency (submit to results on the host) and
    the sustained query rate; every served result is compared with the same
    request run alone."""
    import threading
    from topaz_amd import matrix, synth
    from topaz_amd.serve import Server
    from topaz_amd.sw import _concat
    rng = np.random.default_rng(2026)
    t0 = time.perf_counter()
    n_db = int(100000 * args.serve_scale)
    dl = synth.lognormal_lengths(rng, n_db, 350.0, 250.0, 30, 5000)
    dr, do, dn = synth.concat_random(rng, dl)
    pool = []
    for _ in range(24):
        nq = int(rng.integers(1, 501))
        ql = synth.lognormal_lengths(rng, nq, 350.0, 250.0, 30, 5000)
        qr, qo, qn = synth.concat_random(rng, ql)
        pq = np.repeat(np.arange(nq, dtype=np.int32), 200)
        ps = rng.integers(0, n_db, len(pq)).astype(np.int32)
        # planted homologs: mutated copies of the query written over a candidate
        hom = np.nonzero(rng.random(nq) < 0.5)[0]
        for k in hom:
            j = int(ps[200 * k])
            res, _, hn = synth.mutate(rng, qr, qo[k:k + 1], qn[k:k + 1], rng.uniform(0.25, 0.95, 1))
            m = min(int(hn[0]), int(dn[j]))
            dr[do[j]:do[j] + m] = res[:m]
        pool.append(((qr.astype(np.int8), qo, qn), pq, ps))
    db = (dr.astype(np.int8), do, dn)
    print("[bench] serving workload: %d subjects, %d requests in the pool, generated in %.1f s"
          % (n_db, len(pool), time.perf_counter() - t0), file=sys.stderr, flush=True)
    srv = Server(db, matrix(), 11, 1, memory_budget=1 << 30, max_inflight=2,
                 waves_per_group=args.serve_waves)
    offline = []
    for req in pool:
        offline.append(srv.wait(srv.submit(*req))[:3])
    rows = []
    identical = True
    for rate in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds) + 1))
        arrivals = arrivals[arrivals < seconds]
        picks = rng.integers(0, len(pool), len(arrivals))
        tickets = []
        done = []
        st0 = srv.stats()

        def waiter():
            k = 0
            while k < len(arrivals):
                if k >= len(tickets):
                    time.sleep(0.0005)
                    continue
                sc, re_, rd, lat = srv.wait(tickets[k])
                done.append((picks[k], lat, (sc, re_, rd)))
                k += 1

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for a, p in zip(arrivals, picks):
            delay = start + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(*pool[p]))
        th.join()
        wall = time.perf_counter() - start
        st = srv.stats()
        lats = np.array([d[1] for d in done])
        for p, _, res in done:
            for x, y in zip(res, offline[p]):
                if not np.array_equal(x, y):
                    identical = False
        n_q = int(sum(len(pool[p][0][2]) for p in picks))
        launches = st["launches"] - st0["launches"]
        rows.append({"offered_requests_per_s": rate, "requests": len(arrivals),
                     "queries_per_s": round(n_q / wall, 1),
                     "p50_ms": round(float(np.percentile(lats, 50)), 2),
                     "p99_ms": round(float(np.percentile(lats, 99)), 2),
                     "launches": launches,
                     "requests_per_launch": round(len(arrivals) / max(launches, 1), 2)})
        print("[bench] serving %d/s: %s" % (rate, rows[-1]), file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    return {"workload": "serving: %d resident subjects (lognormal mean 350 sd 250 clipped "
                        "[30,5000]); requests of U{1..500} queries x 200 candidates, Poisson "
                        "arrivals, %.0f s per rate, memory budget 1 GiB, 2 launches in flight, "
                        "%d-wavefront workgroups" % (n_db, seconds, args.serve_waves),
            "rates": rows, "identical_to_offline": identical,
            "peak_admitted_bytes": st["peak_bytes"], "workspace_bytes": st["workspace_bytes"]}


def stats_leg(args, n_queries=200, top=100):
    """Traceback statistics of reported hits (reverse pass, banded_sw): for
    each of the first n_queries queries of the 2M-pair workload, its `top`
    best-scoring pairs, as TOPAZ reports -H 100 hits per query, in one batch.
    A step is one tpz_stats_batch_run: upload, reverse pass, band search,
    trace, three host waits.  The reverse pass covers each hit's prefix
    rectangle, (read_end1 + 1) x (ref_end1 + 1) cells."""
    from topaz_amd import matrix
    from topaz_amd.stats import StatsBatch
    from topaz_amd.sw import SwBatch
    w = make_workload(n_queries / 2000.0)
    b = SwBatch(w["q"], w["s"], w["pair_query"], w["pair_subject"], matrix())
    sc, fe, re_ = b.run().results()
    b.close()
    pick = []
    for qi in range(len(w["q"][2])):
        idx = np.nonzero((w["pair_query"] == qi) & (sc > 0) & (sc < 32767))[0]
        pick.extend(idx[np.argsort(-sc[idx], kind="stable")][:top].tolist())
    pick = np.array(pick, np.int64)
    cells = int(np.sum((re_[pick].astype(np.int64) + 1) * (fe[pick].astype(np.int64) + 1)))
    st = StatsBatch(w["q"], w["s"], w["pair_query"][pick], w["pair_subject"][pick], re_[pick],
                    fe[pick], sc[pick], matrix())
    for _ in range(max(1, args.warmup)):
        st.run()
    ref = st.results()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        st.run()
        times.append(time.perf_counter() - t0)
    got = st.results()
    st.close()
    assert all(np.array_equal(x, y) for x, y in zip(got, ref))
    t = float(np.mean(times))
    print("[bench] stats: %d hits, %.3g reverse-pass cells, %.2f ms per run"
          % (len(pick), cells, t * 1e3), file=sys.stderr, flush=True)
    return {"workload": "traceback statistics: the %d best-scoring pairs of each of %d queries "
                        "of the 2M-pair workload, one batch" % (top, n_queries),
            "hits": int(len(pick)), "reverse_cells": cells, "ms_per_run": round(t * 1e3, 3),
            "hits_per_s": round(len(pick) / t, 1),
            "reverse_gcups": round(cells / t / 1e9, 2)}


def sa_leg(args, t_residues=100_000_000):
    """Suffix array of a database text (the .SA of topaz index, index.c:66):
    a 100 M-byte newline-separated protein text with 1 % of its proteins
    repeated exactly.  A step is one tpz_suffix_array call, text upload and
    suffix-array download (8 bytes per position) included."""
    import ctypes as C
    from topaz_amd import _lib, synth

    class SaStats(C.Structure):
        _fields_ = [("rounds", C.c_int32), ("device_bytes", C.c_int64), ("alloc_s", C.c_double),
                    ("sort_s", C.c_double), ("readback_s", C.c_double)]

    rng = np.random.default_rng(2030)
    text, _, _, _ = synth.internal_text_and_seeds(rng, t_residues, 10, 10)
    t = text.tobytes()
    prots = t.split(b"\n")[:-1]
    t += b"".join(prots[k] + b"\n" for k in rng.integers(0, len(prots), len(prots) // 100))
    L = _lib.lib()
    L.tpz_suffix_array.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(SaStats)]
    L.tpz_suffix_array.restype = C.c_int
    sa = np.empty(len(t), np.int64)
    st = SaStats()
    times, phases = [], []
    for k in range(max(1, args.warmup) + args.steps):
        t0 = time.perf_counter()
        _lib.check(L.tpz_suffix_array(t, len(t), sa.ctypes.data_as(C.POINTER(C.c_int64)),
                                      C.byref(st)), "tpz_suffix_array")
        if k >= max(1, args.warmup):
            times.append(time.perf_counter() - t0)
            phases.append((st.alloc_s, st.sort_s, st.readback_s))
    # a permutation whose neighbouring suffixes are in order on their first 8 bytes
    assert np.array_equal(np.bincount(sa, minlength=len(t)), np.ones(len(t), np.int64))
    pad = np.frombuffer(t + b"\0" * 8, np.uint8)
    key = np.zeros(len(t), np.uint64)
    for b in range(8):
        key = (key << np.uint64(8)) | pad[np.minimum(sa + b, len(t))].astype(np.uint64)
    assert np.all(key[1:] >= key[:-1])
    dt = float(np.mean(times))
    print("[bench] suffix array: %d bytes, %d rounds, %.3f s" % (len(t), st.rounds, dt),
          file=sys.stderr, flush=True)
    return {"workload": "suffix array of a %d-byte protein text (lognormal lengths mean 350, "
                        "1%% of proteins repeated exactly), upload and download included" % len(t),
            "text_bytes": len(t), "rounds": st.rounds, "s_per_build": round(dt, 3),
            "mbytes_per_s": round(len(t) / dt / 1e6, 1), "device_bytes": st.device_bytes,
            "phase_s": dict(zip(("alloc_upload", "rounds", "readback"),
                                (round(float(x), 4) for x in np.mean(phases, axis=0))))}


def _per_query(text):
    """A search's output as {query id: its lines in printed order}: the
    reference prints finished queries' blocks in varying order."""
    blocks = {}
    for line in text.splitlines():
        f = line.split("\t")
        if len(f) >= 2:
            blocks.setdefault(f[0], []).append(line)
    return blocks


def _run_search(exe, work, prefix, qfa, flags, threads, env, label):
    """One `topaz search`: wall seconds, TOPAZ's own elapsed time, stdout, and
    the library's totals when the binary prints them."""
    import subprocess
    # the binary's stderr (per-batch progress lines of the GPU build) goes to
    # a file as it is written, under gpurun_out/ when that folder exists
    logdir = os.path.join(ROOT, "gpurun_out") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) \
        else str(work)
    log = os.path.join(logdir, "search_%s.err" % label.replace(" ", "_"))
    t0 = time.perf_counter()
    with open(log, "w") as ferr:
        r = subprocess.run([str(exe), "search", "-f", str(qfa), "-p", str(prefix), "-T",
                            str(threads)] + flags, stdout=subprocess.PIPE, stderr=ferr, text=True,
                           cwd=str(work), env=dict(os.environ, **env), timeout=1500)
    wall = time.perf_counter() - t0
    err = open(log).read()
    if r.returncode != 0:
        raise SystemExit("%s failed: %s" % (label, err[-3000:]))
    el = [ln for ln in err.split("\n") if ln.startswith("elapsed time:")]
    st = [ln for ln in err.split("\n") if ln.startswith("tpz-stats ")]
    up = [ln for ln in err.split("\n") if ln.startswith("tpz: database resident")]
    info = {"wall_s": round(wall, 3),
            "topaz_elapsed_s": float(el[-1].split()[2]) if el else None,
            "db_upload_s": float(up[-1].split()[-2]) if up else None}
    if st:
        info["library"] = json.loads(st[-1][len("tpz-stats "):])
    print("[bench] search %s: %.2f s wall (%s)" % (label, wall, info["topaz_elapsed_s"]),
          file=sys.stderr, flush=True)
    return r.stdout, info


def search_leg(args):
    """End-to-end `topaz search` on a synthetic database: the reference
    binary on the CPU (all host cores) against TOPAZ with its workers
    searching batches of queries through tpz_search (the GPU), persistent
    and plain schedules; hit lists compared per query on the sample the CPU
    searches.  The index is written by topaz_amd.index.write_index (the
    files `topaz index` writes, suffix array built on the GPU)."""
    import shutil
    import tempfile
    from pathlib import Path
    from tests.test_dropin_gpu import REF, _link_dropin, _link_engine
    from topaz_amd import index, synth
    n_db, n_q = args.search_db, args.search_queries
    sample = min(args.search_sample, n_q)
    hi = 2000 if n_db <= 100000 else 5000
    hom = 0.2 if n_db <= 100000 else 0.1
    # with --workdir the files go to <workdir>/search (the profiles name them there)
    if args.workdir:
        work = Path(args.workdir) / "search"
        work.mkdir(parents=True, exist_ok=True)
    else:
        work = Path(tempfile.mkdtemp(prefix="tpz_search_"))
    out = {"workload": "search: %d database proteins (lognormal mean 350 sd 250 clipped [30,%d]), "
                       "%d queries, %d%% with 1-5 planted homologs at 30-90%% identity; -H 100 "
                       "-E 1e-3, default and -F" % (n_db, hi, n_q, int(hom * 100)),
           "proteins": n_db, "queries": n_q, "sample": sample}
    try:
        t0 = time.perf_counter()
        rng = np.random.default_rng(2040)
        text, lens, q = synth.search_text(rng, n_db, n_q, hom_frac=hom, hi=hi)
        out["text_bytes"] = int(len(text))
        t1 = time.perf_counter()
        print("[bench] search: %d-byte text generated in %.1f s" % (len(text), t1 - t0),
              file=sys.stderr, flush=True)
        st = index.write_index(work / "DB", text, lens, ["s%d" % k for k in range(len(lens))])
        out["index"] = {"write_s": round(time.perf_counter() - t1, 2), "sa_rounds": st.rounds,
                        "sa_s": round(st.alloc_s + st.sort_s + st.readback_s, 2)}
        del text
        print("[bench] search: index written in %.1f s" % out["index"]["write_s"],
              file=sys.stderr, flush=True)
        synth.write_queries_fasta(work / "q.fa", q)
        qs = (q[0], q[1][:sample], q[2][:sample])
        synth.write_queries_fasta(work / "qs.fa", qs)
        exe = _link_engine(work)
        cores = _cores()
        # the GPU runs' engine workers, each searching batches of up to 2,048
        # queries through tpz_search: with 8, some keep the device busy while
        # others are on the host (config 4: 23.7 s with 8 workers, 24.6 s
        # with 4; batches of 4,096 were slower); with fewer queries than 8
        # full batches, 4 workers, whose larger batches pay the per-batch
        # costs fewer times (1,000 queries: 0.40 s with 4, 0.52 s with 8),
        # and one batch on one worker when all queries fit one (1,000
        # queries: 0.26-0.31 s)
        gpu_workers = 8 if n_q >= 8 * 2048 else 4 if n_q > 2048 else 1
        batch = str(max(1, min(2048, (n_q + gpu_workers - 1) // gpu_workers)))
        gpu_env = {"TOPAZ_AMD_SEARCH_STATS": "1", "TOPAZ_AMD_SEARCH_BATCH": batch}
        plain_n = min(args.plain_sample, sample)
        synth.write_queries_fasta(work / "qp.fa", (q[0], q[1][:plain_n], q[2][:plain_n]))
        modes = [m for m in (("default", []), ("fast", ["-F"]), ("blasttab", ["-B"]))
                 if m[0] in args.search_modes.split(",")]
        for mode, flags in modes:
            f = ["-H", "100", "-E", "0.001"] + flags
            ref_txt, ref_i = _run_search(REF / "topaz", work, work / "DB", work / "qs.fa", f, cores,
                                         {}, "%s reference_cpu x%d" % (mode, sample))
            gpu_txt, gpu_i = _run_search(exe, work, work / "DB", work / "qs.fa", f, gpu_workers, gpu_env,
                                         "%s gpu x%d" % (mode, sample))
            pl_txt, pl_i = _run_search(exe, work, work / "DB", work / "qp.fa", f, 1,
                                       dict(gpu_env, TOPAZ_AMD_SEARCH_PLAIN="1"),
                                       "%s gpu plain x%d" % (mode, plain_n))
            ref_b = _per_query(ref_txt)
            pl_b = _per_query(pl_txt)
            row = {"reference_cpu": dict(ref_i, cores=cores,
                                         queries_per_s=round(sample / ref_i["wall_s"], 2)),
                   "gpu_sample": dict(gpu_i, queries_per_s=round(sample / gpu_i["wall_s"], 2)),
                   "gpu_plain_sample": dict(pl_i, queries=plain_n,
                                            queries_per_s=round(plain_n / pl_i["wall_s"], 2)),
                   "identical": _per_query(gpu_txt) == ref_b and
                                all(ref_b.get(k) == v for k, v in pl_b.items()) and
                                all(k in pl_b for k in ref_b if int(k[1:]) < plain_n),
                   "hit_lines": sum(len(v) for v in ref_b.values())}
            if args.dropin_sample > 0:
                # the link-level drop-in (one library call per seed and per
                # hit) and the batch loops (query.c's and suffix.c's loops
                # over the batch entry points) on the first queries
                nd = min(args.dropin_sample, sample)
                synth.write_queries_fasta(work / "qd.fa", (q[0], q[1][:nd], q[2][:nd]))
                ref_d = {k: v for k, v in ref_b.items() if int(k[1:]) < nd}
                for name, batch in (("dropin_per_call", False), ("batch_loops", True)):
                    txt, inf = _run_search(_link_dropin(work, batch=batch), work, work / "DB",
                                           work / "qd.fa", f, cores, {},
                                           "%s %s x%d" % (mode, name, nd))
                    row[name] = dict(inf, queries=nd, ms_per_query=round(1e3 * inf["wall_s"] / nd, 2),
                                     identical=_per_query(txt) == ref_d)
                    row["identical"] = row["identical"] and row[name]["identical"]
            if n_q > sample:
                _, all_i = _run_search(exe, work, work / "DB", work / "q.fa", f, gpu_workers, gpu_env,
                                       "%s gpu x%d" % (mode, n_q))
                row["gpu_all"] = dict(all_i, queries_per_s=round(n_q / all_i["wall_s"], 2))
            big = row.get("gpu_all", row["gpu_sample"])
            lib = big.get("library") or {}
            if lib.get("align_s"):
                row["sw_gcups_stage_wall"] = round(lib["cells"] / lib["align_s"] / 1e9, 2)
            row["speedup_vs_reference"] = round(big["queries_per_s"] /
                                                row["reference_cpu"]["queries_per_s"], 2)
            row["batched_vs_plain"] = round(row["gpu_sample"]["queries_per_s"] /
                                            row["gpu_plain_sample"]["queries_per_s"], 2)
            out[mode] = row
    finally:
        if not args.keep:
            shutil.rmtree(work, ignore_errors=True)
    return out


def serve_search_leg(args, rates=(1, 2, 5, 10, 20, 50, 100, 200)):
    """Serving searches: the search leg's database resident on the device
    (tpz_search_server), requests of U{1..500} queries arriving as a Poisson
    process; for each rate, at least --serve-requests requests: p50 / p99
    latency (submit to hits on the host) and the sustained query rate.
    Every served request's hits are compared with the same queries searched
    offline (one tpz_search batch of all the pool's queries)."""
    import shutil
    import tempfile
    import threading
    from pathlib import Path
    from topaz_amd import index, matrix, synth
    from topaz_amd.search import Database, SearchServer
    n_db, n_q = args.search_db, args.search_queries
    hi = 2000 if n_db <= 100000 else 5000
    hom = 0.2 if n_db <= 100000 else 0.1
    work = Path(args.workdir) / "serve_db" if args.workdir else Path(tempfile.mkdtemp(prefix="tpz_serve_"))
    work.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(2040)
    t0 = time.perf_counter()
    text, lens, q = synth.search_text(rng, n_db, n_q, hom_frac=hom, hi=hi)
    st = index.write_index(work / "DB", text, lens, ["s%d" % k for k in range(len(lens))])
    del text
    print("[bench] serve-search: index of %d proteins written in %.1f s"
          % (n_db, time.perf_counter() - t0), file=sys.stderr, flush=True)
    queries = [bytes(index.INTERNAL_OF_INDEX[q[0][o:o + m].astype(np.int64)])
               for o, m in zip(q[1], q[2])]
    db = Database.from_index(work / "DB", matrix())
    try:
        prng = np.random.default_rng(2041)
        pool = []
        for _ in range(64):
            k = int(prng.integers(1, 501))
            pool.append([int(i) for i in prng.integers(0, len(queries), k)])
        # offline: the pool's distinct queries, in batches of 2048
        used = sorted({i for r in pool for i in r})
        offline = {}
        t1 = time.perf_counter()
        for c in range(0, len(used), 2048):
            part = used[c:c + 2048]
            hits, _ = db.search([queries[i] for i in part])
            offline.update(zip(part, hits))
        t_off = time.perf_counter() - t1
        print("[bench] serve-search: %d queries offline in %.1f s" % (len(used), t_off),
              file=sys.stderr, flush=True)
        srv = SearchServer(db, lanes=args.serve_lanes)
        rows = []
        identical = True
        for rate in rates:
            n_req = args.serve_requests
            arrivals = np.cumsum(prng.exponential(1.0 / rate, n_req))
            picks = prng.integers(0, len(pool), n_req)
            tickets, done = [], []
            st0 = srv.stats()

            def waiter():
                k = 0
                while k < n_req:
                    if k >= len(tickets):
                        time.sleep(0.0005)
                        continue
                    hits, _, lat = srv.wait(tickets[k])
                    done.append((picks[k], lat, hits, time.perf_counter()))
                    k += 1

            th = threading.Thread(target=waiter)
            th.start()
            start = time.perf_counter()
            for a, p in zip(arrivals, picks):
                delay = start + a - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets.append(srv.submit([queries[i] for i in pool[p]]))
            th.join()
            wall = max(d[3] for d in done) - start
            st1 = srv.stats()
            for p, _, hits, _ in done:
                if hits != [offline[i] for i in pool[p]]:
                    identical = False
            lats = np.array([d[1] for d in done])
            n_queries = int(sum(len(pool[p]) for p in picks))
            waves = st1["waves"] - st0["waves"]
            rows.append({"offered_requests_per_s": rate, "requests": n_req,
                         "offered_queries_per_s": round(n_queries / float(arrivals[-1]), 1),
                         "queries_per_s": round(n_queries / wall, 1),
                         "p50_ms": round(float(np.percentile(lats, 50)), 1),
                         "p99_ms": round(float(np.percentile(lats, 99)), 1),
                         "waves": waves, "requests_per_wave": round(n_req / max(waves, 1), 2)})
            print("[bench] serve-search %d/s: %s" % (rate, rows[-1]), file=sys.stderr, flush=True)
        sst = srv.stats()
        srv.close()
    finally:
        db.close()
        if not args.workdir and not args.keep:
            shutil.rmtree(work, ignore_errors=True)
    return {"workload": "serving searches: %d resident proteins (lognormal mean 350 sd 250 clipped "
                        "[30,%d]); requests of U{1..500} queries from %d queries, Poisson arrivals, "
                        "%d requests per rate, %d lanes, budget 16 GiB with the fixed workspace"
                        % (n_db, hi, n_q, args.serve_requests, args.serve_lanes),
            "offline_queries_per_s": round(len(used) / t_off, 1),
            "rates": rows, "identical_to_offline": identical,
            "peak_bytes": sst["peak_bytes"], "fixed_bytes": sst["fixed_bytes"],
            "max_requests_per_wave": sst["max_requests_per_wave"]}


def index_leg(args):
    """`topaz index` of a synthetic database: the reference binary (its own
    libdivsufsort, one core) against TOPAZ linked with the library's
    divsufsort64 (suffix array built on the GPU); both must write identical
    index files.  Lognormal lengths (mean 350, sd 250, clipped [30, 2000]),
    1 % of the proteins repeated exactly, as real databases hold identical
    sequences."""
    import filecmp
    import shutil
    import subprocess
    import tempfile
    from pathlib import Path
    from tests.test_dropin_gpu import REF, _link_index
    from topaz_amd import synth
    if not (REF / "topaz").exists():
        raise SystemExit("reference checker not built (make -C oracle ref)")
    tmp = Path(tempfile.mkdtemp(prefix="tpz_index_", dir=args.workdir))
    try:
        rng = np.random.default_rng(2029)
        t0 = time.perf_counter()
        db, _ = synth.search_database(rng, args.index_db, 1, hom_frac=0.0, hi=2000)
        db += [db[k] for k in rng.integers(0, len(db), max(1, args.index_db // 100))]
        synth.write_fasta(tmp / "db.fa", db, "s")
        print("[bench] index: %d proteins written in %.1f s" % (len(db), time.perf_counter() - t0),
              file=sys.stderr, flush=True)
        exe = _link_index(tmp)
        times = {}
        for name, binary in (("reference_cpu", REF / "topaz"), ("gpu_sa", exe)):
            (tmp / name).mkdir(exist_ok=True)
            t0 = time.perf_counter()
            r = subprocess.run([str(binary), "index", "-f", str(tmp / "db.fa"), "-p",
                                str(tmp / name / "DB")], capture_output=True, text=True,
                               timeout=1000, cwd=str(tmp))
            times[name] = time.perf_counter() - t0
            if r.returncode != 0:
                raise SystemExit("%s index failed: %s" % (name, r.stderr[-2000:]))
            print("[bench] index %s: %.2f s" % (name, times[name]), file=sys.stderr, flush=True)
        files = sorted(p.name for p in (tmp / "reference_cpu").iterdir())
        identical = files == sorted(p.name for p in (tmp / "gpu_sa").iterdir()) and all(
            filecmp.cmp(tmp / "reference_cpu" / f, tmp / "gpu_sa" / f, shallow=False)
            for f in files)
        text_bytes = os.path.getsize(tmp / "reference_cpu" / "DB.psq")
    finally:
        if not args.keep:
            shutil.rmtree(tmp, ignore_errors=True)
    return {"workload": "topaz index of %d proteins (lognormal mean 350 sd 250 clipped [30,2000], "
                        "1%% repeated exactly)" % len(db),
            "proteins": len(db), "text_bytes": text_bytes,
            "index_s": {k: round(v, 2) for k, v in times.items()},
            "speedup": round(times["reference_cpu"] / times["gpu_sa"], 2),
            "files_identical": identical, "files": files}


def copy_leg(args):
    """A plain device-to-device copy of 1 GiB, three times: the calibration
    run for the HBM byte counters (FETCH_SIZE / WRITE_SIZE per dispatch of
    a copy of known size)."""
    import torch
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    return {"metric": "device copy, GB/s (read + write)", "value": round(2 * n / dt / 1e9, 1),
            "unit": "GB/s", "bytes_copied": n, "ms_per_copy": round(dt * 1e3, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs-scale", type=float, default=1.0,
                    help="fraction of the 2,000-query workload (tests only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default configuration")
    ap.add_argument("--leg", choices=["all", "sw", "gapless", "serve", "stats", "sa", "copy",
                                      "search", "serve-search", "index"], default="all")
    ap.add_argument("--search-db", type=int, default=100000,
                    help="database proteins of the search leg")
    ap.add_argument("--search-queries", type=int, default=1000)
    ap.add_argument("--search-sample", type=int, default=1000,
                    help="queries the reference CPU binary (and the plain schedule) search")
    ap.add_argument("--plain-sample", type=int, default=1000,
                    help="queries the plain schedule searches")
    ap.add_argument("--search-modes", default="default,fast",
                    help="search modes to run, comma-separated: default, fast (-F), blasttab (-B)")
    ap.add_argument("--dropin-sample", type=int, default=0,
                    help="queries the link-level drop-in and the batch loops search (0: neither)")
    ap.add_argument("--index-db", type=int, default=1000000, help="proteins of the index leg")
    ap.add_argument("--workdir", default=None, help="where the search leg writes its index")
    ap.add_argument("--keep", action="store_true", help="keep the search leg's files")
    ap.add_argument("--serve-requests", type=int, default=100, help="requests per arrival rate")
    ap.add_argument("--serve-lanes", type=int, default=2, help="batches the search server runs at once")
    ap.add_argument("--serve-waves", type=int, default=8, choices=[4, 8],
                    help="wavefronts per workgroup of the serving loop's launches")
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="fraction of the 100,000-subject serving database")
    ap.add_argument("--no-search-10m", action="store_true",
                    help="leave the 10M-protein search out of the default run")
    ap.add_argument("--gapless-scale", type=float, default=1.0,
                    help="fraction of the 1e9-residue, 5e7-seed gapless workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(queries would be sharded across GPUs with a replicated index); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.leg == "search":
        print(json.dumps({"metric": "topaz search queries/s, GPU vs reference CPU",
                          "search": search_leg(args)}), flush=True)
        return
    if args.leg == "serve-search":
        print(json.dumps({"metric": "serving searches: p50/p99 latency vs sustained queries/s",
                          "serving_search": serve_search_leg(args)}), flush=True)
        return
    if args.leg == "copy":
        print(json.dumps(copy_leg(args)), flush=True)
        return
    if args.leg == "index":
        print(json.dumps({"metric": "topaz index wall seconds, GPU suffix array vs reference CPU",
                          "index": index_leg(args)}), flush=True)
        return
    if args.leg == "sa":
        print(json.dumps({"metric": "suffix array of the database text", "suffix_array": sa_leg(args)}),
              flush=True)
        return
    if args.leg == "stats":
        print(json.dumps({"metric": "traceback statistics per batch of reported hits",
                          "stats": stats_leg(args)}), flush=True)
        return
    if args.leg == "serve":
        sv = serving_leg(args)
        print(json.dumps({"metric": "serving p50/p99 latency vs sustained queries/s",
                          "serving": sv}), flush=True)
        return
    if args.leg in ("all", "sw"):
        out = sw_leg(args)
        if args.leg == "all":
            out["gapless"] = gapless_leg(args)
            out["serving"] = serving_leg(args)
            out["stats"] = stats_leg(args)
            out["suffix_array"] = sa_leg(args)
            # end-to-end `topaz search` on the small configuration: 100,000
            # proteins, 1,000 queries, default and -F, reference CPU vs GPU
            sargs = argparse.Namespace(**vars(args))
            sargs.search_db, sargs.search_queries, sargs.search_sample = 100000, 1000, 1000
            sargs.plain_sample, sargs.search_modes, sargs.dropin_sample = 200, "default,fast", 0
            out["search"] = search_leg(sargs)
            d = out["search"]["default"]
            out["search_queries_per_s"] = d["gpu_sample"]["queries_per_s"]
            out["search_vs_reference_cpu"] = d["speedup_vs_reference"]
            out["search_batched_vs_plain"] = d["batched_vs_plain"]
            out["search_identical"] = all(out["search"][m]["identical"] for m in ("default", "fast"))
            top = [r for r in out["serving"]["rates"] if r["offered_requests_per_s"] == 100]
            if top:
                out["latency_p50_ms"] = top[0]["p50_ms"]
                out["latency_p99_ms"] = top[0]["p99_ms"]
            if not args.no_search_10m:
                # end-to-end `topaz search` at scale: 10,000,000 proteins
                # (3.5 G residues), 100,000 queries, default and -F; the
                # reference on all host cores over the first 1,000 queries,
                # the plain schedule over the first 200
                sargs.search_db, sargs.search_queries, sargs.search_sample = 10000000, 100000, 1000
                sargs.plain_sample, sargs.search_modes = 200, "default,fast"
                try:
                    out["search_10M"] = s10 = search_leg(sargs)
                    for mode in ("default", "fast"):
                        r = s10[mode]
                        out["search_10M_%s" % mode] = {
                            "queries_per_s": r["gpu_all"]["queries_per_s"],
                            "vs_reference_cpu": r["speedup_vs_reference"],
                            "batched_vs_plain": r["batched_vs_plain"],
                            "identical": r["identical"]}
                except (Exception, SystemExit) as e:      # the headline line still prints
                    out["search_10M"] = {"error": str(e)[-2000:]}
                    print("[bench] search 10M failed: %s" % e, file=sys.stderr, flush=True)
    else:
        g = gapless_leg(args)
        m = g["matched"]
        out = {"metric": "gapless HSP extension, achieved HBM GB/s (first-residue-matched seeds)",
               "value": m["roofline"]["achieved"], "unit": "GB/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
               "data": "synthetic", "config": {"workload": g["workload"]},
               "roofline": m["roofline"], "gapless": g}
        if "cpu_baseline" in g:
            out["cpu_baseline"] = g["cpu_baseline"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
