"""tpz_search through the Python host layer, and the search server.

* Offline: Database.search (persistent and plain schedules) prints, line for
  line, what the reference binary prints for the same queries with SEG off
  (-M), default and -B.
* Streams cut short at a tiny cap: the same hits.
* Serving: requests submitted from several threads at once, merged into
  shared batches by the server, each get exactly the hits the same queries
  get searched offline; a request served alone gets them too; the admitted
  device bytes stay within the budget; a request over the budget is refused.
"""
import subprocess
import threading
from pathlib import Path

import numpy as np
import pytest

from topaz_amd import TopazError, index, matrix, synth
from topaz_amd.search import SCHED_PLAIN, Database, SearchServer, format_hits, header_ids

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"


@pytest.fixture(scope="module")
def db_and_queries(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("search")
    rng = np.random.default_rng(61)
    text, lens, q = synth.search_text(rng, 20000, 120, hom_frac=0.5, hi=2000)
    index.write_index(tmp / "DB", text, lens, ["s%d" % k for k in range(len(lens))])
    synth.write_queries_fasta(tmp / "q.fa", q)
    qs = [bytes(index.INTERNAL_OF_INDEX[q[0][o:o + n].astype(np.int64)])
          for o, n in zip(q[1], q[2])]
    db = Database.from_index(tmp / "DB", matrix())
    yield tmp, db, qs
    db.close()


def _ref_lines(tmp, *flags, queries="q.fa", seg_off=True):
    if not (REF / "topaz").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    r = subprocess.run([str(REF / "topaz"), "search", "-f", str(tmp / queries), "-p",
                        str(tmp / "DB"), "-T", "4"] + (["-M"] if seg_off else []) + list(flags),
                       capture_output=True, text=True, timeout=600, cwd=str(tmp))
    assert r.returncode == 0
    blocks = {}
    for line in r.stdout.splitlines():
        f = line.split("\t")
        if len(f) >= 2:
            blocks.setdefault(f[0], []).append(line)
    return blocks


@pytest.mark.parametrize("blasttab", [False, True])
def test_offline_search_matches_reference(db_and_queries, blasttab):
    tmp, db, qs = db_and_queries
    ref = _ref_lines(tmp, *(["-B"] if blasttab else []))
    ids = header_ids(tmp / "DB", np.fromfile(str(tmp / "DB.ind"), index.IND_DTYPE))
    opt = db.default_options(superfast=0 if blasttab else 1)
    for schedule in (1, SCHED_PLAIN):
        hits, st = db.search(qs, opt, schedule)
        ours = {}
        for k, h in enumerate(hits):
            lines = format_hits("q%d" % k, h, ids, superfast=not blasttab)
            if lines:
                ours["q%d" % k] = lines
        assert ours == ref
    assert sum(len(v) for v in ref.values()) > 100 and st["pairs"] > 0


@pytest.mark.parametrize("go,ge,modes", [(3, 3, (False, True)), (6, 6, (True,)),
                                         (1, 4, (False,))])
def test_offline_search_lane_gaps(db_and_queries, go, ge, modes):
    """gap_open <= gap_extend (-o/-e): the alignments and the statistics'
    reverse pass follow the reference's lanes; the printed lines equal the
    reference binary's, default and -B.  (With -B only settings whose
    alignments here stay inside banded_sw's band: for gap-heavy ones the
    reference's traceback reads direction cells its last band pass did not
    write, and its own statistics then vary from process to process.)"""
    tmp, _, qs = db_and_queries
    ids = header_ids(tmp / "DB", np.fromfile(str(tmp / "DB.ind"), index.IND_DTYPE))
    db = Database.from_index(tmp / "DB", matrix(), gap_open=go, gap_extend=ge)
    try:
        for blasttab in modes:
            ref = _ref_lines(tmp, "-o", str(go), "-e", str(ge), *(["-B"] if blasttab else []))
            hits, st = db.search(qs, db.default_options(superfast=0 if blasttab else 1))
            ours = {}
            for k, h in enumerate(hits):
                lines = format_hits("q%d" % k, h, ids, superfast=not blasttab)
                if lines:
                    ours["q%d" % k] = lines
            assert ours == ref, (go, ge, blasttab)
            assert st["pairs"] > 0
    finally:
        db.close()


def test_offline_search_with_seg(db_and_queries):
    """SEG on (TOPAZ's default, no -M): queries read from FASTA and masked by
    the library's SEG (topaz_amd.search.read_queries) give the reference
    binary's lines, default and -B.  Half the queries carry planted
    low-complexity stretches."""
    tmp, db, _ = db_and_queries
    from topaz_amd.search import read_queries
    rng = np.random.default_rng(77)
    recs = []
    for k, (qid, q) in enumerate(read_queries(tmp / "q.fa", use_seg=False)):
        letters = bytes(index.INTERNAL_OF_INDEX.tolist().index(c) for c in q)
        text = synth.to_letters(np.frombuffer(letters, np.uint8)).encode()
        if k % 2 == 0:
            at = int(rng.integers(0, len(text) + 1))
            junk = [b"Q" * 25, b"SG" * 15, b"PAPAPPAAPAPPA" * 2, b"EEEKKEEEKKEKE" * 2][k % 4]
            text = text[:at] + junk + text[at:]
        recs.append((qid, text))
    with open(tmp / "q_seg.fa", "w") as f:
        for qid, text in recs:
            f.write(">%s\n%s\n" % (qid, text.decode()))
    qs = read_queries(tmp / "q_seg.fa")
    masked = sum(q.count(b"V") for _, q in qs)
    assert masked > 500
    ids = header_ids(tmp / "DB", np.fromfile(str(tmp / "DB.ind"), index.IND_DTYPE))
    for blasttab in (False, True):
        ref = _ref_lines(tmp, *(["-B"] if blasttab else []), queries="q_seg.fa", seg_off=False)
        hits, _ = db.search([q for _, q in qs], db.default_options(superfast=0 if blasttab else 1))
        ours = {}
        for (qid, _), h in zip(qs, hits):
            lines = format_hits(qid, h, ids, superfast=not blasttab)
            if lines:
                ours[qid] = lines
        assert ours == ref, blasttab


def test_int64_suffix_array_gives_the_same_hits(db_and_queries, monkeypatch):
    """A text under 4 GiB keeps its suffix array as uint32 on the device;
    TOPAZ_AMD_SA64 keeps the int64 form that larger texts use.  Both give
    the same hits, default and -B."""
    tmp, db, qs = db_and_queries
    monkeypatch.setenv("TOPAZ_AMD_SA64", "1")
    db64 = Database.from_index(tmp / "DB", matrix())
    try:
        for superfast in (1, 0):
            opt = db.default_options(superfast=superfast)
            assert db64.search(qs[:60], opt)[0] == db.search(qs[:60], opt)[0]
    finally:
        db64.close()


def test_stage_s_in_two_parts_gives_the_same_hits(db_and_queries, monkeypatch):
    """A batch from TOPAZ_AMD_SW_PART_PAIRS pairs on (1 M by default, so
    never at this size) submits stage S in two parts of whole queries, each
    part's output stage running while the next is on the device: with the
    threshold at 1 pair every batch splits, and the hits stay the same."""
    tmp, db, qs = db_and_queries
    monkeypatch.setenv("TOPAZ_AMD_SW_PART_PAIRS", "1")
    other = Database.from_index(tmp / "DB", matrix())
    try:
        for superfast in (1, 0):
            opt = db.default_options(superfast=superfast)
            assert other.search(qs[:60], opt)[0] == db.search(qs[:60], opt)[0]
            assert other.search(qs[:1], opt)[0] == db.search(qs[:1], opt)[0]
    finally:
        other.close()


@pytest.mark.parametrize("small_pairs", ["0", "600"])
def test_stage_s_launch_widths_give_the_same_hits(db_and_queries, monkeypatch, small_pairs):
    """The search database's stage S launches run 8 wavefronts per
    workgroup below a pair count and 4 above (one workspace for both): all
    launches with 4 ("0"), and a threshold that splits this workload's
    launches between the two, give the hits of the default (every launch
    of this small database with 8)."""
    tmp, db, qs = db_and_queries
    monkeypatch.setenv("TOPAZ_AMD_SMALL_LAUNCH_PAIRS", small_pairs)
    other = Database.from_index(tmp / "DB", matrix())
    try:
        for superfast in (1, 0):
            opt = db.default_options(superfast=superfast)
            got = [other.search(qs[k:k + 8], opt)[0] for k in range(0, 48, 8)]
            want = [db.search(qs[k:k + 8], opt)[0] for k in range(0, 48, 8)]
            assert got == want
    finally:
        other.close()


@pytest.mark.parametrize("first_chunk", [3, 24, 40, 100])
def test_stream_round_widths_give_the_same_hits(db_and_queries, first_chunk):
    """Stream rounds serve 64 / w streams per wavefront when a round's chunk
    is at most 32 slots (w its power-of-two width), one stream per
    wavefront part otherwise: first chunks that leave lanes of a unit idle
    (3 of 4), fill units of 32, leave a wavefront part-full (40) or take two
    (100) give the hits of the default rounds."""
    tmp, db, qs = db_and_queries
    full, _ = db.search(qs[:40])
    got, st = db.search(qs[:40], db.default_options(first_chunk=first_chunk))
    assert got == full
    assert st["rounds"] > 0


def test_streams_cut_at_the_cap_are_searched_again(db_and_queries):
    """A walk that reaches the end of a stream cut at max_stream_elements
    searches its query again with a higher cap: the hits are those of an
    uncapped search."""
    tmp, db, qs = db_and_queries
    full, _ = db.search(qs[:40])
    opt = db.default_options(max_stream_elements=8, first_chunk=2)
    cut, st = db.search(qs[:40], opt)
    assert cut == full
    assert st["refills"] > 0


def test_server_results_equal_offline(db_and_queries):
    tmp, db, qs = db_and_queries
    offline, _ = db.search(qs)
    rng = np.random.default_rng(7)
    reqs = []
    for _ in range(24):
        n = int(rng.integers(1, 40))
        reqs.append([int(i) for i in rng.integers(0, len(qs), n)])
    srv = SearchServer(db, lanes=3, max_wave_queries=64)
    try:
        # alone first
        hits, _, lat = srv.wait(srv.submit([qs[i] for i in reqs[0]]))
        assert hits == [offline[i] for i in reqs[0]] and lat > 0
        # then all at once from three threads
        out = {}
        errors = []

        def client(part):
            try:
                tickets = [(r, srv.submit([qs[i] for i in reqs[r]])) for r in part]
                for r, t in tickets:
                    out[r] = srv.wait(t)[0]
            except Exception as e:      # reported below
                errors.append(e)

        th = [threading.Thread(target=client, args=(list(range(k, len(reqs), 3)),))
              for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        for r, idx in enumerate(reqs):
            assert out[r] == [offline[i] for i in idx], r
        st = srv.stats()
        assert st["requests_done"] == len(reqs) + 1
        assert st["max_requests_per_wave"] > 1
        assert st["peak_bytes"] <= 16 << 30
    finally:
        srv.close()


def test_server_refuses_request_over_budget(db_and_queries):
    tmp, db, qs = db_and_queries
    srv = None
    try:
        # a budget barely above the fixed workspace: one long query is too much
        probe = SearchServer(db, lanes=1)
        fixed = probe.stats()["fixed_bytes"]
        probe.close()
        srv = SearchServer(db, memory_budget=fixed + 4096, lanes=1)
        with pytest.raises(TopazError):
            srv.submit([qs[0] * 4])
    finally:
        if srv:
            srv.close()
