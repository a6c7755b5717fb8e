"""The seed walk's host logic (tpz_walk_query: runs of the priority queue,
the hit list) against the reference's own walk, on the CPU.

oracle/_ref/ref_walk runs the reference's compiled query_find_suffixes and
query_fetch_alignment_batch (query.c:94-174, priority_queue.c:36-56,
suffix.c:40-65, hit.c:126-193) and prints each query's hit list in
traversal order.  The library's walk gets the same query (after SEG, in the
internal alphabet), the suffix-array positions of its offsets from the
oracle's sa_search, and the extensions from the oracle's sa_hsp through the
extension callback; both hit lists must be equal entry for entry, order
included.  Small hit-list capacities and seed limits force evictions, ties
and an early stop; chunk sizes of 1 force a refill at every element.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.refcheck import Oracle
from topaz_amd import _lib, matrix, synth

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"

EXTEND = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int32, C.c_int64, C.c_int32, C.c_int32,
                     C.POINTER(C.c_int16))


class WalkInput(C.Structure):
    _fields_ = [("n_offsets", C.c_int32), ("offsets", C.c_void_p), ("positions", C.c_void_p),
                ("t_size", C.c_uint64), ("sa", C.c_void_p), ("n_proteins", C.c_uint64),
                ("protein_start", C.c_void_p), ("protein_len", C.c_void_p),
                ("number_of_seeds", C.c_int32), ("capacity", C.c_int32),
                ("max_score", C.c_int32), ("first_chunk", C.c_int32), ("refill", C.c_int32)]


class WalkStats(C.Structure):
    _fields_ = [("pops", C.c_int64), ("runs", C.c_int64), ("refills", C.c_int64),
                ("elements", C.c_int64), ("counted", C.c_int32)]


IND = np.dtype([("header_start", "<u8"), ("protein_start", "<u8"), ("header_len", "<i4"),
                ("protein_len", "<i4")])


def load_db(prefix):
    T = Path(str(prefix) + ".psq").read_bytes()
    SA = np.fromfile(str(prefix) + ".SA", np.int64)
    ind = np.fromfile(str(prefix) + ".ind", IND)
    return T, SA, ind


def ref_walks(prefix, qfa, seeds, alignments):
    r = subprocess.run([str(REF / "ref_walk"), str(prefix), str(qfa), str(seeds),
                        str(alignments)], capture_output=True, timeout=600, check=True)
    out = []
    for line in r.stdout.split(b"\n"):
        if not line:
            continue
        tag, rest = line[:1], line[2:]
        if tag == b"Q":
            out.append({"id": rest.split()[0].decode(), "hits": []})
        elif tag == b"S":
            out[-1]["seq"] = rest
        elif tag == b"H":
            a, b = rest.split()
            out[-1]["hits"].append((int(a), int(b)))
        elif tag == b"C":
            out[-1]["count"] = int(rest)
    return out


def upper_aa_only(c):
    """lexicographical.h:60-74"""
    ch = chr(c)
    return ch if ("A" <= ch <= "Y") or ch in "*\n" else "V"


class OurWalk:
    def __init__(self, T, SA, ind):
        self.L = _lib.lib()
        L = self.L
        L.tpz_walk_query.argtypes = [C.POINTER(WalkInput), EXTEND, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.POINTER(C.c_int32), C.POINTER(WalkStats)]
        L.tpz_walk_query.restype = C.c_int
        self.T, self.SA = T, SA
        self.starts = np.ascontiguousarray(ind["protein_start"], np.uint64)
        self.lens = np.ascontiguousarray(ind["protein_len"], np.int32)
        self.orc = Oracle()
        self.mat = matrix()
        self.calls = 0

    def positions(self, q, offsets):
        pos = np.full(len(offsets), -1, np.int64)
        for k, o in enumerate(offsets):
            rc, idx, _, _ = self.orc.sa_search(self.T, self.SA, q[o:])
            pos[k] = idx if rc == 0 else -1
        return pos

    def run(self, q, seeds, capacity, first_chunk=64, refill=128, min_len=5):
        offsets = np.array([i for i in range(len(q) - min_len) if upper_aa_only(q[i]) != "V"],
                           np.int32)
        pos = self.positions(q, offsets)
        T, SA, n = self.T, self.SA, len(self.T)

        def extend(ctx, offset, first, direction, count, out):
            self.calls += 1
            slots = first + direction * np.arange(count, dtype=np.int64)
            ok = (slots >= 0) & (slots < n)
            seeds_ = np.zeros(int(ok.sum()), dtype=[("query", "<i4"), ("offset", "<i4"),
                                                    ("position", "<i8")])
            seeds_["offset"] = offset
            seeds_["position"] = SA[slots[ok]]
            sc, ln = self.orc.sa_hsp_batch(T, [q], seeds_, self.mat)
            v = np.full(count, -1, np.int16)
            v[ok] = np.where(ln >= min_len, sc, -1)
            C.memmove(out, v.ctypes.data, 2 * count)
            return 0

        cb = EXTEND(extend)
        inp = WalkInput(len(offsets), offsets.ctypes.data, pos.ctypes.data, n,
                        SA.ctypes.data, len(self.starts), self.starts.ctypes.data,
                        self.lens.ctypes.data, seeds, capacity, 255 * 11, first_chunk, refill)
        cap = max(capacity, 1) + 8
        ix = np.zeros(cap, np.uint64)
        sc = np.zeros(cap, np.int32)
        nh = C.c_int32(cap)
        st = WalkStats()
        _lib.check(self.L.tpz_walk_query(C.byref(inp), cb, None, ix.ctypes.data, sc.ctypes.data,
                                         C.byref(nh), C.byref(st)), "tpz_walk_query")
        k = nh.value
        return list(zip(ix[:k].tolist(), sc[:k].tolist())), st


@pytest.fixture(scope="module")
def small_db(tmp_path_factory):
    if not (REF / "ref_walk").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    tmp = tmp_path_factory.mktemp("walk")
    rng = np.random.default_rng(41)
    db, queries = synth.search_database(rng, 3000, 30, hom_frac=0.6, max_hom=5)
    # exact copies (equal scores from several proteins: ties), and queries
    # copied from the database (long exact matches: long runs)
    db += [db[k] for k in rng.integers(0, len(db), 60)]
    queries += [db[k] for k in rng.integers(0, len(db), 6)]
    queries.append(queries[0][:5])                                   # no offsets
    queries.append(np.full(80, 9, np.int8))                          # low complexity: SEG
    synth.write_fasta(tmp / "db.fa", db, "s")
    synth.write_fasta(tmp / "q.fa", queries, "q")
    subprocess.run([str(REF / "topaz"), "index", "-f", str(tmp / "db.fa"), "-p", str(tmp / "DB")],
                   check=True, capture_output=True, timeout=300)
    return tmp


@pytest.mark.parametrize("seeds,alignments,first_chunk,refill", [
    (300000, 5000, 64, 128),        # TOPAZ's defaults
    (100000, 1500, 1, 1),           # --fast, a refill at every element
    (40, 25, 3, 2),                 # early stop, evictions
    (5000, 7, 16, 32),              # a tiny hit list: ties at the eviction edge
])
def test_walk_matches_reference(small_db, seeds, alignments, first_chunk, refill):
    tmp = small_db
    T, SA, ind = load_db(tmp / "DB")
    ref = ref_walks(tmp / "DB", tmp / "q.fa", seeds, alignments)
    ours = OurWalk(T, SA, ind)
    n_hits = 0
    for r in ref:
        q = r["seq"]
        cap = alignments if len(q) <= 3000 else int(1.5 * 100)
        hits, st = ours.run(q, seeds, cap, first_chunk, refill)
        assert hits == r["hits"], r["id"]
        n_hits += len(hits)
        assert st.counted <= seeds
    assert n_hits > 50


def test_walk_matches_reference_100k(tmp_path):
    """The walk over the 100,000-protein end-to-end database (TOPAZ's
    defaults), for 40 of its queries."""
    if not (REF / "ref_walk").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    rng = np.random.default_rng(2027)
    db, queries = synth.search_database(rng, 100000, 1000, hom_frac=0.2, max_hom=5,
                                        ident=(0.3, 0.9), hi=2000)
    synth.write_fasta(tmp_path / "db.fa", db, "s")
    synth.write_fasta(tmp_path / "q.fa", queries[:40], "q")
    subprocess.run([str(REF / "topaz"), "index", "-f", str(tmp_path / "db.fa"), "-p",
                    str(tmp_path / "DB")], check=True, capture_output=True, timeout=600)
    T, SA, ind = load_db(tmp_path / "DB")
    ref = ref_walks(tmp_path / "DB", tmp_path / "q.fa", 300000, 5000)
    ours = OurWalk(T, SA, ind)
    pops = 0
    for r in ref:
        hits, st = ours.run(r["seq"], 300000, 5000)
        assert hits == r["hits"], r["id"]
        pops += st.pops
    assert pops > 40 * 1000
