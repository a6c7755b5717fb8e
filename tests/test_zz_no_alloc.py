"""The no-allocation gapless entry point, tpz_gapless_run_noalloc, against
the batch entry point it varies (tpz_gapless_batch_create/run/results): two
calls in a row on one device, the second with fewer queries and seeds, each
bit-identical to the batch results.  Device memory and the stream come from
the HIP runtime the library links; the call goes through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from topaz_amd import _lib, matrix, synth
from topaz_amd.gapless import SEED_DTYPE, GaplessBatch, Text
from topaz_amd.matrices import INTERNAL_ORDER, MATRIX_ORDER

pytestmark = pytest.mark.gpu


def _declare(L):
    VP = C.c_void_p
    L.tpz_gapless_run_noalloc.argtypes = [VP, C.c_int32, VP, C.c_int64, VP, VP, C.c_int64, VP,
                                          C.POINTER(C.c_int8), VP, VP, VP, VP]
    L.tpz_gapless_run_noalloc.restype = C.c_int


def _internal(idx):
    return bytes(ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[i]) for i in idx)


def _workload(rng, T, n_queries, n_seeds):
    Tn = np.frombuffer(T, np.uint8)
    qs = []
    for L in synth.lognormal_lengths(rng, n_queries, 200, 150, 1, 900):
        if rng.random() < 0.5:
            st = int(rng.integers(0, len(T) - 1))
            q = bytearray(T[st:st + L].replace(b"\n", b"G"))
            for k in np.nonzero(rng.random(len(q)) < 0.2)[0]:
                q[k] = ord("A") + int(rng.integers(0, 25))
            qs.append(bytes(q) or b"A")
        else:
            qs.append(_internal(synth.random_residues(rng, L)))
    seeds = np.zeros(n_seeds, SEED_DTYPE)
    seeds["query"] = rng.integers(0, len(qs), n_seeds)
    seeds["offset"] = [int(rng.integers(0, len(qs[q]))) for q in seeds["query"]]
    pos = rng.integers(0, len(T), n_seeds)
    # half of them on a matching first residue, as suffix-array seeds are
    for k in range(0, n_seeds, 2):
        c = qs[seeds["query"][k]][seeds["offset"][k]]
        hits = np.nonzero(Tn == c)[0]
        if len(hits):
            pos[k] = hits[int(rng.integers(0, len(hits)))]
    pos[:8] = [0, 1, 2, 127, 128, len(T) - 1, len(T) - 2, 129]
    seeds["position"] = pos
    return qs, seeds


class Hip:
    """Device buffers and a stream from the HIP runtime the library itself
    links (one runtime in the process)."""

    def __init__(self):
        _lib.lib()                                  # loads libamdhip64 with the library
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def stream(self):
        s = C.c_void_p()
        assert self.L.hipStreamCreate(C.byref(s)) == 0
        return s

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(hip, st, text, qs, seeds, mat):
    L = _lib.lib()
    _declare(L)
    ql = np.array([len(q) for q in qs], np.int32)
    # queries back to back with a gap of 3 bytes between them
    qo = np.zeros(len(qs), np.int64)
    qo[1:] = np.cumsum(ql[:-1].astype(np.int64) + 3)
    buf = bytearray(int(qo[-1] + ql[-1]) + 5)
    for o, q in zip(qo, qs):
        buf[o:o + len(q)] = q
    d_q = hip.put(np.frombuffer(bytes(buf), np.uint8))
    d_qo = hip.put(qo)
    d_ql = hip.put(ql)
    d_seeds = hip.put(seeds.view(np.uint8))
    fill = np.full(len(seeds), 12345, np.int32)
    d_score = hip.put(fill)
    d_len = hip.put(fill)
    d_status = hip.put(np.array([7], np.int32))
    m = np.ascontiguousarray(mat, np.int8).reshape(-1)
    rc = L.tpz_gapless_run_noalloc(text._h, len(qs), d_q, len(buf), d_qo, d_ql, len(seeds),
                                   d_seeds, m.ctypes.data_as(C.POINTER(C.c_int8)), d_score,
                                   d_len, d_status, st)
    _lib.check(rc, "tpz_gapless_run_noalloc")
    assert hip.L.hipStreamSynchronize(st) == 0
    out = hip.get(d_score, fill), hip.get(d_len, fill), int(hip.get(d_status, fill[:1])[0])
    hip.free()
    return out


@pytest.mark.parametrize("use_sa", [False, True])
def test_noalloc_matches_batch_twice(use_sa):
    hip = Hip()
    st = hip.stream()
    rng = np.random.default_rng(77)
    lens = synth.lognormal_lengths(rng, 600, 200, 150, 1, 900)
    T = b"".join(_internal(synth.random_residues(rng, L)) + b"\n" for L in lens)
    sa = None
    if use_sa:
        Tn = np.frombuffer(T, np.uint8)
        # suffix array by sorting (small text): suffixes as byte strings
        sa = np.array(sorted(range(len(T)), key=lambda i: T[i:i + 64]), np.int64)
        assert len(sa) == len(Tn)
    text = Text(T, sa)
    mat = matrix()
    for n_q, n_s in ((300, 60000), (17, 4000)):       # second call smaller
        qs, seeds = _workload(rng, T, n_q, n_s)
        if use_sa:
            seeds["position"] = rng.integers(0, len(T), n_s)
        b = GaplessBatch(text, qs, seeds, mat).run()
        want_s, want_l = b.results()
        b.close()
        got_s, got_l, status = _noalloc(hip, st, text, qs, seeds, mat)
        assert status == 0
        assert np.array_equal(got_l, want_l)
        assert np.array_equal(got_s, want_s)
        assert (want_l != -1).sum() > (n_s // 50 if use_sa else n_s // 10)
    hip.L.hipStreamDestroy(st)


def test_noalloc_rejects_too_large():
    L = _lib.lib()
    _declare(L)
    rng = np.random.default_rng(3)
    T = _internal(synth.random_residues(rng, 1000)) + b"\n"
    text = Text(T)
    m = np.ascontiguousarray(matrix(), np.int8).reshape(-1)
    rc = L.tpz_gapless_run_noalloc(text._h, 0, None, 40 << 20, None, None, 0, None,
                                   m.ctypes.data_as(C.POINTER(C.c_int8)), None, None, None, None)
    assert rc != 0
    assert b"scratch" in L.tpz_last_error()
