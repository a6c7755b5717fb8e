"""SEG masking (tpz_seg) against the reference's own run_seg
(seg_wrapper.c:15-49, compiled from its sources into
oracle/_ref/libtopazref.so), byte for byte.

Host logic: no device is used.  The sequences mix random residues with
planted low-complexity stretches (single-letter runs, short tandem repeats,
two- and three-letter mixtures), lower-case letters, letters outside the
20 standard ones, sequences shorter than the 12-letter window, and stretches
longer than the 100 letters a trim may remove.
"""
import ctypes as C

import numpy as np
import pytest

from tests.refcheck import REF_SO, SeqT
from topaz_amd import _lib

STD = b"ACDEFGHIKLMNPQRSTVWY"


def _ref():
    if not REF_SO.exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    L = C.CDLL(str(REF_SO))
    L.initialise_seg.restype = None
    L.run_seg.argtypes = [C.POINTER(SeqT)]
    L.run_seg.restype = None
    L.initialise_seg()
    return L


def _ours():
    L = _lib.lib()
    L.tpz_seg.argtypes = [C.c_char_p, C.c_int32]
    L.tpz_seg.restype = C.c_int32
    return L


def ref_seg(L, seq):
    buf = C.create_string_buffer(bytes(seq), len(seq) + 1)
    s = SeqT()
    s.seq_buf.buf = C.cast(buf, C.c_char_p)
    s.seq_buf.len = len(seq) + 1
    s.seq_len = len(seq)
    L.run_seg(C.byref(s))
    return buf.raw[:len(seq)]


def our_seg(L, seq):
    buf = C.create_string_buffer(bytes(seq), len(seq) + 1)
    n = L.tpz_seg(buf, len(seq))
    out = buf.raw[:len(seq)]
    # n counts the masked positions, some of which may have held 'X' already
    assert sum(a != b for a, b in zip(seq, out)) <= n <= out.count(b"X")
    return out


def pick(rng, letters, m):
    return np.array(rng.choice(list(letters), m), np.uint8).tobytes()


def sequences(rng, n):
    out = []
    for _ in range(n):
        parts = []
        for _ in range(int(rng.integers(1, 6))):
            kind = rng.integers(0, 7)
            m = int(rng.integers(1, 90))
            if kind == 0:
                parts.append(pick(rng, STD, m))
            elif kind == 1:                      # one letter
                parts.append(bytes([STD[rng.integers(0, 20)]]) * m)
            elif kind == 2:                      # a short tandem repeat
                unit = pick(rng, STD, int(rng.integers(2, 5)))
                parts.append((unit * (m // len(unit) + 1))[:m])
            elif kind == 3:                      # a two- or three-letter mixture
                alpha = list(rng.choice(list(STD), int(rng.integers(2, 4)), replace=False))
                parts.append(pick(rng, alpha, m))
            elif kind == 4:                      # lower case
                parts.append(pick(rng, STD, m).lower())
            elif kind == 5:                      # letters outside the 20
                parts.append(pick(rng, b"XBZJUO*", m))
            else:                                # a long biased stretch
                alpha = list(rng.choice(list(STD), 4, replace=False))
                parts.append(pick(rng, alpha, m * 4))
        out.append(b"".join(parts))
    return out


def test_seg_matches_reference():
    R, O = _ref(), _ours()
    rng = np.random.default_rng(11)
    seqs = sequences(rng, 2500)
    seqs += [b"", b"A", b"ACDEFGHIKLM", b"Q" * 11, b"Q" * 12, b"Q" * 13, b"P" * 400,
             (b"AS" * 300), pick(rng, STD, 3000)]
    masked = 0
    for s in seqs:
        a, b = ref_seg(R, s), our_seg(O, s)
        assert a == b, s
        masked += a.count(b"X") - s.count(b"X")
    assert masked > 5000


def test_seg_rejects_bad_arguments():
    O = _ours()
    assert O.tpz_seg(None, 5) == -1
    assert O.tpz_seg(b"AAAA", -1) == -1
    assert O.tpz_seg(None, 0) == 0
