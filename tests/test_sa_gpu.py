"""Suffix array of the database text on the GPU (tpz_suffix_array and the
divsufsort64 drop-in) against the reference's own divsufsort64, compiled
from its sources by oracle/Makefile, and against brute-force sorting of the
suffixes for small texts.  The suffix array of a string is unique, so all
three must agree entry for entry."""
import ctypes as C

import numpy as np
import pytest

from tests.refcheck import REF_SO
from topaz_amd import _lib, synth

pytestmark = pytest.mark.gpu


class SaStats(C.Structure):
    _fields_ = [("rounds", C.c_int32), ("device_bytes", C.c_int64), ("alloc_s", C.c_double),
                ("sort_s", C.c_double), ("readback_s", C.c_double)]


def _gpu_sa(text, via_dropin=False):
    L = _lib.lib()
    L.tpz_suffix_array.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_int64),
                                   C.POINTER(SaStats)]
    L.tpz_suffix_array.restype = C.c_int
    L.divsufsort64.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
    L.divsufsort64.restype = C.c_int32
    sa = np.full(max(len(text), 1), -7, np.int64)
    st = SaStats()
    if via_dropin:
        assert L.divsufsort64(text, sa.ctypes.data_as(C.POINTER(C.c_int64)), len(text)) == 0
    else:
        rc = L.tpz_suffix_array(text, len(text), sa.ctypes.data_as(C.POINTER(C.c_int64)),
                                C.byref(st))
        assert rc == 0, _lib.lib().tpz_last_error()
    return sa[:len(text)], st


def _ref_sa(text):
    if not REF_SO.exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    R = C.CDLL(str(REF_SO))
    R.divsufsort64.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
    R.divsufsort64.restype = C.c_int32
    sa = np.zeros(max(len(text), 1), np.int64)
    assert R.divsufsort64(text, sa.ctypes.data_as(C.POINTER(C.c_int64)), len(text)) == 0
    return sa[:len(text)]


def _brute_sa(text):
    return np.array(sorted(range(len(text)), key=lambda i: text[i:]), np.int64)


@pytest.mark.parametrize("text", [b"", b"a", b"ab", b"ba", b"aa", b"aaa", b"abab", b"banana",
                                  b"mississippi", bytes([255, 0, 255, 0, 1]), b"\n\n\n",
                                  # one repeated pair of ten: the later rounds sort the
                                  # repeated suffixes only
                                  b"abcdefghab", b"xyzxyzxyzxyzxyzxyzq", b"abracadabra" * 3])
def test_tiny_texts_brute_force(text):
    sa, _ = _gpu_sa(text)
    assert sa.tolist() == _brute_sa(text).tolist()
    if text:
        assert _gpu_sa(text, via_dropin=True)[0].tolist() == sa.tolist()


def test_random_and_repetitive_texts_against_reference():
    rng = np.random.default_rng(51)
    block = rng.integers(0, 256, 1500, dtype=np.uint8).tobytes()
    cases = {
        "random bytes": rng.integers(0, 256, 200_000, dtype=np.uint8).tobytes(),
        "two letters": rng.integers(65, 67, 100_000, dtype=np.uint8).tobytes(),
        "one letter": b"A" * 5000,
        "period 2": b"AB" * 4000,
        "block x 40": block * 40,                 # repeats 58,500 bytes long
    }
    for name, text in cases.items():
        sa, st = _gpu_sa(text)
        assert np.array_equal(sa, _ref_sa(text)), name
        if len(text) <= 5000:
            assert np.array_equal(sa, _brute_sa(text)), name
    # 'A' * 5000: ranks all distinct only once the prefix length reaches 5000
    st = _gpu_sa(b"A" * 5000)[1]
    assert st.rounds == 13
    assert min(st.alloc_s, st.sort_s, st.readback_s) >= 0


def test_protein_text_against_reference():
    """A TOPAZ .psq-like text of 10 M bytes: internal letters with '\\n' after
    each protein, mutated copies of queries, and 200 exact copies of
    proteins at the end (databases hold identical sequences), so that
    repeats run to thousands of bytes."""
    rng = np.random.default_rng(52)
    text, _, _, _ = synth.internal_text_and_seeds(rng, 10_000_000, 10, 10)
    t = text.tobytes()
    prots = t.split(b"\n")[:-1]
    t += b"".join(prots[k] + b"\n" for k in rng.integers(0, len(prots), 200))
    sa, st = _gpu_sa(t)
    assert st.rounds >= 10
    assert np.array_equal(sa, _ref_sa(t))


def _distinct_pairs(t):
    b = np.frombuffer(t, np.uint8).astype(np.int64)
    nxt = np.r_[b[1:] + 1, 0]
    return len(np.unique((b + 1) * 257 + nxt))


@pytest.mark.parametrize("target", [511, 512])
def test_byte_pair_count_at_the_bitmap_limit(target):
    """Round 2 reads four-byte ranks off a bitmap when the text has at most
    511 distinct byte pairs (the last byte with the end counted), and sorts
    otherwise.  Texts with exactly 511 and 512 pairs: 22 letters in random
    order (485 pairs), then one fresh byte value per missing pair (each
    adds the pairs (last, fresh) and (fresh, end) and drops (last, end))."""
    rng = np.random.default_rng(53 + target)
    letters = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWYBZ", np.uint8)
    t = bytes(letters[rng.integers(0, len(letters), 200_000)])
    assert _distinct_pairs(t) == len(letters) ** 2 + 1
    fresh = iter(range(128, 256))
    while _distinct_pairs(t) < target:
        t += bytes([next(fresh)])
    assert _distinct_pairs(t) == target
    t = t[:100_000] + t[:5_000] + t[100_000:]       # a 5,000-byte repeat
    assert _distinct_pairs(t) == target
    sa, st = _gpu_sa(t)
    assert np.array_equal(sa, _ref_sa(t))
