"""TOPAZ's index files written from proteins held in memory, with the suffix
array built on the GPU: what `topaz index` writes (demux.c:15-49,
index.c:24-172), without going through a FASTA file.

  .psq  the proteins in the internal alphabet (aa2internal,
        lexicographical.h:135-192), each followed by '\\n' (seq_write,
        seq.c:131-144): the text T of the suffix array
  .phr  each protein's id and '\\n'
  .pin  per protein "%ld %d %ld %d\\n": start in .psq, length, start in
        .phr, id length (demux.c:29)
  .ind  position_t per protein (db.h:10-17): header_start, protein_start
        (uint64), header_len, protein_len (int32), 24 bytes
  .SA   the suffix array of T, int64 per byte (index.c:47-69); built by
        tpz_suffix_array into the memory-mapped file
  .ISA  a file of the same size that `topaz index` maps and never writes
        (index.c:57-61, without USE_LCP): created here at that size, empty
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from .matrices import INTERNAL_ORDER, MATRIX_ORDER

IND_DTYPE = np.dtype([("header_start", "<u8"), ("protein_start", "<u8"), ("header_len", "<i4"),
                      ("protein_len", "<i4")])

# matrix index -> the internal letter aa2internal gives its amino-acid letter
INTERNAL_OF_INDEX = np.array([ord("A") + INTERNAL_ORDER.index(MATRIX_ORDER[k]) for k in range(25)],
                             np.uint8)


class SaStats(C.Structure):
    _fields_ = [("rounds", C.c_int32), ("device_bytes", C.c_int64), ("alloc_s", C.c_double),
                ("sort_s", C.c_double), ("readback_s", C.c_double)]


def text_of(res, lens, chunk=1 << 26):
    """The .psq text of proteins given as matrix indices stored back to
    back (res) with lengths lens: internal letters, '\\n' after each."""
    lens = np.asarray(lens, np.int64)
    T = np.empty(int(lens.sum()) + len(lens), np.uint8)
    ends = np.cumsum(lens)
    k0 = 0
    while k0 < len(lens):
        r0 = int(ends[k0 - 1]) if k0 else 0
        k1 = min(len(lens), max(k0 + 1, int(np.searchsorted(ends, r0 + chunk, side="right"))))
        L = lens[k0:k1]
        n = int(L.sum())
        inc = np.zeros(n + 1, np.int32)
        starts = np.cumsum(L) - L
        inc[starts[1:]] = 1                   # one '\n' before every protein after the first
        dest = np.cumsum(inc[:n]) + np.arange(n, dtype=np.int64)
        seg = T[r0 + k0:r0 + k0 + n + (k1 - k0)]
        seg[dest] = INTERNAL_OF_INDEX[np.asarray(res[r0:r0 + n], np.int64)]
        seg[starts + L + np.arange(k1 - k0)] = ord("\n")
        k0 = k1
    return T


def write_index(prefix, text, lens, ids, sa=None):
    """Write TOPAZ's index files for the proteins whose .psq text is `text`
    (uint8, internal letters, '\\n' after each protein), lengths `lens` and
    ids `ids` (str).  The suffix array comes from the GPU (tpz_suffix_array)
    unless `sa` is given.  Returns the suffix-array build statistics (None
    when sa was given)."""
    prefix = str(prefix)
    text = np.ascontiguousarray(text, np.uint8)
    lens = np.asarray(lens, np.int64)
    n = len(text)
    if int(lens.sum()) + len(lens) != n:
        raise ValueError("the text does not hold the proteins and their newlines")
    starts = np.cumsum(lens + 1) - (lens + 1)
    if len(lens) and not np.all(text[starts + lens] == ord("\n")):
        raise ValueError("a protein is not followed by a newline")
    text.tofile(prefix + ".psq")
    id_bytes = [i.encode() for i in ids]
    hl = np.array([len(b) for b in id_bytes], np.int64)
    hstart = np.cumsum(hl + 1) - (hl + 1)
    with open(prefix + ".phr", "wb") as f:
        f.write(b"".join(b + b"\n" for b in id_bytes))
    with open(prefix + ".pin", "w") as f:
        f.write("".join("%d %d %d %d\n" % t for t in zip(starts.tolist(), lens.tolist(),
                                                         hstart.tolist(), hl.tolist())))
    ind = np.zeros(len(lens), IND_DTYPE)
    ind["header_start"] = hstart
    ind["protein_start"] = starts
    ind["header_len"] = hl
    ind["protein_len"] = lens
    ind.tofile(prefix + ".ind")
    st = None
    out = np.memmap(prefix + ".SA", np.int64, "w+", shape=(n,)) if n else None
    if sa is not None:
        out[:] = sa
    elif n:
        L = _lib.lib()
        L.tpz_suffix_array.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(SaStats)]
        L.tpz_suffix_array.restype = C.c_int
        st = SaStats()
        _lib.check(L.tpz_suffix_array(text.ctypes.data, n, out.ctypes.data, C.byref(st)),
                   "tpz_suffix_array")
    if out is not None:
        out.flush()
        del out
    else:
        open(prefix + ".SA", "wb").close()
    with open(prefix + ".ISA", "wb") as f:
        f.truncate(8 * n)
    return st
