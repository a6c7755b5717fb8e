"""The C-ABI library loads and exports every entry point its headers declare
(no compute call: this runs without a GPU)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from topaz_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # (a function-pointer typedef, "type (*name)(...)", declares no function)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\((?!\s*\*)[^;{]*\)\s*;", text):
            name = m.group(1)
            if name in ("if", "for", "while", "return", "sizeof"):
                continue
            names.append((h.name, name))
    return names


def test_headers_declare_something():
    assert len(declared_functions()) >= 8


@pytest.mark.parametrize("header,name", declared_functions())
def test_symbol_exported(header, name):
    L = C.CDLL(str(_lib.LIB_PATH))
    assert hasattr(L, name), "%s declares %s but the library does not export it" % (header, name)


def test_version_string():
    assert b"gfx950" in _lib.lib().tpz_version()


def test_cigar_helpers_of_the_compat_header(tmp_path):
    """The header's inline CIGAR helpers (ssw.h:148-194) follow the BAM
    encoding of the SAM specification: ops M I D N S H P = X are 0..8,
    length in the bits above 4; unknown letters encode as M and codes above
    8 decode as M."""
    src = tmp_path / "cigar.c"
    src.write_text(r'''
#include <stdio.h>
#include "topaz_ssw_compat.h"
int main(void) {
    const char* letters = "MIDNSHP=XQm";
    for (const char* c = letters; *c; ++c) {
        uint32_t v = to_cigar_int(37u, *c);
        printf("%c %u %c %u\n", *c, v, cigar_int_to_op(v), cigar_int_to_len(v));
    }
    for (uint32_t op = 9; op < 16; ++op) printf("%c", cigar_int_to_op((7u << 4) | op));
    printf("\n");
    return 0;
}
''')
    exe = tmp_path / "cigar"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    spec = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4, "H": 5, "P": 6, "=": 7, "X": 8}
    for line, letter in zip(out, "MIDNSHP=XQm"):
        c, v, back, n = line.split(" ")
        op = spec.get(letter, 0)
        assert (c, int(v), back, int(n)) == (letter, 37 * 16 + op, "MIDNSHP=X"[op], 37)
    assert out[11] == "M" * 7


BATCH = ROOT / "oracle" / "_ref" / "batch"


@pytest.mark.skipif(not (BATCH / "query_batch.o").exists() or not (BATCH / "suffix_batch.o").exists(),
                    reason="patched TOPAZ objects not built (make -C oracle ref)")
def test_patched_topaz_objects_call_the_batch_entry_points():
    """oracle/patch_query.py and patch_suffix.py applied to TOPAZ's query.c
    and suffix.c: the objects call the batch entry points the compat header
    declares, and query_batch.o no longer calls the per-hit or per-seed ones
    from its loops (perform_alignment_with_profile is still defined there,
    only no longer called by query_perform_alignments)."""
    def symbols(obj):
        out = subprocess.run(["nm", str(obj)], check=True, capture_output=True, text=True).stdout
        undef = {l.split()[-1] for l in out.splitlines() if l.split()[0] == "U"}
        defined = {l.split()[-1] for l in out.splitlines() if len(l.split()) == 3}
        return undef, defined
    qu, qd = symbols(BATCH / "query_batch.o")
    su, sd = symbols(BATCH / "suffix_batch.o")
    assert {"align_batch_using_ssw_profile", "stats_batch_using_ssw_profile", "sa_hsp_batch",
            "sa_search_batch", "suffix_prefetched"} <= qu
    assert not {"stats_using_ssw_profile", "sa_hsp", "sa_search"} & qu
    assert {"query_find_suffixes", "query_perform_alignments", "query_perform_statistics"} <= qd
    assert "sa_hsp_batch" in su and "suffix_prefetched" in sd
    assert {"suffix_alloc", "suffix_increment", "suffix_duplicate", "suffix_free"} <= sd
    lib = C.CDLL(str(_lib.LIB_PATH))
    for name in ("align_batch_using_ssw_profile", "stats_batch_using_ssw_profile", "sa_hsp_batch",
                 "sa_search_batch", "db_index", "divsufsort64"):
        assert hasattr(lib, name)


def test_db_index_equals_reference():
    """db_index (host code, no GPU): the protein holding each text position,
    from the table of every 64th position, equals the reference's own binary
    search (db.c:176-205, compiled from its sources) at every position of a
    tiled index (short, long and empty proteins); an index that does not
    tile the text takes the binary search step for step."""
    from tests.refcheck import REF_SO, DbT
    from tests.test_compat_gpu import PositionT
    if not REF_SO.exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    import numpy as np
    L, R = C.CDLL(str(_lib.LIB_PATH)), C.CDLL(str(REF_SO))
    for f in (L.db_index, R.db_index):
        f.argtypes = [C.POINTER(DbT), C.c_uint64, C.POINTER(C.c_int64)]
        f.restype = C.c_int
    rng = np.random.default_rng(81)
    lens = np.concatenate([rng.integers(0, 5, 40), rng.integers(30, 700, 300),
                           rng.integers(0, 3, 20)]).astype(np.int64)
    rng.shuffle(lens)
    starts = np.concatenate([[0], np.cumsum(lens + 1)[:-1]])
    for tiled in (True, False):
        index = (PositionT * len(lens))()
        for k, (s0, n) in enumerate(zip(starts, lens)):
            index[k].protein_start = int(s0) + (0 if tiled or k % 7 else 3)
            index[k].protein_len = int(n)
        db = DbT()
        db.index = C.cast(index, C.c_void_p)
        db.Psize = len(lens)
        db.Tsize = int(starts[-1] + lens[-1] + 1)
        for pos in list(range(db.Tsize + 2)):
            a, b = C.c_int64(-5), C.c_int64(-5)
            ra, rb = L.db_index(C.byref(db), pos, C.byref(a)), R.db_index(C.byref(db), pos, C.byref(b))
            assert (ra, a.value) == (rb, b.value), (tiled, pos)


def test_suffix_array_argument_checks_match_the_reference():
    """divsufsort64's argument checks (divsufsort.c:338-339) come before any
    device work: NULL text or array or n < 0 give -1, n = 0 gives 0.  The
    library returns the same codes as the reference's own compiled
    divsufsort64, and tpz_suffix_array rejects texts too long for 32-bit
    positions with a message, without a GPU."""
    from tests.refcheck import REF_SO
    L = _lib.lib()
    L.divsufsort64.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
    L.divsufsort64.restype = C.c_int32
    L.tpz_suffix_array.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p]
    L.tpz_suffix_array.restype = C.c_int
    sa = (C.c_int64 * 4)()
    cases = [(None, sa, 3), (b"abc", None, 3), (b"abc", sa, -1), (b"abc", sa, 0), (None, None, 0)]
    ours = [L.divsufsort64(t, a, n) for t, a, n in cases]
    assert ours == [-1, -1, -1, 0, -1]
    if REF_SO.exists():
        R = C.CDLL(str(REF_SO))
        R.divsufsort64.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.c_int64]
        R.divsufsort64.restype = C.c_int32
        assert [R.divsufsort64(t, a, n) for t, a, n in cases] == ours
    assert L.tpz_suffix_array(None, 3, sa, None) == -1
    assert b"bad arguments" in L.tpz_last_error()
    assert L.tpz_suffix_array(b"abc", 1 << 32, sa, None) == -1
    assert b"32-bit" in L.tpz_last_error()
    assert L.tpz_suffix_array(b"", 0, sa, None) == 0
