"""topaz_amd.index.write_index writes `topaz index`'s files byte for byte.

The reference's own indexer (oracle/_ref/topaz index, index.c:24-172 and
demux.c:15-49 over a FASTA file) and write_index (from the same proteins in
memory) must produce identical .psq, .phr, .pin, .ind, .SA and .ISA files.
On the CPU the suffix array handed to write_index is the oracle's (itself
pinned against the reference's libdivsufsort); the GPU test builds it with
tpz_suffix_array.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.refcheck import Oracle
from topaz_amd import index, synth

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"
FILES = (".psq", ".phr", ".pin", ".ind", ".SA", ".ISA")


def _reference_and_ours(tmp, n_db, gpu):
    if not (REF / "topaz").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    rng = np.random.default_rng(55)
    (res, off, lens), _ = synth.write_search_fasta(tmp / "db.fa", tmp / "q.fa", rng, n_db, 20,
                                                    hom_frac=0.5, hi=3000)
    subprocess.run([str(REF / "topaz"), "index", "-f", str(tmp / "db.fa"), "-p",
                    str(tmp / "REF")], check=True, capture_output=True, timeout=600)
    text = index.text_of(res, lens)
    ids = ["s%d" % k for k in range(len(lens))]
    sa = None if gpu else Oracle().suffix_array(text.tobytes())
    index.write_index(tmp / "OURS", text, lens, ids, sa=sa)
    for ext in FILES:
        a = (tmp / ("REF" + ext)).read_bytes()
        b = (tmp / ("OURS" + ext)).read_bytes()
        assert a == b, ext


def test_index_files_identical_cpu(tmp_path):
    _reference_and_ours(tmp_path, 2000, gpu=False)


@pytest.mark.gpu
def test_index_files_identical_gpu(tmp_path):
    _reference_and_ours(tmp_path, 30000, gpu=True)


def test_text_of_matches_letters():
    rng = np.random.default_rng(3)
    lens = np.array([1, 59, 60, 61, 300, 2], np.int64)
    res = synth.random_residues(rng, int(lens.sum()))
    text = index.text_of(res, lens, chunk=100)
    want = b"".join(bytes(index.INTERNAL_OF_INDEX[res[o:o + n].astype(np.int64)]) + b"\n"
                    for o, n in zip(np.cumsum(lens) - lens, lens))
    assert text.tobytes() == want
