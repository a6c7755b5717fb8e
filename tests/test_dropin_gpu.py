"""End-to-end drop-in check: TOPAZ's own search driver linked against
libtopaz_amd.so in place of ssw.o, ssw_wrapper.o, search.o and alignment.o
prints exactly the hit lists of the stock reference binary.

Both binaries come from the reference sources compiled by oracle/Makefile
into oracle/_ref/ (the checker); only the alignment and suffix-search
boundary differs.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from topaz_amd import _lib, synth

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"
OBJ = REF / "obj"
# what the library replaces: every symbol these objects define
REPLACED = {"ssw.o", "ssw_wrapper.o", "search.o", "alignment.o"}


def _link_dropin(tmp, batch=False):
    """TOPAZ's objects minus ssw.o, ssw_wrapper.o, search.o and
    alignment.o, linked against the library (sa_search, sa_hsp and the
    alignment calls then run on the GPU).  batch: query.o and suffix.o
    replaced by query_batch.o and suffix_batch.o, whose per-seed and per-hit
    loops call the batch entry points."""
    drop = REPLACED | ({"query.o", "suffix.o"} if batch else set())
    objs = sorted(str(p) for p in OBJ.rglob("*.o"))
    keep = [o for o in objs if Path(o).name not in drop]
    extra = []
    if batch:
        keep += [str(REF / "batch" / "query_batch.o"), str(REF / "batch" / "suffix_batch.o")]
        # the batch build also takes db_index from the library (db.o keeps the rest)
        db_local = str(tmp / "db_local.o")
        subprocess.run(["objcopy", "-L", "db_index", str(OBJ / "db.o"), db_local], check=True)
        keep = [o for o in keep if Path(o).name != "db.o"]
        extra.append(db_local)
    exe = tmp / ("topaz_batch" if batch else "topaz_dropin")
    libdir = str(_lib.LIB_PATH.parent)
    subprocess.run(["g++", "-o", str(exe)] + keep + extra + ["-L" + libdir,
                    "-ltopaz_amd", "-Wl,-rpath," + libdir, "-lm", "-lpthread"], check=True)
    return exe


@pytest.fixture(scope="module")
def workspace(tmp_path_factory):
    if not (REF / "topaz").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    tmp = tmp_path_factory.mktemp("dropin")
    rng = np.random.default_rng(2024)
    db, queries = synth.search_database(rng, 1500, 24, hom_frac=0.5, max_hom=4)
    synth.write_fasta(tmp / "db.fa", db, "s")
    synth.write_fasta(tmp / "q.fa", queries, "q")
    subprocess.run([str(REF / "topaz"), "index", "-f", str(tmp / "db.fa"), "-p", str(tmp / "DB")],
                   check=True, capture_output=True)
    built = all((REF / "batch" / f).exists() for f in ("query_batch.o", "suffix_batch.o"))
    batch = _link_dropin(tmp, batch=True) if built else None
    return tmp, _link_dropin(tmp), batch


def _search(exe, tmp, *flags):
    r = subprocess.run([str(exe), "search", "-f", str(tmp / "q.fa"), "-p", str(tmp / "DB"),
                        "-T", "1"] + list(flags), capture_output=True, text=True, timeout=600,
                       cwd=str(tmp))
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def _per_query(text):
    """The hit list as {query id: its lines in printed order}, plus the
    lines before the first hit.  The reference prints each query's block
    when that query finishes, and the order of the blocks varies between
    runs of the stock binary itself (the same input printed q0's block
    second or eighth, 3 runs in 30), so blocks are compared per query."""
    head, blocks = [], {}
    for line in text.splitlines():
        fields = line.split("\t")
        if len(fields) < 2:
            head.append(line)
        else:
            blocks.setdefault(fields[0], []).append(line)
    return head, blocks


@pytest.mark.parametrize("flags", [[], ["-B"], ["-F"], ["-o", "2", "-e", "4"],
                                   ["-B", "-o", "3", "-e", "3"]])
def test_hit_lists_identical(workspace, flags):
    tmp, dropin, _ = workspace
    ref = _search(REF / "topaz", tmp, *flags)
    ours = _search(dropin, tmp, *flags)
    assert ref.count("\n") > 20
    assert _per_query(ours) == _per_query(ref)


@pytest.mark.parametrize("flags", [[], ["-B"], ["-F"], ["-B", "-T", "4"],
                                   ["-B", "-o", "3", "-e", "3"]])
def test_batch_loops_hit_lists_identical(workspace, flags):
    """query.c and suffix.c with their per-seed and per-hit loops replaced by
    sa_hsp_batch (seeds of a query, and the seed walk prefetched in batches),
    align_batch_using_ssw_profile and stats_batch_using_ssw_profile
    (oracle/patch_query.py, oracle/patch_suffix.py) print the stock binary's
    hit lists, also with several worker threads calling at once."""
    tmp, _, batch = workspace
    if batch is None:
        pytest.skip("patched query.o not built (make -C oracle ref)")
    ref = _search(REF / "topaz", tmp, *flags)
    ours = _search(batch, tmp, *flags)
    assert ref.count("\n") > 20
    assert _per_query(ours) == _per_query(ref)


@pytest.fixture(scope="module")
def full_config(tmp_path_factory):
    """The small end-to-end configuration at its full size: 100,000 database
    proteins (lognormal mean 350, sd 250, clipped [30, 2000]) and 1,000
    queries, 20 % of them with 1-5 planted homologs at 30-90 % identity."""
    if not (REF / "topaz").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    if not all((REF / "batch" / f).exists() for f in ("query_batch.o", "suffix_batch.o")):
        pytest.skip("patched query.o not built (make -C oracle ref)")
    tmp = tmp_path_factory.mktemp("config1")
    rng = np.random.default_rng(2027)
    db, queries = synth.search_database(rng, 100000, 1000, hom_frac=0.2, max_hom=5,
                                        ident=(0.3, 0.9), hi=2000)
    synth.write_fasta(tmp / "db.fa", db, "s")
    synth.write_fasta(tmp / "q.fa", queries, "q")
    subprocess.run([str(REF / "topaz"), "index", "-f", str(tmp / "db.fa"), "-p", str(tmp / "DB")],
                   check=True, capture_output=True, timeout=600)
    return tmp, _link_dropin(tmp, batch=True)


@pytest.mark.parametrize("flags", [[], ["-B"], ["-F"]])
def test_batch_loops_full_size_config(full_config, flags):
    """The batch build against the stock binary on the whole 100,000 x 1,000
    configuration (-H 100 -E 1e-3), default, -B and -F; both with 16 worker
    threads, the GPU build's calls spread over 8 hardware queues."""
    tmp, batch = full_config
    common = ["-H", "100", "-E", "0.001", "-T", "16"] + list(flags)
    ref = _search(REF / "topaz", tmp, *common)
    env = dict(os.environ, GPU_MAX_HW_QUEUES="8")
    r = subprocess.run([str(batch), "search", "-f", str(tmp / "q.fa"), "-p", str(tmp / "DB")] +
                       common, capture_output=True, text=True, timeout=600, cwd=str(tmp), env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    head, blocks = _per_query(ref)
    assert len(blocks) > 150 and sum(len(v) for v in blocks.values()) > 400
    assert _per_query(r.stdout) == (head, blocks)


def _link_engine(tmp):
    """TOPAZ's objects with engine.o replaced by engine_batch.o, whose
    workers search batches of queries through tpz_search
    (oracle/patch_engine.py), linked against the library."""
    if not (REF / "batch" / "engine_batch.o").exists():
        pytest.skip("patched engine.o not built (make -C oracle ref)")
    objs = [o for o in sorted(str(p) for p in OBJ.rglob("*.o")) if Path(o).name != "engine.o"]
    exe = tmp / "topaz_search_gpu"
    libdir = str(_lib.LIB_PATH.parent)
    subprocess.run(["g++", "-o", str(exe)] + objs + [str(REF / "batch" / "engine_batch.o"),
                    "-L" + libdir, "-ltopaz_amd", "-Wl,-rpath," + libdir, "-lm", "-lpthread"],
                   check=True)
    return exe


def _run(exe, tmp, args, env=None):
    r = subprocess.run([str(exe), "search", "-f", str(tmp / "q.fa"), "-p", str(tmp / "DB")] +
                       list(args), capture_output=True, text=True, timeout=900, cwd=str(tmp),
                       env=dict(os.environ, **(env or {})))
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("flags,env", [
    ([], {}), (["-B"], {}), (["-F"], {}),
    (["-T", "3"], {"TOPAZ_AMD_SEARCH_BATCH": "5"}),      # several batches at once
    (["-B"], {"TOPAZ_AMD_SEARCH_PLAIN": "1"}),           # one query at a time
    (["-B", "-o", "3", "-e", "3"], {}),                  # gap_open <= gap_extend: the lanes
    (["-o", "1", "-e", "2"], {}),
])
def test_search_engine_hit_lists_identical(workspace, flags, env):
    """TOPAZ with its workers searching whole batches through tpz_search
    (suffix-array search, stream expansion, walk, SW, statistics) prints the
    stock binary's hit lists."""
    tmp = workspace[0]
    exe = _link_engine(tmp)
    ref_flags = []
    for k, f in enumerate(flags):
        if f == "-T" or (k > 0 and flags[k - 1] == "-T"):
            continue
        ref_flags.append(f)
    ref = _search(REF / "topaz", tmp, *ref_flags)
    ours = _run(exe, tmp, flags, env)
    assert ref.count("\n") > 20
    assert _per_query(ours) == _per_query(ref)


@pytest.mark.parametrize("flags", [[], ["-B"], ["-F"]])
def test_search_engine_full_size_config(full_config, flags):
    """tpz_search through TOPAZ's own driver on the whole 100,000 x 1,000
    configuration (-H 100 -E 1e-3), default, -B and -F, against the stock
    binary; batches of 256 queries from 2 worker threads."""
    tmp, _ = full_config
    exe = _link_engine(tmp)
    common = ["-H", "100", "-E", "0.001"] + list(flags)
    ref = _search(REF / "topaz", tmp, *(common + ["-T", "16"]))
    ours = _run(exe, tmp, common + ["-T", "2"], {"TOPAZ_AMD_SEARCH_BATCH": "256"})
    head, blocks = _per_query(ref)
    assert len(blocks) > 150
    assert _per_query(ours) == (head, blocks)


def _link_index(tmp):
    """TOPAZ's objects minus divsufsort.o, sssort.o and trsort.o (the
    vendored libdivsufsort), linked against the library's divsufsort64."""
    drop = {"divsufsort.o", "sssort.o", "trsort.o"}
    objs = [o for o in sorted(str(p) for p in OBJ.rglob("*.o")) if Path(o).name not in drop]
    exe = tmp / "topaz_index_gpu"
    libdir = str(_lib.LIB_PATH.parent)
    subprocess.run(["g++", "-o", str(exe)] + objs + ["-L" + libdir, "-ltopaz_amd",
                    "-Wl,-rpath," + libdir, "-lm", "-lpthread"], check=True)
    return exe


def test_index_files_identical(tmp_path):
    """`topaz index` with the suffix array built on the GPU writes the stock
    binary's index files byte for byte (.SA included), and a search over
    that index prints the same hit lists."""
    if not (REF / "topaz").exists():
        pytest.skip("reference checker not built (make -C oracle ref)")
    rng = np.random.default_rng(2031)
    db, queries = synth.search_database(rng, 3000, 12, hom_frac=0.5, max_hom=4)
    db += [db[k] for k in rng.integers(0, len(db), 50)]     # identical sequences
    synth.write_fasta(tmp_path / "db.fa", db, "s")
    synth.write_fasta(tmp_path / "q.fa", queries, "q")
    exe = _link_index(tmp_path)
    for name, binary in (("ref", REF / "topaz"), ("gpu", exe)):
        (tmp_path / name).mkdir()
        subprocess.run([str(binary), "index", "-f", str(tmp_path / "db.fa"), "-p",
                        str(tmp_path / name / "DB")], check=True, capture_output=True,
                       cwd=str(tmp_path), timeout=300)
    files = sorted(p.name for p in (tmp_path / "ref").iterdir())
    assert "DB.SA" in files and files == sorted(p.name for p in (tmp_path / "gpu").iterdir())
    for f in files:
        assert (tmp_path / "ref" / f).read_bytes() == (tmp_path / "gpu" / f).read_bytes(), f
    out = {}
    for name in ("ref", "gpu"):
        r = subprocess.run([str(REF / "topaz"), "search", "-f", str(tmp_path / "q.fa"), "-p",
                            str(tmp_path / name / "DB"), "-T", "1"], capture_output=True,
                           text=True, timeout=300, cwd=str(tmp_path))
        assert r.returncode == 0
        out[name] = r.stdout
    # per query: the reference prints the queries' blocks in varying order
    assert out["gpu"].count("\n") >= 3 and _per_query(out["gpu"]) == _per_query(out["ref"])
