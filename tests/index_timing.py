"""`topaz index` timing on a machine with one MI355X: the stock reference binary (its own
libdivsufsort, one core) against TOPAZ linked with libtopaz_amd's
divsufsort64 (suffix array built on the GPU).  Both must write identical
index files; prints one JSON line.

    python tests/index_timing.py [--db 1000000]

The database follows the end-to-end configurations: lognormal protein
lengths (mean 350, sd 250, clipped [30, 2000]), with 1 % of the proteins
repeated exactly, as real databases hold identical sequences.  The
binaries come from the reference sources compiled by oracle/Makefile (the
checker), so this needs `make -C oracle ref` first.
"""
import argparse
import filecmp
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

from tests.test_dropin_gpu import REF, _link_index  # noqa: E402
from topaz_amd import synth  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--db", type=int, default=1000000)
    ap.add_argument("--timeout", type=int, default=900)
    ap.add_argument("--workdir", default=None, help="keep the files here (default: a temporary directory)")
    args = ap.parse_args()
    if not (REF / "topaz").exists():
        raise SystemExit("reference checker not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as d:
        tmp = Path(args.workdir or d)
        tmp.mkdir(parents=True, exist_ok=True)
        rng = np.random.default_rng(2029)
        t0 = time.perf_counter()
        db, _ = synth.search_database(rng, args.db, 1, hom_frac=0.0, hi=2000)
        db += [db[k] for k in rng.integers(0, len(db), max(1, args.db // 100))]
        synth.write_fasta(tmp / "db.fa", db, "s")
        print("[index] %d proteins written in %.1f s" % (len(db), time.perf_counter() - t0),
              file=sys.stderr, flush=True)
        exe = _link_index(tmp)
        times = {}
        for name, binary in (("reference_cpu", REF / "topaz"), ("gpu_sa", exe)):
            (tmp / name).mkdir(exist_ok=True)
            t0 = time.perf_counter()
            r = subprocess.run([str(binary), "index", "-f", str(tmp / "db.fa"), "-p",
                                str(tmp / name / "DB")], capture_output=True, text=True,
                               timeout=args.timeout, cwd=str(tmp))
            times[name] = time.perf_counter() - t0
            if r.returncode != 0:
                raise SystemExit("%s index failed: %s" % (name, r.stderr[-2000:]))
            print("[index] %s: %.2f s" % (name, times[name]), file=sys.stderr, flush=True)
        files = sorted(p.name for p in (tmp / "reference_cpu").iterdir())
        identical = files == sorted(p.name for p in (tmp / "gpu_sa").iterdir()) and all(
            filecmp.cmp(tmp / "reference_cpu" / f, tmp / "gpu_sa" / f, shallow=False)
            for f in files)
        text_bytes = os.path.getsize(tmp / "reference_cpu" / "DB.psq")
    res = {"proteins": len(db), "text_bytes": text_bytes,
           "index_s": {k: round(v, 2) for k, v in times.items()},
           "speedup": round(times["reference_cpu"] / times["gpu_sa"], 2),
           "files_identical": identical, "files": files}
    print(json.dumps(res))
    return 0 if identical else 1


if __name__ == "__main__":
    sys.exit(main())
