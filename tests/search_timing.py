"""End-to-end `topaz search` timing on a machine with one MI355X: the stock reference
binary (CPU, the reference's own ssw.c / search.c) against TOPAZ linked with
libtopaz_amd, both as the link-level drop-in (one GPU call per seed and per
hit) and with query.c's and suffix.c's per-seed and per-hit loops replaced by
the batch entry points.  All must print the same hit lists; prints one JSON
line with the times.

    python tests/search_timing.py [--db 100000] [--queries 1000] [--threads 16] [--blasttab]

The database follows the small end-to-end configuration: lognormal lengths
(mean 350, sd 250, clipped [30, 2000]), 20 % of the queries with 1-5 planted
homologs at 30-90 % identity; search with 100 hits per query and E < 1e-3.
The binaries come from the reference sources compiled by oracle/Makefile
(the checker), so this needs `make -C oracle ref` first.
"""
import argparse
import json
import os
import resource
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

from tests.test_dropin_gpu import REF, _link_dropin, _per_query  # noqa: E402
from topaz_amd import synth  # noqa: E402


def run(cmd, cwd, timeout, out_path=None, env=None):
    """Run cmd; its stdout goes to out_path as it is printed (so a long
    search shows progress there) and is returned with the wall time, stderr
    and the child's CPU seconds (user + system)."""
    r0 = resource.getrusage(resource.RUSAGE_CHILDREN)
    t0 = time.perf_counter()
    if out_path is None:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=cwd, env=env)
        out, err, rc = r.stdout, r.stderr, r.returncode
    else:
        with open(out_path, "w") as f:
            r = subprocess.run(cmd, stdout=f, stderr=subprocess.PIPE, text=True,
                               timeout=timeout, cwd=cwd, env=env)
        out, err, rc = open(out_path).read(), r.stderr, r.returncode
    dt = time.perf_counter() - t0
    r1 = resource.getrusage(resource.RUSAGE_CHILDREN)
    cpu = (r1.ru_utime - r0.ru_utime) + (r1.ru_stime - r0.ru_stime)
    if rc != 0:
        raise SystemExit("%s failed: %s" % (cmd[0], err[-2000:]))
    return out, dt, err, cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--db", type=int, default=100000)
    ap.add_argument("--queries", type=int, default=1000)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--gpu-threads", type=int, default=None,
                    help="worker threads of the GPU-backed binaries (default: 4 x --threads); "
                         "their threads mostly wait on the GPU, so more of them than cores "
                         "keeps more requests in each launch")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for the GPU-backed binaries (HIP's default is 4; "
                         "with 8 the library spreads its calls over two streams)")
    ap.add_argument("--blasttab", action="store_true",
                    help="search with -B (start coordinates and traceback statistics per hit)")
    ap.add_argument("--timeout", type=int, default=900)
    ap.add_argument("--skip", nargs="*", default=[], help="binaries not to run (dropin_gpu, ...)")
    ap.add_argument("--out", default=None, help="directory for each search's output as printed")
    ap.add_argument("--workdir", default=None,
                    help="keep the database, index and binaries here (default: a temporary directory)")
    args = ap.parse_args()
    batch_objs = [REF / "batch" / f for f in ("query_batch.o", "suffix_batch.o")]
    if not (REF / "topaz").exists() or not all(p.exists() for p in batch_objs):
        raise SystemExit("reference checker not built (make -C oracle ref)")
    if args.out:
        os.makedirs(args.out, exist_ok=True)
    with tempfile.TemporaryDirectory() as d:
        tmp = Path(args.workdir or d)
        tmp.mkdir(parents=True, exist_ok=True)
        rng = np.random.default_rng(2027)
        t0 = time.perf_counter()
        db, queries = synth.search_database(rng, args.db, args.queries, hom_frac=0.2, max_hom=5,
                                            ident=(0.3, 0.9), hi=2000)
        synth.write_fasta(tmp / "db.fa", db, "s")
        synth.write_fasta(tmp / "q.fa", queries, "q")
        print("[search] %d subjects, %d queries written in %.1f s"
              % (len(db), len(queries), time.perf_counter() - t0), file=sys.stderr, flush=True)
        _, t_index, _, _ = run([str(REF / "topaz"), "index", "-f", str(tmp / "db.fa"),
                          "-p", str(tmp / "DB")], str(tmp), args.timeout)
        print("[search] index built in %.1f s" % t_index, file=sys.stderr, flush=True)
        exes = {"reference_cpu": REF / "topaz", "dropin_gpu": _link_dropin(tmp),
                "batch_gpu": _link_dropin(tmp, batch=True)}
        exes = {k: v for k, v in exes.items() if k not in args.skip}
        flags = ["-H", "100", "-E", "0.001"] + (["-B"] if args.blasttab else [])
        gpu_threads = args.gpu_threads or 4 * args.threads
        out, times, inner, cpu_s = {}, {}, {}, {}
        gpu_env = dict(os.environ, GPU_MAX_HW_QUEUES=str(args.hw_queues))
        for name, exe in exes.items():
            t = args.threads if name == "reference_cpu" else gpu_threads
            text, dt, err, cpu_s[name] = run([str(exe), "search", "-f", str(tmp / "q.fa"), "-p", str(tmp / "DB"),
                            "-T", str(t)] + flags, str(tmp), args.timeout,
                           os.path.join(args.out, name + ".txt") if args.out else None,
                           None if name == "reference_cpu" else gpu_env)
            out[name], times[name] = text, dt
            # TOPAZ's own clock: from the search's start, after the index is loaded
            el = [ln for ln in err.split("\n") if ln.startswith("elapsed time:")]
            inner[name] = float(el[-1].split()[2]) if el else None
            print("[search] %s: %.2f s (TOPAZ's elapsed time %s s), %d lines"
                  % (name, dt, inner[name], text.count("\n")), file=sys.stderr, flush=True)
        ref = _per_query(out["reference_cpu"])
        identical = {k: _per_query(v) == ref for k, v in out.items()}
    res = {"subjects": len(db), "queries": len(queries), "threads": args.threads,
           "gpu_threads": gpu_threads, "gpu_hw_queues": args.hw_queues,
           "flags": " ".join(flags), "index_s": round(t_index, 2),
           "search_s": {k: round(v, 2) for k, v in times.items()},
           "topaz_elapsed_s": inner,
           "cpu_s": {k: round(v, 2) for k, v in cpu_s.items()},
           "queries_per_s": {k: round(len(queries) / v, 1) for k, v in times.items()},
           "hit_lines": out["reference_cpu"].count("\n"), "identical": identical,
           "host_cpus": os.cpu_count()}
    print(json.dumps(res))
    return 0 if all(identical.values()) else 1


if __name__ == "__main__":
    sys.exit(main())
