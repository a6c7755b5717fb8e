"""bench.py's contract pieces that run without a GPU."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_measured_traffic_record():
    for key in ("sw_forward_2Mpairs", "gapless_uniform_50M", "gapless_matched_50M"):
        v = bench.measured_traffic(key)
        assert isinstance(v, int) and v > 10**9, key
    assert bench.measured_traffic("no such kernel") is None
