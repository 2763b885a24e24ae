"""bench.py's host-side helpers, on the CPU: the launch-shape key, the traffic and limiter lines it
reads from the committed rocprofv3 summaries, and the refusal of more than one GPU."""
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_shape_key_orders_the_launch_fields():
    launch = {"gridBlocksClosest": 256, "blockThreads": 512, "topNodes": 0, "refillMin": 8,
              "helperWaves": 4, "raysPerFetch": 128, "stackLdsEntries": 24, "other": 1}
    assert bench.shape_key(launch) == "256,512,0,8,4,128,24"


def test_traffic_is_read_only_for_a_profiled_shape():
    for lds in (8, 16, 24, 32):
        t = bench.profiled_traffic(f"256,512,0,8,4,128,{lds}")
        # HBM bytes per launch: well below the 1.75 GB of algorithmic bytes, above the 32 MB of rays
        assert isinstance(t, int) and 32_000_000 < t < 1_754_000_000
    assert bench.profiled_traffic("256,512,0,8,4,128,20") is None
    assert bench.profiled_traffic("1,2,3,4,5,6,7") is None


def test_limiter_lines_are_present():
    for path in (bench.PROFILE, bench.FUSED_PROFILE):
        text = bench.profile_line("limiter:", path)
        assert text and len(text) > 20


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr
