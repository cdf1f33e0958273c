"""bench.py's constants and CLI contract (CPU only)."""
import json
import os
import subprocess
import sys

from tests.conftest import ROOT


def test_algorithmic_flop_constant_matches_op_count(orc):
    import bench
    oc = orc.opcount()
    assert bench.ALGO_FLOPS_PER_CELL == oc["ALGO_FLOPS_PER_CELL"]
    assert bench.ALGO_FLOPS_PER_DIRECTION == oc["per_direction"]
    assert bench.ALGO_FLOPS_PER_CELL_SHARED_EPS == oc["ALGO_FLOPS_PER_CELL_SHARED_EPS"]


def test_multi_gpu_request_exits_with_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in json.loads(p.stdout.strip().splitlines()[-1])["error"]
