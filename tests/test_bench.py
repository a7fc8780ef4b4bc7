"""bench.py's host side on the CPU: the figures it reads from the committed
rocprofv3 summaries, its algorithmic byte count and its one-GPU guard."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

NOPS = 1048576 * 16  # the throughput workload: 1,048,576 transactions x 16 ops


def test_profile_traffic_per_launch():
    t = bench.traffic_per_launch()
    assert t is not None, "profiles/txn_exec_config2_pmc.md holds traffic_bytes_per_launch"
    # every op moves at least one 64-byte line, and no more than a few dozen
    assert 64 * NOPS <= t <= 64 * 64 * NOPS


def test_profile_atomics_per_launch():
    a = bench.atomics_per_launch()
    assert a is not None, "profiles/txn_exec_config2_atomics.md holds atomics_per_launch"
    # the eight protocols issue between 0.1 (Silo) and 2 (2PL) atomics per op
    assert 0.1 * NOPS <= a <= 2.5 * NOPS


def test_algorithmic_bytes():
    # per transaction 8 B; per op 8 + 8 + 16 + 16 B; per read 8 B
    assert bench.algorithmic_bytes(0, 0, 0) == 0
    assert bench.algorithmic_bytes(1, 16, 14) == 8 + 16 * 48 + 14 * 8
    assert bench.algorithmic_bytes(1, 0, 0) == 8


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
