"""The entry points the tooling calls, checked without a device: bench.py's
refusal of more than one GPU, its algorithmic byte counts, and the build /
smoke functions of __graft_entry__."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_more_than_one_gpu(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_bytes_per_update():
    sys.path.insert(0, ROOT)
    import bench
    Q = bench.Q
    assert bench.bytes_per_lup({"path": Q.PATH_SPLIT, "compressed": 1}) == 17.0
    assert bench.bytes_per_lup({"path": Q.PATH_SPLIT, "compressed": 0}) == 40.0
    assert bench.bytes_per_lup({"path": Q.PATH_FUSED, "compressed": 1}) == 13.0
    assert bench.bytes_per_lup({"path": Q.PATH_PLAIN, "compressed": 0}) == 64.0


def test_graft_entry_exposes_build_and_smoke():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as g
    assert callable(g.build) and callable(g.smoke)
