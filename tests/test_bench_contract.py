"""bench.py runs on one GPU only: any other GPU count exits with a message before touching a
device (checked on the CPU)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "--gpus 1" in r.stderr + r.stdout
    assert not r.stdout.strip().startswith("{")  # no benchmark line


def test_graph_chunk_divides_timed_steps():
    """The headline's graph chunk is the largest divisor of the timed steps within the limit,
    so every timed MCS replays a graph captured before the timed region."""
    sys.path.insert(0, ROOT)
    import bench
    assert bench.graph_chunk_for(20, 50) == 20
    assert bench.graph_chunk_for(500, 50) == 50
    assert bench.graph_chunk_for(120, 50) == 40
    assert bench.graph_chunk_for(53, 50) == 1
    assert bench.graph_chunk_for(20, 0) == 0
