"""bench.py's command line without a GPU: more than one GPU is refused with a message
(every box here has one MI355X), and the options parse."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_options_parse():
    sys.path.insert(0, REPO)
    import bench
    old = sys.argv
    try:
        sys.argv = ["bench.py", "--steps", "3", "--warmup", "1", "--no-cpu", "--no-bodies", "--no-variants"]
        a = bench.parse()
    finally:
        sys.argv = old
    assert (a.gpus, a.steps, a.warmup, a.no_cpu, a.no_bodies, a.no_variants) == (1, 3, 1, True, True, True)
