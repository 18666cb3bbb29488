"""bench.py's command line without a GPU: it measures one GPU and says so
when asked for more (the path shards with no communication, so a figure
at several GPUs would be the one-GPU figure times N)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                            "--gpus", n], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode != 0
        assert "one GPU" in (r.stdout + r.stderr)
        assert not r.stdout.lstrip().startswith("{")  # no result line
