"""bench.py's contract on the CPU: more than one GPU is refused with a message and a
non-zero exit (this build measures one GPU), before any device work; the multi-GPU
design is documented, not run."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
    assert not r.stdout.strip().startswith("{")  # no result line
