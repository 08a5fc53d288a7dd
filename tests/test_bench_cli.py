"""bench.py contract on the CPU: more than one GPU is refused with a message."""
import os
import subprocess
import sys

from conftest import ROOT


def test_multi_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
