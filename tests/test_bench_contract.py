"""bench.py's command-line contract that needs no GPU: it measures one GPU
only, so a request for several exits with status 2 and says why, before it
loads the library or touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2, (p.returncode, p.stderr)
    assert "one GPU only" in p.stderr
    assert p.stdout == ""
