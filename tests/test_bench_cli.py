"""bench.py's command line on the CPU: more than one GPU is refused with a
message, before anything touches a GPU."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "one GPU" in r.stderr and r.stdout == ""
