"""bench.py's command line without a GPU: a request for more than one GPU is
refused (exit status 2, a message on stderr, no JSON line on stdout) before
anything touches a device."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"),
                            "--gpus", str(n)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 2, r.stderr
        assert r.stdout == ""
        assert f"--gpus {n}" in r.stderr
