"""bench.py's command line on the CPU: one GPU only (a run asked for more
exits with status 2 and says why), and every mode is listed."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_modes_listed():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--projection", "--ad", "--mgrids", "--config", "--serve-only", "--kernel-only"):
        assert opt in r.stdout
