"""bench.py's command line on the CPU: more than one GPU is refused with a
message and exit status 2 before any GPU work; --help lists the legs."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_help_lists_the_legs():
    r = _run("--help")
    assert r.returncode == 0
    for leg in ("c2_pt", "c4_reference_compat", "c5_batched", "rmse", "f_rows"):
        assert leg in r.stdout
