"""bench.py's command line without a GPU: the help lists every workload and
the observables switch, and a request for more than one GPU ends with a
message and exit status 2 before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_help_lists_workloads_and_observables():
    out = _run("--help")
    assert out.returncode == 0
    for word in ("config2", "config3", "config4", "--observables", "--converge", "--saturate"):
        assert word in out.stdout


def test_more_than_one_gpu_is_refused():
    out = _run("--gpus", "2")
    assert out.returncode == 2
    assert "one GPU" in out.stderr
