"""bench.py's argument handling, which runs before any device is touched."""
import os
import subprocess
import sys

from conftest import ROOT


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_reason():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_unknown_mode_and_bad_counts_are_refused():
    r = _bench("--modes", "fp64,fp16")
    assert r.returncode == 2 and "fp16" in r.stderr
    r = _bench("--steps", "0")
    assert r.returncode == 2 and "--steps" in r.stderr
    r = _bench("--workload", "c9")
    assert r.returncode == 2
