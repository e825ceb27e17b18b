"""The bench's command-line contract on a host without a GPU: more than one
GPU is refused with a message, and without a usable device the bench stops
with an error instead of measuring anything on the CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=300)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert not r.stdout.strip().startswith("{")


def test_no_device_stops_the_bench():
    import lsamd

    if lsamd.load().ls_amd_device_available() == 0:
        pytest.skip("a usable device is present: the bench would run")
    r = run_bench("--steps", "1", "--warmup", "0")
    assert r.returncode != 0
    assert "no usable gfx950 device" in r.stderr
    assert "{" not in r.stdout
