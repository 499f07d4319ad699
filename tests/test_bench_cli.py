"""bench.py's command line parses for every workload, and more than one GPU is refused with a message."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORKLOADS = ["two_body", "precision", "batch", "batch_precision", "contact", "sdf_build", "ray_cast"]


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("workload", WORKLOADS)
def test_bench_arguments_parse(workload):
    r = run("--workload", workload, "--help")
    assert r.returncode == 0, r.stderr


def test_bench_refuses_several_gpus():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
