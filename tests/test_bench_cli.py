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


def test_traffic_profiles_parse_and_match_their_workload():
    """The committed --pmc summaries give the bench line's `traffic` only for the workload they were measured on."""
    sys.path.insert(0, ROOT)
    import bench

    tet = bench.traffic_from_profile()
    assert tet is not None and 1.0 < tet / 147468104 < 1.2
    for path in (bench.PROFILE_PMC_BATCH, bench.PROFILE_PMC_CONTACT):
        workload = next(line.split(":", 1)[1].strip() for line in open(path) if line.startswith("workload:"))
        traffic = bench.traffic_from_profile(path, workload)
        assert traffic is not None and traffic > 1e9
        assert bench.traffic_from_profile(path, workload + " (other)") is None
