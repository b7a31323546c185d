"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message, and the help text names the options."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="4", RANK="0", LOCAL_RANK="0")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120, env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_the_workload_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--mesh", "--sgs", "--precision", "--mach", "--no-compare"):
        assert opt in r.stdout


def test_cpu_baselines_on_a_small_sample():
    """Both CPU legs of the bench run on the host (no GPU): the compiled
    reference (where oracle/_ref was built) and the oracle."""
    import pytest
    sys.path.insert(0, ROOT)
    import bench
    from oracle_lib import ref_available
    port = bench.cpu_baseline(steps=1, mesh=(16, 12, 8))
    assert port["kind"] == "port" and port["cores"] == 1 and port["value"] > 0
    if not ref_available("dyn"):
        pytest.skip("compiled reference not built here")
    ref = bench.cpu_reference(steps=1, mesh=(16, 12, 8))
    assert ref["kind"] == "reference" and ref["cores"] == 1 and ref["value"] > 0
