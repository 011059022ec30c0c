"""bench.py's command line without a GPU: more than one GPU is refused
before anything touches a device, and the default step counts are the ones
the help text states."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_world_size_above_one_refused():
    env = dict(os.environ, WORLD_SIZE="4")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")], capture_output=True, text=True,
                       timeout=120, cwd=ROOT, env=env)
    assert r.returncode == 2
    assert "WORLD_SIZE=4" in r.stderr


def test_help_states_defaults():
    r = run_bench("--help")
    assert r.returncode == 0
    assert "50 at c1" in r.stdout and "10 at c1" in r.stdout
