"""bench.py's command-line contract that needs no GPU: one GPU only (a request for
more exits with status 2 and says why, before anything touches a device), and the
option set the driver and the profiles use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, env=env, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="2")
    r = _run("--gpus", "1", env=env)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_options():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-plain", "--workload",
                "--calibrate-copy"):
        assert opt in r.stdout, opt
    for w in ("glv10", "vdp_obs", "glv128_shared"):
        assert w in r.stdout, w
