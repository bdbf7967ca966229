"""bench.py's command-line contract on the CPU: a run on more than one GPU is
refused with a message before anything touches the library or a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_a_distributed_launch_is_refused():
    r = _run(["--steps", "1", "--warmup", "0"], {"WORLD_SIZE": "4", "RANK": "0"})
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
