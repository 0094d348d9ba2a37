"""bench.py's command-line contract that needs no GPU: it measures one GPU
and refuses more, with a message, before touching the device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=e)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_a_multi_rank_launch_is_refused_too():
    r = _bench("--gpus", "1", env={"WORLD_SIZE": "8"})
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_steps_that_are_not_whole_bold_samples_are_refused():
    for steps in ("0", "5", "25"):
        r = _bench("--steps", steps)
        assert r.returncode == 2 and "multiple of 10" in r.stderr and r.stdout == "", steps
