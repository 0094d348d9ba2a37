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


def test_roofline_inputs_are_the_documented_byte_counts():
    """The neuron kernel's algorithmic bytes per neuron (78 with parameter sets,
    133 with per-neuron constants; pair weights and direct delivery add their
    arrivals), and the delivery traffic model read from the committed PMC
    summary: its fixed accumulator write plus bytes per event."""
    sys.path.insert(0, ROOT)
    import bench

    assert bench.neuron_bytes({"neuron_psets": 4, "deliver_mode": 1}) == 78
    assert bench.neuron_bytes({"neuron_psets": 0, "deliver_mode": 1}) == 133
    assert bench.neuron_bytes({"neuron_psets": 4, "deliver_mode": 1}, per_neuron_noise=False) == 74
    assert bench.neuron_bytes({"neuron_psets": 4, "deliver_mode": 1}, pair=True) == 86
    assert bench.neuron_bytes({"neuron_psets": 4, "deliver_mode": 0}) == 86  # + the arrival clear
    txt = open(os.path.join(ROOT, "profiles", "delivery_large_pmc.md")).read()
    assert "traffic_fixed_bytes_per_launch: 4e+07" in txt and "traffic_bytes_per_event: 32.82" in txt
    assert abs(bench.pmc_traffic("delivery", "large", 1e7) - (4e7 + 32.82e7)) < 1.0
    assert bench.pmc_traffic("delivery", "no_such_workload", 1e7) is None
