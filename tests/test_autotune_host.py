"""CPU checks of the delivery autotuner's host logic (digital_twin_brain_amd/autotune.py)
on a stand-in block whose costs are a fixed function of its launch shape: the
tuner keeps the cheapest shape it measured, decides segment lists and the
persistent launch by whole-step time, records the per-kernel shape's mode
beside a persistent choice, and its cache restores the same launch."""
import numpy as np
import pytest

from digital_twin_brain_amd import autotune
from digital_twin_brain_amd._native import DTBError

EVENTS, SPIKES = 100_000, 100
KEYS = ("deliver_mode", "deliver_lanes", "deliver_grid", "stage_segs", "stage_ept", "stage_variant", "accum_occ")


class StandInBlock:
    """n, nnz, max_delay and the launch interface of NativeBlock; the delivery
    phase of a step costs cost(shape) us per 1e6 events."""

    def __init__(self, persist_us=None, seg_on_us=50.0, seg_off_us=60.0, n=200_000, D=10):
        self.n, self.nnz, self.max_delay = n, n * 100, D
        self.launch = dict(deliver_mode=1, deliver_lanes=64, deliver_grid=1024, stage_segs=128, stage_ept=16,
                           stage_variant=0, accum_occ=1, segment_lists=0)
        self.persist_us, self.seg_on_us, self.seg_off_us = persist_us, seg_on_us, seg_off_us
        self.measured = []

    def cost(self, shape):
        mode, lanes, grid, sg, ept, var, occ = shape
        if mode == 0:
            return 300.0 + lanes + abs(np.log2(grid)) * 3.0
        return (100.0 + 4.0 * abs(np.log2(grid / 1024.0)) + {64: 6.0, 128: 0.0, 256: 3.0}[sg] +
                (0.0 if ept == 16 else 2.0) + {0: 2.5, 1: 1.5, 2: 0.0}[var] + (0.5 if occ == 2 else 0.0))

    def shape(self):
        return tuple(self.launch[k] for k in KEYS)

    def set_launch(self, **kw):
        if kw.get("deliver_mode") == 2 and self.persist_us is None:
            raise DTBError(1, "persistent delivery does not fit this block")
        self.launch.update(kw)

    def get_launch(self):
        return dict(self.launch)

    def run(self, steps, graph=True, kernel_ms=None, spikes=None, events=None, **kw):
        if spikes is not None:
            spikes[:] = SPIKES
        if events is not None:
            events[:] = EVENTS
        if kernel_ms is not None:
            self.measured.append(self.shape())
            kernel_ms[:, 0] = 0.05
            kernel_ms[:, 1] = self.cost(self.shape()) * (EVENTS / 1e6) / 1e3

    def step_us(self):
        if self.launch["deliver_mode"] == 2:
            return self.persist_us
        return self.seg_on_us if self.launch["segment_lists"] else self.seg_off_us


@pytest.fixture
def tuner(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_step_us", lambda nb, kw, steps: nb.step_us())
    return autotune.tune_delivery


def _expected_best(nb):
    shapes = [s for s in set(nb.measured) if s[0] in (0, 1)]
    return min(shapes, key=nb.cost)


def test_keeps_the_cheapest_measured_shape_and_segment_lists_by_step_time(tuner):
    nb = StandInBlock(persist_us=None)
    best, table = tuner(nb, {}, steps=20)
    want = _expected_best(nb)
    assert tuple(best[k] for k in KEYS) == want
    assert best["us_per_1e6_events"] == pytest.approx(nb.cost(want), rel=1e-5)
    assert best["segment_lists"] == 1 and nb.launch["segment_lists"] == 1
    assert nb.shape() == want
    assert "persistent step us" not in table  # the block does not fit the persistent launch
    # the six finalists were measured a second time, over five times the steps
    assert sum(1 for k in table if k.endswith("*")) == 6


def test_persistent_launch_when_faster_keeps_the_shape_mode(tuner):
    nb = StandInBlock(persist_us=20.0, seg_on_us=70.0, seg_off_us=40.0)
    best, table = tuner(nb, {}, steps=20)
    want = _expected_best(nb)
    assert best["deliver_mode"] == 2 and best["shape_mode"] == want[0] == 1
    assert best["segment_lists"] == 0 and best["us_per_step"] == 20.0
    assert nb.launch["deliver_mode"] == 2 and table["persistent step us"] < table["best shape step us"]


def test_persistent_launch_when_slower_is_not_kept(tuner):
    nb = StandInBlock(persist_us=90.0)
    best, table = tuner(nb, {}, steps=20)
    assert best["deliver_mode"] == 1 and "shape_mode" not in best
    assert nb.launch["deliver_mode"] == 1 and nb.shape() == _expected_best(nb)


@pytest.mark.parametrize("persist_us", [None, 20.0])
def test_cache_restores_the_same_launch(tuner, persist_us):
    first = StandInBlock(persist_us=persist_us)
    best, _ = tuner(first, {}, steps=20)
    again = StandInBlock(persist_us=persist_us)
    best2, table2 = tuner(again, {}, steps=20)
    assert "cached" in table2 and best2 == best
    assert again.get_launch() == first.get_launch()
    assert len(again.measured) == 1  # only the probe run that sizes the cache key
