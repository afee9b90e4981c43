"""The autotuner's coordinate search on the host: with the library calls replaced by a
timing model, every value of every axis is timed once, each axis keeps its best value so
far, and the fastest setting seen is the one applied last."""
import itertools

import pytest
import torch

from gtls_amd import tune

AXES = (tune.TOEP_Q, tune.CHUNK_MB, tune.WG_PER_CU, tune.QUAD_Q)


class FakeBatch:
    """Stands in for a DeviceBatch and its sample batch: counts search calls."""

    class _Dev:
        index = 0

    def __init__(self):
        self.dev = self._Dev()
        self.calls = 0

    def search(self, *args, **kwargs):
        self.calls += 1


def run_autotune(monkeypatch, cost):
    """autotune() with set_tuning / kernel_ms / the sample batch replaced: a search call
    under a setting takes cost(setting) milliseconds. Returns (choice, times, applied)."""
    applied, state = [], {}

    def set_tuning(dev, *setting):
        applied.append(tuple(setting))
        state["cur"] = tuple(setting)

    def kernel_ms(dev):
        return (cost(state["cur"]), 0.0, 0.0), None

    fb = FakeBatch()
    monkeypatch.setattr(tune, "set_tuning", set_tuning)
    monkeypatch.setattr(tune, "kernel_ms", kernel_ms)
    monkeypatch.setattr(tune, "set_kernel_timing", lambda dev, on: None)
    monkeypatch.setattr(tune, "sample_batch", lambda db, n: fb)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    choice, times = tune.autotune(fb)
    return choice, times, applied, fb


def separable(best):
    """A cost with its unique minimum at `best`: the sum over the axes of the value's
    distance (in list positions) from the best one."""
    def cost(s):
        return 100.0 + sum(abs(AXES[a].index(s[a]) - AXES[a].index(best[a]))
                           for a in range(4))
    return cost


@pytest.mark.parametrize("best", [
    (12, 64 << 10, 2, 96), (8, 8 << 10, 1, 1 << 20), (24, 64 << 10, 1, 64),
    (16, 8 << 10, 2, 128)])
def test_separable_cost_reaches_its_minimum(monkeypatch, best):
    choice, times, applied, fb = run_autotune(monkeypatch, separable(best))
    assert choice == best
    assert applied[-1] == best
    # every value of every axis timed, each setting once
    for a, values in enumerate(AXES):
        assert {s[a] for s in times} == set(values)
    assert len(times) <= sum(len(v) for v in AXES)
    # one untimed call, then one call per timed setting (repeats = 1)
    assert fb.calls == 1 + len(times)


def test_the_fastest_setting_seen_is_applied(monkeypatch):
    """A cost with interactions: whatever the path, the applied setting is the fastest of
    those timed."""
    table = {s: float(hash(s) % 1000) + 1.0 for s in itertools.product(*AXES)}
    choice, times, applied, _ = run_autotune(monkeypatch, lambda s: table[s])
    assert times[choice] == min(times.values())
    assert applied[-1] == choice
    assert all(times[s] == table[s] for s in times)


def test_repeats_take_the_best_of_each_setting(monkeypatch):
    """With repeats, each setting's figure is its fastest call."""
    calls = {}

    def cost(s):
        calls[s] = calls.get(s, 0) + 1
        return 50.0 if calls[s] == 2 else 80.0   # the second call of a setting is fastest

    applied, state = [], {}
    monkeypatch.setattr(tune, "set_tuning",
                        lambda dev, *s: (applied.append(tuple(s)), state.update(cur=tuple(s))))
    monkeypatch.setattr(tune, "kernel_ms", lambda dev: ((cost(state["cur"]), 0.0, 0.0), None))
    monkeypatch.setattr(tune, "set_kernel_timing", lambda dev, on: None)
    fb = FakeBatch()
    monkeypatch.setattr(tune, "sample_batch", lambda db, n: fb)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    _, times = tune.autotune(fb, repeats=3)
    assert set(times.values()) == {50.0}
