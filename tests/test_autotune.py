"""The autotuner's search (CPU, with a synthetic cost in place of timed GPU
runs): coordinate descent revisits a knob whose best value depends on
another knob that moved later in the sweep, ignores changes within the
noise margin, and times every candidate once."""
import pytest


def test_coordinate_descent_revisits_interacting_knobs(monkeypatch):
    from cudamonsel_amd import autotune
    space = {"a": [0, 1, 2], "b": [0, 1]}

    # a's best value depends on b: with b = 0 it is a = 1, with b = 1 it is a = 2
    def cost(t):
        a, b = t["a"], t["b"]
        table = {(0, 0): 10.0, (1, 0): 9.0, (2, 0): 9.5, (0, 1): 9.8, (1, 1): 8.5, (2, 1): 7.0}
        return table[(a, b)]
    calls = []

    def fake_time(workload, tuning, device, repeats=1):
        calls.append(dict(tuning))
        return cost(tuning)
    monkeypatch.setattr(autotune, "_time", fake_time)
    best, trials = autotune.autotune(None, space=space)
    assert best == {"a": 2, "b": 1}
    assert len(calls) == len({tuple(sorted(c.items())) for c in calls})   # each candidate once
    # a single sweep would have stopped at a = 1, b = 1 (8.5)
    assert min(t for _, t in trials) == 7.0


def test_noise_margin(monkeypatch):
    from cudamonsel_amd import autotune
    space = {"a": [0, 1]}
    monkeypatch.setattr(autotune, "_time", lambda w, t, d, repeats=1: 1.0 if t["a"] == 0 else 0.995)
    best, _ = autotune.autotune(None, space=space, min_gain=0.01)
    assert best == {"a": 0}
