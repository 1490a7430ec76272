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


def test_linescan_workload_uploads_once_per_engine():
    """The line-scan workload uploads its scene and runs its warm-up job once
    per engine, then submits its jobs with at most `inflight` unfinished."""
    from cudamonsel_amd import autotune

    class FakeEngine:
        def __init__(self):
            self.uploads, self.open, self.peak, self.submitted, self.next = 0, set(), 0, [], 1

        def upload(self, cs):
            self.uploads += 1
            return 7

        def submit(self, sid, centers, per, e, sigma, seed):
            assert sid == 7 and centers.shape == (128, 3) and per == 2000 and e == 500.0
            self.next += 1
            self.open.add(self.next)
            self.peak = max(self.peak, len(self.open))
            self.submitted.append(seed)
            return self.next

        def result(self, jid):
            self.open.remove(jid)
            return {}

    w = autotune.linescan_workload(jobs=6, inflight=4)
    a, b = FakeEngine(), FakeEngine()
    w(a)
    w(a)
    w(b)
    assert a.uploads == 1 and b.uploads == 1
    assert a.submitted == [999] + list(range(1000, 1006)) * 2 and a.peak == 4 and not a.open
