"""Launch-shape autotuner for the batched form.

Coordinate search over the engine's tuning knobs -- workgroup size, regroup
interval (steps per launch), queue depth (new primaries admitted per launch),
jobs in flight, tail-compaction threshold, secondary handoff and the lanes a
wave gathers for findEndOfStep -- timing a
caller-supplied workload on a fresh context for each candidate.  The winner can be stored as JSON and passed to
`abi.Engine(**tuning)`.
"""
import json
import os
import time

from .abi import Engine

SPACE = {
    "block_size": [256, 128],
    "steps_per_launch": [0, 512, 128],
    "primaries_per_launch": [0, 1 << 20, 1 << 17],
    "max_jobs_in_flight": [0, 256, 64],
    "tail_requeue": [0, 16, 48, -1],
    "spawn_handoff": [0, -1],
    "geometry_batch": [0, 4, 8],
}


def _time(workload, tuning, device):
    with Engine(device, **tuning) as eng:
        t0 = time.perf_counter()
        workload(eng)
        return time.perf_counter() - t0


def autotune(workload, space=None, device=0, log=None):
    """workload(engine) runs the jobs to be tuned for.  Returns (best tuning
    dict, [(tuning, seconds), ...]) from a one-knob-at-a-time search that
    starts at the defaults (all zeros)."""
    space = space or SPACE
    best = {k: v[0] for k, v in space.items()}
    trials = []
    best_t = _time(workload, best, device)
    trials.append((dict(best), best_t))
    if log:
        log(f"autotune start {best} {best_t:.3f}s")
    for knob, values in space.items():
        for v in values[1:]:
            cand = dict(best)
            cand[knob] = v
            t = _time(workload, cand, device)
            trials.append((cand, t))
            if log:
                log(f"autotune {knob}={v}: {t:.3f}s")
            if t < best_t:
                best, best_t = cand, t
    return best, trials


def save(path, tuning, trials):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"tuning": tuning, "trials": [{"tuning": t, "seconds": s} for t, s in trials]}, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)["tuning"]
