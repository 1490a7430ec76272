"""Launch-shape autotuner for the batched form.

Coordinate descent over the engine's tuning knobs -- workgroup size, regroup
interval (steps per launch), queue depth (new primaries admitted per launch),
jobs in flight, tail-compaction threshold, secondary handoff, the lanes a
wave gathers for findEndOfStep, the between-launch queue sort and the
admission order -- timing a caller-supplied workload on a fresh context for
each candidate (best of `repeats` runs).  Sweeps repeat from the current best
until one changes nothing, so that a knob whose best value depends on
another's is revisited after that one moved; a candidate replaces the best
only when it is faster by more than `min_gain` (the run-to-run noise).  The
winner can be stored as JSON and passed to `abi.Engine(**tuning)`.

Usage: python -m cudamonsel_amd.autotune batch [out.json]
       (the 512-job batch workload of serving.batch_workload)
       python -m cudamonsel_amd.autotune linescan [out.json]
       (line-scan jobs of 128 pixels x 2000 primaries across a PMMA line on
       Si at 0.5 keV, submitted back to back with 4 in flight)
"""
import json
import os
import time

from .abi import Engine

SPACE = {
    "block_size": [256, 128],
    "steps_per_launch": [0, 512, 128],
    # 0 is the engine's default (262,144); 1 << 30 admits all the queue holds
    "primaries_per_launch": [0, 1 << 20, 1 << 17, 1 << 30],
    "max_jobs_in_flight": [0, 256, 64],
    "tail_requeue": [0, 16, 48, -1],
    "spawn_handoff": [0, -1],
    "geometry_batch": [0, 4, 8],
    "sort_queue": [0, -1],
    "admission_order": [0, 1],
}


def _time(workload, tuning, device, repeats=1):
    with Engine(device, **tuning) as eng:
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            workload(eng)
            best = min(best, time.perf_counter() - t0)
        return best


def autotune(workload, space=None, device=0, log=None, repeats=2, min_gain=0.01, max_sweeps=4):
    """workload(engine) runs the jobs to be tuned for.  Returns (best tuning
    dict, [(tuning, seconds), ...]) from coordinate descent that starts at
    the defaults (the first value of every knob) and sweeps all knobs until a
    sweep changes nothing (at most max_sweeps)."""
    space = space or SPACE
    best = {k: v[0] for k, v in space.items()}
    trials = []
    seen = {}

    def timed(cand):
        key = tuple(sorted(cand.items()))
        if key not in seen:
            seen[key] = _time(workload, cand, device, repeats)
            trials.append((dict(cand), seen[key]))
        return seen[key]

    best_t = timed(best)
    if log:
        log(f"autotune start {best} {best_t:.3f}s")
    for sweep in range(max_sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best)
                cand[knob] = v
                t = timed(cand)
                if log:
                    log(f"autotune sweep {sweep} {knob}={v}: {t:.3f}s (best {best_t:.3f}s)")
                if t < best_t * (1.0 - min_gain):
                    best, best_t = cand, t
                    changed = True
        if not changed:
            break
    return best, trials


def save(path, tuning, trials):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"tuning": tuning, "trials": [{"tuning": t, "seconds": s} for t, s in trials]}, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)["tuning"]


# the line scan starts from its measured tuning; the regroup interval and
# the tail threshold get finer steps (their best values moved with the cost
# of the between-launch queue sort)
LINESCAN_SPACE = {
    "primaries_per_launch": [262144, 131072, 524288, 1 << 30],
    "steps_per_launch": [0, 512, 256, 128, 64],
    "tail_requeue": [48, 32, 64, 16, -1],
    "geometry_batch": [8, 4, 12, 16],
    "sort_queue": [0, -1],
    "spawn_handoff": [0, -1],
    "admission_order": [0, 1],
}


def linescan_workload(jobs=6, inflight=4, per_pixel=2000):
    """workload(engine) for the line scan: the scene is uploaded once per
    engine, then `jobs` scan jobs run with up to `inflight` on the device."""
    import numpy as np
    from . import scenes
    sc = scenes.line_on_substrate(h_nm=50.0, w_nm=20.0, theta_deg=3.0, r_nm=5.0, length_nm=1000.0, line="PMMA",
                                  substrate="Si", beam_eV=500.0, sigma_nm=0.5, compat=True)
    centers = scenes.scan_centers(np.arange(-64.0, 64.0, 1.0), [0.0], sc.beam_z)
    cs = sc.monte.compileScene()

    def workload(eng):
        sid = getattr(eng, "_linescan_scene", None)
        if sid is None:
            sid = eng._linescan_scene = eng.upload(cs)
            eng.result(eng.submit(sid, centers, per_pixel, 500.0, 0.5e-9, 999))  # warm-up
        ids = []
        for k in range(jobs):
            ids.append(eng.submit(sid, centers, per_pixel, 500.0, 0.5e-9, 1000 + k))
            if len(ids) >= inflight:
                eng.result(ids.pop(0))
        while ids:
            eng.result(ids.pop(0))
    return workload


def _linescan_main(out):
    import sys
    log = lambda m: print(m, file=sys.stderr, flush=True)
    best, trials = autotune(linescan_workload(), space=LINESCAN_SPACE, log=log)
    save(out, best, trials)
    print(json.dumps(best))


def _batch_main(out):
    import sys
    from . import serving
    jobs = serving.batch_workload(512, seed=4)

    def workload(eng):
        serving.run_batch(eng, jobs)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    best, trials = autotune(workload, log=log)
    save(out, {k: v for k, v in best.items()}, trials)
    print(json.dumps(best))


if __name__ == "__main__":
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    if len(sys.argv) > 1 and sys.argv[1] == "batch":
        _batch_main(sys.argv[2] if len(sys.argv) > 2 else os.path.join(here, "data", "tuning_batch.json"))
    elif len(sys.argv) > 1 and sys.argv[1] == "linescan":
        _linescan_main(sys.argv[2] if len(sys.argv) > 2 else os.path.join(here, "data", "tuning_linescan.json"))
    else:
        print(__doc__, file=sys.stderr)
        sys.exit(2)
