"""Batched and served forms of the transport engine.

- `batch_workload` / `run_batch`: many single-pixel yield jobs of uneven size
  and energy (flat or step-edge Si, PMMA, Cu) submitted together, so that
  they share transport launches, against the same jobs run one at a time.
- `request_stream` / `serve`: a Poisson stream of SEM-image requests (random
  line arrays or contact holes, PMMA on Si) served with continuous batching:
  requests are admitted while earlier ones are on the GPU as long as the
  device-memory footprint of everything admitted fits a budget; each
  request's latency runs from its arrival to its result.  `serve(...,
  plain=True)` runs the same stream one request at a time;
  `per_pixel_plain` estimates it run as the reference does, one job per
  pixel.

Scenes are built with the reference's classes (`scenes.py`); jobs go through
the C-ABI (`abi.Engine`).  Results of a job are the same alone and in any mix
(every electron's random stream depends only on its job's seed, primary index
and ancestry; tallies are integers).
"""
import heapq
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import scenes

MATERIALS = ("Si", "PMMA", "Cu")


@dataclass
class YieldJob:
    material: str
    step_edge: bool
    beam_eV: float
    primaries: int
    seed: int


def batch_workload(n_jobs=512, seed=4, e_range=(200.0, 5000.0), n_range=(100, 100000)):
    """Jobs with E0 log-uniform, material uniform over Si/PMMA/Cu, flat or
    with a 20 nm step edge, primaries log-uniform."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_jobs):
        jobs.append(YieldJob(material=MATERIALS[rng.integers(3)], step_edge=bool(rng.integers(2)),
                             beam_eV=float(np.exp(rng.uniform(*np.log(e_range)))),
                             primaries=int(np.exp(rng.uniform(*np.log(n_range)))), seed=1000 + k))
    return jobs


class SceneCache:
    """Reference-style scenes (kept alive) and their uploaded ids."""

    def __init__(self, engine):
        self.engine = engine
        self.held = {}

    def get(self, key, build):
        if key not in self.held:
            sc = build()
            cs = sc.monte.compileScene()
            self.held[key] = (sc, cs, self.engine.upload(cs))
        return self.held[key]


def _yield_scene(cache, job):
    key = ("yield", job.material, job.step_edge)
    if job.step_edge:
        return cache.get(key, lambda: scenes.step_edge(job.material))
    return cache.get(key, lambda: scenes.flat(job.material))


def run_batch(engine, jobs, plain=False):
    """Runs the jobs together (or one at a time when plain) and returns
    (results, wall seconds)."""
    cache = SceneCache(engine)
    for j in jobs:
        _yield_scene(cache, j)
    t0 = time.perf_counter()
    results = []
    if plain:
        for j in jobs:
            sc, cs, sid = _yield_scene(cache, j)
            results.append(engine.result(engine.submit(sid, [[0.0, 0.0, sc.beam_z]], j.primaries, j.beam_eV,
                                                       0.5e-9, j.seed)))
    else:
        ids = []
        for j in jobs:
            sc, cs, sid = _yield_scene(cache, j)
            ids.append(engine.submit(sid, [[0.0, 0.0, sc.beam_z]], j.primaries, j.beam_eV, 0.5e-9, j.seed))
        results = [engine.result(i) for i in ids]
    return results, time.perf_counter() - t0


@dataclass(order=True)
class Request:
    arrival: float
    rid: int = field(compare=False)
    side: int = field(compare=False)
    per_pixel: int = field(compare=False)
    beam_eV: float = field(compare=False)
    geometry: str = field(compare=False)
    geo_seed: int = field(compare=False)

    @property
    def primaries(self):
        return self.side * self.side * self.per_pixel


def request_stream(rate_per_s=2.0, duration_s=600.0, seed=5, sides=(32, 64, 128, 256), per_pixel=(50, 500),
                   e_range=(300.0, 2000.0), n_geometries=8):
    """Poisson arrivals; side uniform over `sides`, primaries/pixel and beam
    energy uniform; geometry = one of n_geometries seeded random-line arrays
    or contact-hole layers."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t >= duration_s:
            break
        geo = "lines" if rng.integers(2) == 0 else "holes"
        out.append(Request(arrival=t, rid=len(out), side=int(sides[rng.integers(len(sides))]),
                           per_pixel=int(rng.integers(per_pixel[0], per_pixel[1] + 1)),
                           beam_eV=float(rng.uniform(*e_range)), geometry=geo,
                           geo_seed=int(rng.integers(n_geometries))))
    return out


def _request_scene(cache, req):
    key = (req.geometry, req.geo_seed)
    if req.geometry == "lines":
        return cache.get(key, lambda: scenes.random_lines(seed=100 + req.geo_seed))
    rng = np.random.default_rng(200 + req.geo_seed)
    radius = float(rng.uniform(8.0, 25.0))
    depth = float(rng.uniform(30.0, 100.0))
    return cache.get(key, lambda: scenes.contact_holes(radius_nm=radius, depth_nm=depth))


def raster_centers(side, z):
    """side x side pixel centres at 1 nm pitch around the origin (m)."""
    xs = (np.arange(side) - (side - 1) / 2.0) * 1e-9
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    return np.stack([X.ravel(), Y.ravel(), np.full(side * side, z)], 1)


def percentile(values, q):
    """Nearest-rank percentile (q in [0, 100])."""
    v = sorted(values)
    if not v:
        return float("nan")
    k = max(0, min(len(v) - 1, int(math.ceil(q / 100.0 * len(v))) - 1))
    return float(v[k])


class Admission:
    """Admits requests while the summed device footprint of the admitted
    and unfinished ones stays within the budget (always at least one)."""

    def __init__(self, budget_bytes):
        self.budget = budget_bytes
        self.used = 0
        self.inflight = 0

    def fits(self, nbytes):
        return self.inflight == 0 or self.used + nbytes <= self.budget

    def admit(self, nbytes):
        self.used += nbytes
        self.inflight += 1

    def release(self, nbytes):
        self.used -= nbytes
        self.inflight -= 1


def per_pixel_plain(engine, requests, sample=16, log=None):
    """The reference's call pattern on the GPU: each request as one job per
    pixel (GaussianBeam::setCenter + runMultipleTrajectories per pixel,
    LinesOnLayers.cu:350-380), one job at a time.  Timed on `sample` pixels of
    every request (spread over the image); a request's service time is the
    mean pixel time x its pixels, and the stream is replayed on those service
    times in arrival order (one request at a time) for latencies."""
    cache = SceneCache(engine)
    for r in requests:
        _request_scene(cache, r)
    service = []
    for r in sorted(requests, key=lambda q: q.arrival):
        sc, cs, sid = _request_scene(cache, r)
        cen = raster_centers(r.side, sc.beam_z)
        idx = np.linspace(0, cen.shape[0] - 1, min(sample, cen.shape[0])).astype(int)
        t0 = time.perf_counter()
        for k in idx:
            engine.result(engine.submit(sid, cen[k:k + 1], r.per_pixel, r.beam_eV, 0.5e-9, 5000 + r.rid))
        dt = (time.perf_counter() - t0) / len(idx)
        service.append((r, dt * cen.shape[0]))
        if log:
            log(f"per-pixel plain: request {r.rid} ({r.side}x{r.side}), {dt * 1e3:.3f} ms per pixel")
    t = 0.0
    lat = []
    for r, sv in service:
        t = max(t, r.arrival) + sv
        lat.append(t - r.arrival)
    wall = t - min(r.arrival for r in requests)
    return dict(requests=len(requests), wall_s=wall, p50_s=percentile(lat, 50), p99_s=percentile(lat, 99),
                images_per_hour=3600.0 * len(requests) / wall, pixels_sampled_per_request=sample)


def per_pixel_check(engine, requests, sample=16):
    """Checks the per-pixel estimate on the smallest request of the stream:
    its estimate (sample pixels timed, mean x pixels) against all its pixels
    run one job each."""
    cache = SceneCache(engine)
    r = min(requests, key=lambda q: (q.side * q.side * q.per_pixel, q.rid))
    sc, cs, sid = _request_scene(cache, r)
    cen = raster_centers(r.side, sc.beam_z)
    engine.result(engine.submit(sid, cen[:1], r.per_pixel, r.beam_eV, 0.5e-9, 5000 + r.rid))   # warm-up
    idx = np.linspace(0, cen.shape[0] - 1, min(sample, cen.shape[0])).astype(int)
    t0 = time.perf_counter()
    for k in idx:
        engine.result(engine.submit(sid, cen[k:k + 1], r.per_pixel, r.beam_eV, 0.5e-9, 5000 + r.rid))
    est = (time.perf_counter() - t0) / len(idx) * cen.shape[0]
    t0 = time.perf_counter()
    for k in range(cen.shape[0]):
        engine.result(engine.submit(sid, cen[k:k + 1], r.per_pixel, r.beam_eV, 0.5e-9, 5000 + r.rid))
    run = time.perf_counter() - t0
    return dict(request=r.rid, side=r.side, per_pixel=r.per_pixel, estimated_s=est, measured_s=run,
                ratio=est / run)


def serve(engine, requests, plain=False, hbm_budget=64 << 30, time_scale=1.0, poll_s=0.0005, log=None,
          keep_results=False):
    """Replays the request stream in real time (arrival times divided by
    time_scale) and returns per-request latencies and totals (requests,
    primaries and electron steps per second of wall time); log(msg), if
    given, hears of every completed request.  With keep_results the
    returned dict also maps every request id to its tallies ("results")."""
    cache = SceneCache(engine)
    for r in requests:
        _request_scene(cache, r)
    pending = list(requests)
    heapq.heapify(pending)
    waiting = []          # arrived, not admitted (FIFO)
    running = {}          # job id -> (request, footprint)
    adm = Admission(hbm_budget)
    lat = {}
    kept = {}
    peak = 0
    steps = 0
    t0 = time.perf_counter()
    done = 0
    while done < len(requests):
        now = (time.perf_counter() - t0) * time_scale
        while pending and pending[0].arrival <= now:
            waiting.append(heapq.heappop(pending))
        # admission: FIFO, by footprint; plain = one request at a time
        while waiting and (not plain or not running):
            r = waiting[0]
            sc, cs, sid = _request_scene(cache, r)
            cen = raster_centers(r.side, sc.beam_z)
            need = engine.footprint(cen.shape[0], r.per_pixel, r.beam_eV, 50)
            if not adm.fits(need):
                break
            waiting.pop(0)
            jid = engine.submit(sid, cen, r.per_pixel, r.beam_eV, 0.5e-9, 5000 + r.rid)
            running[jid] = (r, need)
            adm.admit(need)
            peak = max(peak, len(running))
        finished = [j for j in running if engine.poll(j)]
        for j in finished:
            r, need = running.pop(j)
            res = engine.result(j)
            steps += int(res.get("stats", {}).get("steps", 0))
            if keep_results:
                kept[r.rid] = res
            adm.release(need)
            lat[r.rid] = (time.perf_counter() - t0) * time_scale - r.arrival
            done += 1
            if log:
                log(f"request {r.rid} ({r.side}x{r.side}, {r.per_pixel}/px, {r.beam_eV:.0f} eV, {r.geometry}) "
                    f"done, latency {lat[r.rid]:.2f} s ({done}/{len(requests)})")
        if not finished:
            if not running and not waiting and pending:
                time.sleep(max(0.0, (pending[0].arrival - now) / time_scale))
            else:
                time.sleep(poll_s)
    wall = time.perf_counter() - t0
    vals = [lat[r.rid] for r in requests]
    prim = sum(r.primaries for r in requests)
    out = dict(requests=len(requests), wall_s=wall, p50_s=percentile(vals, 50), p99_s=percentile(vals, 99),
               images_per_hour=3600.0 * len(requests) / wall, primaries=prim, primaries_per_s=prim / wall,
               electron_steps=steps, electron_steps_per_s=steps / wall, latencies=vals, max_in_flight=peak)
    if keep_results:
        out["results"] = kept
    return out


def batch_timing(engine, jobs):
    """run_batch's batched form with the host side of its wall time split
    out: the submit loop, the last job's completion on the engine (its
    submit time plus its submit-to-done time), and the reading of the
    results after it."""
    cache = SceneCache(engine)
    for j in jobs:
        _yield_scene(cache, j)
    t0 = time.perf_counter()
    ids, t_sub = [], []
    for j in jobs:
        sc, cs, sid = _yield_scene(cache, j)
        t_sub.append(time.perf_counter())
        ids.append(engine.submit(sid, [[0.0, 0.0, sc.beam_z]], j.primaries, j.beam_eV, 0.5e-9, j.seed))
    t1 = time.perf_counter()
    results = [engine.result(i) for i in ids]
    wall = time.perf_counter() - t0
    last = max(ts + r["stats"]["total_ms"] * 1e-3 for ts, r in zip(t_sub, results)) - t0
    prim = sum(j.primaries for j in jobs)
    return results, dict(jobs=len(jobs), primaries=prim, wall_s=wall, submit_s=t1 - t0, last_done_s=last,
                         read_after_last_s=wall - last, primaries_per_s=prim / wall)


def image_timing(engine, side=256, per_pixel=300, beam_eV=500.0, geo_seed=3, seed=3):
    """One side x side image of random_lines(geo_seed) at 1 nm pitch as one
    job, after a warm-up job of 256 pixels x 50 primaries; returns its
    tallies and the timing."""
    sc = scenes.random_lines(seed=geo_seed, beam_eV=beam_eV)
    sid = engine.upload(sc.monte.compileScene())
    cen = raster_centers(side, sc.beam_z)
    engine.result(engine.submit(sid, cen[:256], 50, beam_eV, 0.5e-9, seed + 1))
    t0 = time.perf_counter()
    r = engine.result(engine.submit(sid, cen, per_pixel, beam_eV, 0.5e-9, seed), timeout_s=600)
    wall = time.perf_counter() - t0
    prim = cen.shape[0] * per_pixel
    return r, dict(pixels=int(cen.shape[0]), primaries=prim, wall_s=wall, primaries_per_s=prim / wall,
                   electron_steps=int(r["stats"]["steps"]), launches=int(r["stats"]["transport_launches"]))


def main(argv=None):
    """python -m cudamonsel_amd.serving batch [--jobs N] [--tuning FILE]:
    the batch workload (N single-pixel yield jobs, seed 4) on a tuned engine,
    after a warm-up of its first 8 jobs; prints one JSON line.
    python -m cudamonsel_amd.serving image [--side N]: the 256 x 256 x 300
    image of three seeded lines at 0.5 keV as one job on the engine's
    default settings; prints one JSON line."""
    import argparse
    import json
    import os
    from .abi import Engine
    ap = argparse.ArgumentParser(prog="python -m cudamonsel_amd.serving")
    ap.add_argument("form", choices=["batch", "image"])
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--tuning", default=os.path.join(os.path.dirname(__file__), "data", "tuning_batch.json"))
    a = ap.parse_args(argv)
    if a.form == "image":
        with Engine() as eng:
            _, t = image_timing(eng, side=a.side)
        print(json.dumps(t), flush=True)
        return
    tuning = json.load(open(a.tuning))["tuning"] if os.path.exists(a.tuning) else {}
    jobs = batch_workload(a.jobs, seed=4)
    with Engine(**tuning) as eng:
        run_batch(eng, jobs[:8])
        _, t = batch_timing(eng, jobs)
    print(json.dumps(dict(t, tuning=tuning)), flush=True)


if __name__ == "__main__":
    main()
