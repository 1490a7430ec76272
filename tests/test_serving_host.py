"""Host logic of the batched and served forms (no GPU): workload
generators, admission by footprint, percentiles, raster geometry."""
import numpy as np

from cudamonsel_amd import serving


def test_batch_workload_matches_the_configuration():
    jobs = serving.batch_workload(512, seed=4)
    assert len(jobs) == 512
    e = np.array([j.beam_eV for j in jobs])
    n = np.array([j.primaries for j in jobs])
    assert e.min() >= 200 and e.max() <= 5000 and n.min() >= 100 and n.max() <= 100000
    # log-uniform: the median of log E sits near the middle of the log range
    assert abs(np.median(np.log(e)) - 0.5 * (np.log(200) + np.log(5000))) < 0.25
    assert {j.material for j in jobs} == {"Si", "PMMA", "Cu"}
    assert 0.4 < np.mean([j.step_edge for j in jobs]) < 0.6
    assert len({j.seed for j in jobs}) == 512


def test_request_stream_is_poisson_with_the_configured_mix():
    reqs = serving.request_stream(rate_per_s=2.0, duration_s=600.0, seed=5)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert 1100 < len(reqs) < 1300
    assert abs(gaps.mean() - 0.5) < 0.05 and abs(gaps.std() - 0.5) < 0.06   # exponential: mean = sd
    assert {r.side for r in reqs} == {32, 64, 128, 256}
    assert all(50 <= r.per_pixel <= 500 and 300 <= r.beam_eV <= 2000 for r in reqs)
    assert {r.geometry for r in reqs} == {"lines", "holes"}
    mean_prim = np.mean([r.primaries for r in reqs])
    assert 4.5e6 < mean_prim < 7.5e6   # E[side^2] * E[per pixel] = 21760 * 275


def test_admission_by_footprint():
    a = serving.Admission(100)
    assert a.fits(500)          # an idle server admits any one request
    a.admit(60)
    assert a.fits(40) and not a.fits(41)
    a.admit(40)
    a.release(60)
    assert a.fits(60) and a.inflight == 1 and a.used == 40


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert serving.percentile(v, 50) == 50 and serving.percentile(v, 99) == 99
    assert serving.percentile(v, 100) == 100 and serving.percentile([3.0], 99) == 3.0


def test_raster_centers_pitch_and_centring():
    c = serving.raster_centers(4, -7e-8)
    assert c.shape == (16, 3)
    assert np.allclose(np.unique(np.round(c[:, 0] * 1e9, 6)), [-1.5, -0.5, 0.5, 1.5])
    assert np.allclose(c[:, 2], -7e-8) and np.allclose(c[:, :2].mean(0), 0.0)


def test_per_pixel_plain_replays_service_times(monkeypatch):
    """The per-pixel estimate times `sample` single-pixel jobs per request,
    scales to the request's pixels and replays the stream one request at a
    time: with a fake engine whose jobs take a fixed time, the service times
    and queueing delays follow in closed form."""
    import time as _time

    class FakeEngine:
        def __init__(self):
            self.jobs = 0

        def upload(self, cs):
            return 1

        def submit(self, *a, **k):
            self.jobs += 1
            return self.jobs

        def result(self, jid):
            return {}

    clock = [0.0]
    monkeypatch.setattr(serving.time, "perf_counter", lambda: clock[0])
    eng = FakeEngine()
    orig = eng.result

    def timed_result(jid):
        clock[0] += 0.001   # 1 ms per pixel job
        return orig(jid)
    eng.result = timed_result
    monkeypatch.setattr(serving, "_request_scene", lambda cache, r: (type("S", (), {"beam_z": 0.0})(), None, 1))
    reqs = [serving.Request(arrival=0.0, rid=0, side=32, per_pixel=50, beam_eV=500.0, geometry="lines", geo_seed=0),
            serving.Request(arrival=0.5, rid=1, side=64, per_pixel=50, beam_eV=500.0, geometry="lines", geo_seed=0)]
    r = serving.per_pixel_plain(eng, reqs, sample=8)
    assert eng.jobs == 16
    s0, s1 = 1024 * 0.001, 4096 * 0.001          # pixels x 1 ms
    assert abs(r["wall_s"] - (s0 + s1)) < 1e-9        # the second request waits for the first
    lat = sorted([s0, s0 + s1 - 0.5])
    assert abs(r["p50_s"] - lat[0]) < 1e-9 and abs(r["p99_s"] - lat[1]) < 1e-9
    assert abs(r["images_per_hour"] - 3600.0 * 2 / (s0 + s1)) < 1e-6
