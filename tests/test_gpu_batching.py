"""The batching engine: jobs submitted together share transport launches
(one batch scene holding their scenes side by side, every electron tagged
with its job).  Every electron's random stream depends only on its job's
seed, its primary index and its ancestry, and all tallies are integers (the
energy sums are fixed-point), so a job's results must be bit-identical
alone, in a mixed batch, and under any regroup interval or admission
schedule."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("back", "fwd", "se", "se_sq", "bse_sq", "back_energy", "back_energy_hist", "fwd_energy_hist",
        "elevation_hist", "azimuth_hist")


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    for k in ("steps", "electrons", "escapes", "watchdog_drops", "queue_overflow"):
        assert a["stats"][k] == b["stats"][k], k


def _jobs(M):
    """(compiled scene, centers, per-pixel, eV, seed) for a mix of scenes and energies."""
    from cudamonsel_amd import scenes
    out = []
    specs = [("PMMA", 500.0, 300, 11), ("Si", 1000.0, 200, 12), ("Cu", 2000.0, 100, 13), ("PMMA", 300.0, 500, 14)]
    for mat, e, per, seed in specs:
        sc = scenes.flat(mat, beam_eV=e)
        out.append((sc, sc.monte.compileScene(), np.array([[0.0, 0.0, sc.beam_z]] * 4), per, e, seed))
    sc = scenes.line_on_substrate()
    out.append((sc, sc.monte.compileScene(), scenes.scan_centers(np.linspace(-15, 15, 8), [0.0], sc.beam_z), 150,
                500.0, 15))
    return out


def test_mixed_batch_matches_jobs_alone(M):
    from cudamonsel_amd.abi import Engine
    jobs = _jobs(M)
    with Engine() as eng:
        sids = [eng.upload(cs) for (_, cs, _, _, _, _) in jobs]
        alone = []
        for sid, (sc, cs, cen, per, e, seed) in zip(sids, jobs):
            alone.append(eng.result(eng.submit(sid, cen, per, e, 0.5e-9, seed)))
        ids = [eng.submit(sid, cen, per, e, 0.5e-9, seed) for sid, (sc, cs, cen, per, e, seed) in zip(sids, jobs)]
        mixed = [eng.result(i) for i in ids]
    for a, b in zip(alone, mixed):
        _same(a, b)
        assert a["stats"]["steps"] > 0 and a["stats"]["queue_overflow"] == 0


@pytest.mark.parametrize("tuning", [dict(steps_per_launch=16), dict(primaries_per_launch=97),
                                    dict(max_jobs_in_flight=2, steps_per_launch=5), dict(block_size=128),
                                    dict(admission_order=1, primaries_per_launch=300), dict(sort_queue=-1),
                                    dict(geometry_batch=1), dict(geometry_batch=64, steps_per_launch=7),
                                    dict(spawn_handoff=-1),
                                    dict(spawn_handoff=-1, primaries_per_launch=500, steps_per_launch=9)],
                         ids=["regroup16", "admit97", "two-slots-regroup5", "block128", "fewest-left-first",
                              "unsorted-queue", "geometry-every-step", "geometry-full-wave-regroup7",
                              "no-handoff", "no-handoff-admit500-regroup9"])
def test_results_independent_of_tuning(M, tuning):
    from cudamonsel_amd.abi import Engine
    jobs = _jobs(M)[:3]
    with Engine() as ref, Engine(**tuning) as eng:
        for (sc, cs, cen, per, e, seed) in jobs:
            a = ref.result(ref.submit(ref.upload(cs), cen, per, e, 0.5e-9, seed))
            b = eng.result(eng.submit(eng.upload(cs), cen, per, e, 0.5e-9, seed))
            _same(a, b)


def test_tail_compaction_keeps_results(M):
    """A launch large enough for tail compaction (>= 65536 items): waves down
    to a few live lanes queue them for the next launch; the results do not
    depend on the threshold.  Regrouping every 7 steps with secondaries handed
    to free lanes, twice: whether an electron is queued again (which runs the
    containment test at its claim) or keeps its lane does not change it.
    The requeued queue (over 65536 electrons) is sorted by region and energy
    class between launches; the same run with the queue left unsorted
    matches, so the sort neither drops nor repeats an electron."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.line_on_substrate()
    cs = sc.monte.compileScene()
    cen = scenes.scan_centers(np.linspace(-20, 20, 64), [0.0], sc.beam_z)
    res = []
    for tuning in (dict(tail_requeue=-1), dict(), dict(tail_requeue=48), dict(tail_requeue=48, geometry_batch=8),
                   dict(steps_per_launch=7, spawn_handoff=-1), dict(steps_per_launch=7, spawn_handoff=-1),
                   dict(steps_per_launch=7, sort_queue=-1)):
        with Engine(**tuning) as eng:
            r = eng.result(eng.submit(eng.upload(cs), cen, 1100, 500.0, 0.5e-9, 21))
            assert r["stats"]["queue_overflow"] == 0 and r["stats"]["watchdog_drops"] == 0
            res.append(r)
    assert res[2]["stats"]["transport_launches"] > res[0]["stats"]["transport_launches"]
    for r in res[1:]:
        _same(res[0], r)


def test_many_small_jobs_of_uneven_size(M):
    """64 single-pixel jobs, 1e2..1e4 primaries, 0.2..5 keV, three materials,
    flat or step edge, all in flight together; each matches its run alone."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    rng = np.random.default_rng(4)
    scs = {}
    for mat in ("Si", "PMMA", "Cu"):
        for edge in (False, True):
            sc = scenes.step_edge(mat) if edge else scenes.flat(mat)
            scs[(mat, edge)] = (sc, sc.monte.compileScene())
    specs = []
    for k in range(64):
        mat = ("Si", "PMMA", "Cu")[rng.integers(3)]
        edge = bool(rng.integers(2))
        e = float(np.exp(rng.uniform(np.log(200.0), np.log(5000.0))))
        n = int(np.exp(rng.uniform(np.log(100), np.log(10000))))
        specs.append((mat, edge, e, n, 100 + k))
    with Engine() as eng:
        sids = {key: eng.upload(cs) for key, (sc, cs) in scs.items()}
        ids = []
        for mat, edge, e, n, seed in specs:
            sc = scs[(mat, edge)][0]
            ids.append(eng.submit(sids[(mat, edge)], [[0.0, 0.0, sc.beam_z]], n, e, 0.5e-9, seed))
        res = [eng.result(i, timeout_s=120) for i in ids]
        for j in (0, 17, 63):
            mat, edge, e, n, seed = specs[j]
            sc = scs[(mat, edge)][0]
            alone = eng.result(eng.submit(sids[(mat, edge)], [[0.0, 0.0, sc.beam_z]], n, e, 0.5e-9, seed))
            _same(alone, res[j])
    for r, (mat, edge, e, n, seed) in zip(res, specs):
        assert r["stats"]["primaries"] == n and r["stats"]["queue_overflow"] == 0
        assert 0 < r["back"][0] + r["fwd"][0]


def test_dead_items_do_not_end_a_launch(M):
    """A launch whose items include long runs of electrons that are dead when
    claimed -- primaries of a job whose beam starts outside the chamber, and
    the queued electrons of a job stopped by the generation limit -- still
    claims every item: a healthy job in the same launches matches its run
    alone bit for bit, the dead job reports nothing, and the stopped job
    counts its dropped electrons."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    pm = scenes.flat("PMMA", beam_eV=500.0)
    cu = scenes.flat("Cu", beam_eV=5000.0)
    # the Cu job needs 20 launches to admit its primaries (3,000 per launch)
    # and is stopped after 16; the healthy job, admitted first, completes in
    # a few launches alone and mixed
    tuning = dict(max_generations=16, primaries_per_launch=3000, admission_order=1)
    healthy = ([[0.0, 0.0, pm.beam_z]] * 4, 150, 500.0, 31)
    with Engine(**tuning) as eng:
        s_pm, s_cu = eng.upload(pm.monte.compileScene()), eng.upload(cu.monte.compileScene())
        alone = eng.result(eng.submit(s_pm, healthy[0], healthy[1], healthy[2], 0.5e-9, healthy[3]))
        outside = eng.submit(s_pm, [[0.0, 0.0, -0.5]] * 8, 20000, 500.0, 0.5e-9, 32)   # 0.5 m: beyond the chamber
        stopped = eng.submit(s_cu, [[0.0, 0.0, cu.beam_z]], 60000, 5000.0, 0.5e-9, 33)
        mixed = eng.submit(s_pm, healthy[0], healthy[1], healthy[2], 0.5e-9, healthy[3])
        r_out, r_stop, r_mix = eng.result(outside), eng.result(stopped), eng.result(mixed)
    assert r_mix["stats"]["queue_overflow"] == 0, (alone["stats"], r_mix["stats"])
    _same(alone, r_mix)
    assert alone["stats"]["queue_overflow"] == 0 and alone["stats"]["steps"] > 0
    assert r_out["stats"]["primaries"] == 160000 and r_out["stats"]["steps"] == 0
    assert int(r_out["back"].sum() + r_out["fwd"].sum()) == 0
    assert r_stop["stats"]["queue_overflow"] > 0


def test_cancel(M):
    """monsel_job_cancel: a queued job ends at once; a running one ends after
    its queued electrons are dropped, freeing its slot; jobs sharing its
    launches, and jobs after it, match their runs alone."""
    import time
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.line_on_substrate()
    cen = scenes.scan_centers(np.linspace(-20, 20, 64), [0.0], sc.beam_z)
    small = ([[0.0, 0.0, sc.beam_z]] * 4, 200, 500.0, 0.5e-9, 41)
    with Engine() as eng:
        sid = eng.upload(sc.monte.compileScene())
        alone = eng.result(eng.submit(sid, *small))
        # running: a 2.56 M-primary scan cancelled while a small job shares its launches
        big = eng.submit(sid, cen, 40000, 500.0, 0.5e-9, 42)
        mix = eng.submit(sid, *small)
        time.sleep(0.02)
        eng.cancel(big)
        assert eng.wait_status(big, timeout_s=60) in (-8, 0)
        _same(alone, eng.result(mix))
        eng.release(big)
        # after: the engine runs on with the freed slot and reservation
        _same(alone, eng.result(eng.submit(sid, *small)))
    with Engine(max_jobs_in_flight=1) as eng:
        sid = eng.upload(sc.monte.compileScene())
        first = eng.submit(sid, cen, 3000, 500.0, 0.5e-9, 43)
        queued = eng.submit(sid, *small)
        eng.cancel(queued)
        assert eng.wait_status(queued, timeout_s=60) == -8
        eng.release(queued)
        r = eng.result(first)
        assert r["stats"]["primaries"] == 64 * 3000 and r["stats"]["queue_overflow"] == 0
        # cancelling a completed job changes nothing
        done = eng.submit(sid, *small)
        assert eng.wait_status(done, timeout_s=60) == 0
        eng.cancel(done)
        _same(alone, eng.result(done))


def test_queue_sort_with_many_regions(M):
    """The between-launch sort on a scene of 148 regions (12 x 12 contact
    holes: 296 sort keys, a key x chunk table of many scan tiles), with
    every electron requeued after 7 steps so that each launch's queue is
    sorted: results equal the run with the queue left unsorted."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.contact_holes(radius_nm=6.0, depth_nm=40.0, pitch_nm=16.0, n=12)
    cs = sc.monte.compileScene()
    xs = np.linspace(-80.0, 80.0, 8)
    cen = scenes.scan_centers(xs, xs, sc.beam_z)
    res = []
    for tuning in (dict(steps_per_launch=7), dict(steps_per_launch=7, sort_queue=-1)):
        with Engine(**tuning) as eng:
            r = eng.result(eng.submit(eng.upload(cs), cen, 1200, 500.0, 0.5e-9, 31))
            assert r["stats"]["queue_overflow"] == 0 and r["stats"]["watchdog_drops"] == 0
            res.append(r)
    assert res[0]["stats"]["transport_launches"] > 3
    _same(res[0], res[1])


def test_scan_split_by_pixel_range_equals_whole(M):
    """How the path shards across GPUs: a scan split into pixel ranges, each
    a job whose primary_offset is the index of its first primary, runs the
    same electrons as the whole scan, so the per-pixel tallies concatenate and
    the histograms and counters add up to the whole scan's bit for bit."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    sc = scenes.line_on_substrate()
    cen = scenes.scan_centers(np.linspace(-24, 24, 48), [0.0], sc.beam_z)
    per = 400
    cuts = [0, 13, 30, 48]
    with Engine() as eng:
        sid = eng.upload(sc.monte.compileScene())
        whole = eng.result(eng.submit(sid, cen, per, 500.0, 0.5e-9, 41))
        ids = [eng.submit(sid, cen[a:b], per, 500.0, 0.5e-9, 41, primary_offset=a * per)
               for a, b in zip(cuts[:-1], cuts[1:])]
        parts = [eng.result(i) for i in ids]
    for k in ("back", "fwd", "se", "se_sq", "bse_sq", "back_energy"):
        assert np.array_equal(np.concatenate([p[k] for p in parts]), whole[k]), k
    for k in ("back_energy_hist", "fwd_energy_hist", "elevation_hist", "azimuth_hist"):
        assert np.array_equal(sum(p[k] for p in parts), whole[k]), k
    for k in ("primaries", "steps", "electrons", "escapes"):
        assert sum(p["stats"][k] for p in parts) == whole["stats"][k], k


def test_jobs_of_both_reference_modes_in_one_context(M):
    """Scenes of the reference-compatible and the fixed mode cannot share a
    batch scene: jobs submitted alternately in the two modes run in turns
    (a job of the other mode waits until the batch in flight drains), and
    each equals its run alone."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    cen = scenes.scan_centers(np.linspace(-15, 15, 8), [0.0], 0.0)
    specs = []
    for k, compat in enumerate((True, False, True, False, False, True)):
        sc = scenes.line_on_substrate(compat=compat)
        c = cen.copy()
        c[:, 2] = sc.beam_z
        specs.append((sc, sc.monte.compileScene(), c, 120 + 10 * k, 50 + k))
    with Engine() as eng:
        sids = [eng.upload(cs) for (_, cs, _, _, _) in specs]
        ids = [eng.submit(sid, c, per, 500.0, 0.5e-9, seed) for sid, (_, _, c, per, seed) in zip(sids, specs)]
        mixed = [eng.result(i, timeout_s=120) for i in ids]
        alone = [eng.result(eng.submit(sid, c, per, 500.0, 0.5e-9, seed))
                 for sid, (_, _, c, per, seed) in zip(sids, specs)]
    for a, b in zip(alone, mixed):
        _same(a, b)
