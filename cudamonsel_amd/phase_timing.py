"""Per-phase wave timing of the transport kernel on the line-scan workload.

Needs a library built with `python -m cudamonsel_amd.build --phase-timing`
(csrc/phase_timing.h).  Runs the 0.5 keV PMMA-line-on-Si scan (128 pixels at
1 nm pitch, primaries per pixel from the command line), reads the kernel's
counters and prints, per phase, its share of the wave cycles of the whole
loop, the cycles per entry and the lane utilisation on entry.

Usage: python -m cudamonsel_amd.phase_timing [per_pixel] [repeats] [--default-engine]
       python -m cudamonsel_amd.phase_timing --batch   (the 512-job batch)
(the engine runs the line-scan tuning of data/tuning_linescan.json unless
--default-engine is given)
"""
import ctypes
import sys

import numpy as np

PHASES = ["claim", "contain", "rates", "geom", "csd", "select", "elastic", "fitted", "polaron", "phonon",
          "tabulated", "barrier", "step", "tail", "rates_uni", "geom_uni", "safety", "uniforms", "rate_loop",
          "geom_iv", "geom_walk", "geom_loc"]


def read_counters(lib, device=0):
    fn = getattr(lib, "monsel_phase_timing_read", None)
    if fn is None:
        raise RuntimeError("libmonsel_gpu.so has no phase counters: build it with "
                           "`python -m cudamonsel_amd.build --phase-timing`")
    out = (ctypes.c_uint64 * (3 * len(PHASES)))()
    if fn(ctypes.c_int(device), out, ctypes.c_int(3 * len(PHASES))) != 0:
        raise RuntimeError("monsel_phase_timing_read failed")
    return np.array(out[:], dtype=np.float64).reshape(len(PHASES), 3)


def read_spans(lib, device=0, max_spans=1 << 18):
    """(step histogram [24], wave spans [n, 2] in 100 MHz ticks) recorded
    since the last read."""
    hist = (ctypes.c_uint64 * 24)()
    spans = (ctypes.c_uint64 * (2 * max_spans))()
    n = ctypes.c_int(0)
    if lib.monsel_phase_timing_spans(ctypes.c_int(device), hist, spans, ctypes.c_int(max_spans), ctypes.byref(n)) != 0:
        raise RuntimeError("monsel_phase_timing_spans failed")
    sp = np.array(spans[:2 * n.value], dtype=np.float64).reshape(-1, 2)
    return np.array(hist[:], dtype=np.float64), sp


def launches(spans):
    """Splits wave spans into launches (a launch's waves start together) and
    returns per launch (waves, duration ms, wave-slot occupancy, end times of
    the waves at 50/90/99/100 % as fractions of the duration)."""
    order = np.argsort(spans[:, 0])
    sp = spans[order]
    out = []
    i = 0
    while i < len(sp):
        j = i + 1
        end = sp[i, 1]
        while j < len(sp) and sp[j, 0] <= end:
            end = max(end, sp[j, 1])
            j += 1
        w = sp[i:j]
        t0, t1 = w[:, 0].min(), w[:, 1].max()
        dur = t1 - t0
        occ = float(np.sum(w[:, 1] - w[:, 0]) / (len(w) * dur)) if dur > 0 else 1.0
        q = np.quantile((w[:, 1] - t0) / dur, [0.5, 0.9, 0.99, 1.0]) if dur > 0 else np.ones(4)
        out.append((len(w), dur / 1e5, occ, q))
        i = j
    return out


def table(c):
    """Rows (phase, share of loop cycles, cycles per entry, entries, lane utilisation).
    The loop's cycles are those of its four top-level phases: the claim, the
    step's start (rates), the step's end (step) and the tail; the other
    phases lie within one of them."""
    loop = (c[PHASES.index("claim"), 0] + c[PHASES.index("rates"), 0] + c[PHASES.index("step"), 0] +
            c[PHASES.index("tail"), 0])
    rows = []
    for k, name in enumerate(PHASES):
        cyc, ent, lanes = c[k]
        rows.append((name, cyc / loop if loop else 0.0, cyc / ent if ent else 0.0, int(ent),
                     lanes / (64.0 * ent) if ent else 0.0))
    return rows


def main():
    import json
    import os
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    per_pixel = int(args[0]) if len(args) > 0 else 2000
    repeats = int(args[1]) if len(args) > 1 else 2
    import cudamonsel_amd as M
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import Engine
    tuning = {}
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning_linescan.json")
    if "--default-engine" not in sys.argv and os.path.exists(tfile):
        tuning = json.load(open(tfile))["tuning"]
    print(f"engine tuning: {tuning or 'defaults'}")
    sc = scenes.line_on_substrate(h_nm=50.0, w_nm=20.0, theta_deg=3.0, r_nm=5.0, length_nm=1000.0,
                                  line="PMMA", substrate="Si", beam_eV=500.0, sigma_nm=0.5, compat=True)
    centers = scenes.scan_centers(np.arange(-64.0, 64.0, 1.0), [0.0], sc.beam_z)
    lib = M.gpu_lib()
    steps = 0
    with Engine(**tuning) as eng:
        sid = eng.upload(sc.monte.compileScene())
        eng.result(eng.submit(sid, centers, per_pixel, 500.0, 0.5e-9, 1))   # warm-up
        read_counters(lib)
        read_spans(lib)
        for k in range(repeats):
            r = eng.result(eng.submit(sid, centers, per_pixel, 500.0, 0.5e-9, 2 + k))
            steps += r["stats"]["steps"]
    c = read_counters(lib)
    print(f"line scan 128 px x {per_pixel} x {repeats}: {steps} electron steps")
    print(f"{'phase':<10} {'share':>7} {'cyc/entry':>10} {'entries':>12} {'lanes':>6}")
    for name, share, cpe, ent, util in table(c):
        print(f"{name:<10} {share:7.3f} {cpe:10.0f} {ent:12d} {util:6.3f}")
    report(lib, c)


def report(lib, c):
    """Prints the phase table, the step-count histogram and the launches."""
    print(f"{'phase':<10} {'share':>7} {'cyc/entry':>10} {'entries':>12} {'lanes':>6}")
    for name, share, cpe, ent, util in table(c):
        print(f"{name:<10} {share:7.3f} {cpe:10.0f} {ent:12d} {util:6.3f}")
    hist, spans = read_spans(lib)
    print("steps per electron and launch (log2 buckets):",
          " ".join(f"{k}:{int(v)}" for k, v in enumerate(hist) if v))
    ls = launches(spans)
    tot = sum(ms for _, ms, _, _ in ls)
    if tot > 0:
        print(f"{len(ls)} launches, {tot:.1f} ms, time-weighted occupancy "
              f"{sum(ms * occ for _, ms, occ, _ in ls) / tot:.3f}")
    print(f"{'launch':>6} {'waves':>6} {'ms':>8} {'occupancy':>9}  wave ends at (50/90/99/100 %)")
    for k, (nw, ms, occ, q) in enumerate(ls):
        print(f"{k:6d} {nw:6d} {ms:8.3f} {occ:9.3f}  " + " ".join(f"{x:.2f}" for x in q))


def batch_main():
    """The 512-job batch workload on its tuning (data/tuning_batch.json)."""
    import json
    import os
    import cudamonsel_amd as M
    from cudamonsel_amd import serving
    from cudamonsel_amd.abi import Engine
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning_batch.json")
    tuning = json.load(open(tfile))["tuning"] if os.path.exists(tfile) else {}
    print(f"engine tuning: {tuning or 'defaults'}")
    jobs = serving.batch_workload(512, seed=4)
    lib = M.gpu_lib()
    with Engine(**tuning) as eng:
        serving.run_batch(eng, jobs[:8])   # warm-up
        read_counters(lib)
        read_spans(lib)
        res, t = serving.run_batch(eng, jobs)
    c = read_counters(lib)
    print(f"batch of 512 jobs: {sum(r['stats']['steps'] for r in res)} electron steps in {t:.3f} s")
    report(lib, c)


if __name__ == "__main__":
    if "--batch" in sys.argv:
        batch_main()
    else:
        main()
