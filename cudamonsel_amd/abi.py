"""ctypes declarations of the C-ABI (include/monsel_gpu.h) for the parts the
pybind host API does not cover: direct job submission and the probes."""
import ctypes

import numpy as np

from . import gpu_lib, last_error

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)


class JobDesc(ctypes.Structure):
    _fields_ = [
        ("beam_energy", ctypes.c_double), ("beam_width", ctypes.c_double),
        ("n_pixels", ctypes.c_int32), ("primaries_per_pixel", ctypes.c_int32),
        ("pixel_center", c_double_p), ("seed", ctypes.c_uint64),
        ("n_energy_bins", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("se_cutoff_ev", ctypes.c_double), ("primary_offset", ctypes.c_uint64),
        ("log_detected", ctypes.c_int64),
    ]


class Detected(ctypes.Structure):
    """monsel_detected"""
    _fields_ = [
        ("electron", ctypes.c_uint64), ("primary", ctypes.c_uint32), ("step", ctypes.c_uint32),
        ("energy_ev", ctypes.c_double), ("position", ctypes.c_double * 3), ("theta", ctypes.c_double),
        ("phi", ctypes.c_double),
    ]


DETECTED_DTYPE = np.dtype([("electron", np.uint64), ("primary", np.uint32), ("step", np.uint32),
                           ("energy_ev", np.float64), ("position", np.float64, (3,)), ("theta", np.float64),
                           ("phi", np.float64)])


class Tallies(ctypes.Structure):
    _fields_ = [
        ("back", c_u64_p), ("fwd", c_u64_p), ("se", c_u64_p),
        ("se_sq", c_double_p), ("bse_sq", c_double_p), ("back_energy", c_double_p),
        ("back_energy_hist", c_u64_p), ("fwd_energy_hist", c_u64_p),
        ("elevation_hist", c_u64_p), ("azimuth_hist", c_u64_p),
    ]


class JobStats(ctypes.Structure):
    _fields_ = [
        ("primaries", ctypes.c_uint64), ("electrons", ctypes.c_uint64), ("steps", ctypes.c_uint64),
        ("escapes", ctypes.c_uint64), ("watchdog_drops", ctypes.c_uint64), ("queue_overflow", ctypes.c_uint64),
        ("generations", ctypes.c_int32), ("transport_launches", ctypes.c_int32),
        ("device_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("transport_ms", ctypes.c_double),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("max_generations", ctypes.c_int32),
        ("queue_capacity", ctypes.c_int32), ("max_jobs_in_flight", ctypes.c_int32),
        ("primaries_per_launch", ctypes.c_int32), ("steps_per_launch", ctypes.c_int32),
        ("admission_order", ctypes.c_int32), ("sort_queue", ctypes.c_int32),
        ("geometry_batch", ctypes.c_int32), ("spawn_handoff", ctypes.c_int32),
        ("tail_requeue", ctypes.c_int32),
    ]


SYMBOLS = [
    "monsel_last_error", "monsel_abi_version", "monsel_create", "monsel_destroy", "monsel_device_info",
    "monsel_ctx_counters",
    "monsel_scene_footprint", "monsel_scene_upload", "monsel_scene_release", "monsel_job_footprint",
    "monsel_job_submit", "monsel_job_poll", "monsel_job_wait", "monsel_job_read", "monsel_job_stats_get",
    "monsel_job_release", "monsel_run", "monsel_probe_rates", "monsel_probe_elastic_angle", "monsel_probe_csd",
    "monsel_probe_find_end_of_step", "monsel_probe_locate", "monsel_probe_nu_table", "monsel_probe_scatter",
    "monsel_probe_barrier", "monsel_probe_safety", "monsel_job_cancel", "monsel_probe_philox",
    "monsel_probe_hist_bin", "monsel_probe_primaries", "monsel_run_static", "monsel_job_read_detected",
]


def _fp(a):
    return a.ctypes.data_as(c_float_p)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


class Probe:
    """Device probes on one uploaded scene of a context."""

    def __init__(self, ctx_address, scene_id):
        self.L = gpu_lib()
        self.ctx = ctypes.c_void_p(ctx_address)
        self.scene = ctypes.c_uint64(scene_id)

    def rates(self, msm, energies_eV, n_mech):
        e = np.ascontiguousarray(energies_eV, dtype=np.float32)
        out = np.zeros((e.size, n_mech + 1), np.float32)
        _check(self.L.monsel_probe_rates(self.ctx, self.scene, int(msm), e.size, _fp(e), _fp(out)), "probe_rates")
        return out

    def elastic_angle(self, msm, element, energies_eV, u):
        e = np.ascontiguousarray(energies_eV, dtype=np.float32)
        uu = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        out = np.zeros(e.size, np.float32)
        _check(self.L.monsel_probe_elastic_angle(self.ctx, self.scene, int(msm), int(element), e.size, _fp(e), _fp(uu),
                                                 _fp(out)), "probe_elastic_angle")
        return out

    def csd(self, msm, len_nm, energies_eV):
        l = np.ascontiguousarray(len_nm, dtype=np.float32)
        e = np.ascontiguousarray(energies_eV, dtype=np.float32)
        out = np.zeros(e.size, np.float32)
        _check(self.L.monsel_probe_csd(self.ctx, self.scene, int(msm), e.size, _fp(l), _fp(e), _fp(out)), "probe_csd")
        return out

    def find_end_of_step(self, region, p0_nm, p1_nm):
        a = np.ascontiguousarray(p0_nm, dtype=np.float32).reshape(-1, 3)
        b = np.ascontiguousarray(p1_nm, dtype=np.float32).reshape(-1, 3)
        n = a.shape[0]
        u = np.zeros(n, np.float32)
        nxt = np.zeros(n, np.int32)
        nrm = np.zeros((n, 3), np.float32)
        _check(self.L.monsel_probe_find_end_of_step(self.ctx, self.scene, int(region), n, _fp(a), _fp(b), _fp(u),
                                                    nxt.ctypes.data_as(c_int_p), _fp(nrm)), "probe_find_end_of_step")
        return u, nxt, nrm

    def safety(self, region, pos_nm):
        p = np.ascontiguousarray(pos_nm, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.float32)
        _check(self.L.monsel_probe_safety(self.ctx, self.scene, int(region), p.shape[0], _fp(p), _fp(out)),
               "probe_safety")
        return out

    def locate(self, pos_nm):
        p = np.ascontiguousarray(pos_nm, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.int32)
        _check(self.L.monsel_probe_locate(self.ctx, self.scene, p.shape[0], _fp(p), out.ctypes.data_as(c_int_p)),
               "probe_locate")
        return out

    def nu_table(self, table, order, x, dim):
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, dim)
        out = np.zeros(xx.shape[0], np.float32)
        _check(self.L.monsel_probe_nu_table(self.ctx, self.scene, int(table), int(order), xx.shape[0], _fp(xx), _fp(out)),
               "probe_nu_table")
        return out

    def scatter(self, msm, mech, prev_eV, E_eV, dirs, u):
        """Same layout as the oracle's scatter probe (dict of arrays)."""
        pe = np.ascontiguousarray(prev_eV, dtype=np.float32)
        e = np.ascontiguousarray(E_eV, dtype=np.float32)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        uu = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 6)
        n = e.size
        out = np.zeros((n, 11), np.float32)
        _check(self.L.monsel_probe_scatter(self.ctx, self.scene, int(msm), int(mech), n, _fp(pe), _fp(e), _fp(d),
                                           _fp(uu), _fp(out)), "probe_scatter")
        return dict(dir=out[:, 0:3], E=out[:, 3], has=out[:, 4] > 0, sE=out[:, 5], sdir=out[:, 6:9],
                    trapped=out[:, 9] > 0)

    def primaries(self, centers_m, per_pixel, beam_eV, sigma_m, seed, n, primary_offset=0):
        """Start positions (nm, n x 3) and regions of the job's first n
        primaries, as the transport kernel creates them."""
        c = np.ascontiguousarray(centers_m, dtype=np.float64).reshape(-1, 3)
        jd = JobDesc(beam_eV * 1.60217653e-19, sigma_m, c.shape[0], int(per_pixel), c.ctypes.data_as(c_double_p),
                     int(seed), 50, 0, 50.0, int(primary_offset))
        out = np.zeros((n, 4), np.float32)
        _check(self.L.monsel_probe_primaries(self.ctx, self.scene, ctypes.byref(jd), int(n), _fp(out)),
               "probe_primaries")
        return out[:, :3], out[:, 3].astype(int)

    def barrier(self, region_from, region_to, dirs, normals, E_eV, u):
        """Same layout as the oracle's barrier probe (dict of arrays)."""
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        nb = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        e = np.ascontiguousarray(E_eV, dtype=np.float32)
        uu = np.ascontiguousarray(u, dtype=np.float32)
        n = e.size
        out = np.zeros((n, 6), np.float32)
        _check(self.L.monsel_probe_barrier(self.ctx, self.scene, int(region_from), int(region_to), n, _fp(d), _fp(nb),
                                           _fp(e), _fp(uu), _fp(out)), "probe_barrier")
        return dict(dir=out[:, 0:3], E=out[:, 3], region=out[:, 4].astype(int), sign=out[:, 5].astype(int))


def philox(ctx_address, ctr, key):
    """Device Philox4x32-10 words (n x 4, uint32) and their FP32 uniforms."""
    L = gpu_lib()
    c = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    k = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    n = c.shape[0]
    out = np.zeros((n, 4), np.uint32)
    u = np.zeros((n, 4), np.float32)
    p32 = ctypes.POINTER(ctypes.c_uint32)
    _check(L.monsel_probe_philox(ctypes.c_void_p(ctx_address), n, c.ctypes.data_as(p32), k.ctypes.data_as(p32),
                                 out.ctypes.data_as(p32), _fp(u)), "probe_philox")
    return out, u


def hist_bin(ctx_address, values, lo, hi, nbins, compat):
    """Device Histogram::bin of each value (int array)."""
    L = gpu_lib()
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros(v.size, np.int32)
    _check(L.monsel_probe_hist_bin(ctypes.c_void_p(ctx_address), v.size, v.ctypes.data_as(c_double_p),
                                   ctypes.c_double(lo), ctypes.c_double(hi), int(nbins), int(bool(compat)),
                                   out.ctypes.data_as(c_int_p)), "probe_hist_bin")
    return out


def probe_for(scene_obj, device=0):
    """Upload a cudamonsel_amd.scenes.Scene and return (Probe, CompiledScene)."""
    from . import Device
    cs = scene_obj.monte.compileScene()
    dev = Device.get(device)
    sid = dev.upload(cs)
    return Probe(dev.ctx_address(), sid), cs


class Engine:
    """A device context of its own (with launch tuning) driven through the
    C-ABI: scene upload, asynchronous job submission, results as arrays."""

    def __init__(self, device=0, **tuning):
        self.L = gpu_lib()
        self.ctx = ctypes.c_void_p()
        t = Tuning(**tuning)
        _check(self.L.monsel_create(int(device), ctypes.byref(t), ctypes.byref(self.ctx)), "monsel_create")
        self._keep = {}

    def close(self):
        if self.ctx:
            self.L.monsel_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, compiled):
        sid = ctypes.c_uint64()
        _check(self.L.monsel_scene_upload(self.ctx, ctypes.c_void_p(compiled.desc_address()), ctypes.byref(sid)),
               "monsel_scene_upload")
        return sid.value

    def submit(self, scene_id, centers_m, per_pixel, beam_eV, sigma_m, seed, nbins=50, se_cutoff_eV=50.0,
               max_steps=0, primary_offset=0, log_detected=0):
        c = np.ascontiguousarray(centers_m, dtype=np.float64).reshape(-1, 3)
        jd = JobDesc(beam_eV * 1.60217653e-19, sigma_m, c.shape[0], int(per_pixel), c.ctypes.data_as(c_double_p),
                     int(seed), int(nbins), int(max_steps), float(se_cutoff_eV), int(primary_offset),
                     int(log_detected))
        jid = ctypes.c_uint64()
        _check(self.L.monsel_job_submit(self.ctx, ctypes.c_uint64(scene_id), ctypes.byref(jd), ctypes.byref(jid)),
               "monsel_job_submit")
        self._keep[jid.value] = (c.shape[0], int(nbins), int(log_detected))
        return jid.value

    def counters(self):
        """(transport launches, electron steps of completed jobs, transport
        kernel ms) so far on this context."""
        la, st, ms = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        _check(self.L.monsel_ctx_counters(self.ctx, ctypes.byref(la), ctypes.byref(st), ctypes.byref(ms)),
               "monsel_ctx_counters")
        return la.value, st.value, ms.value

    def cancel(self, job_id):
        _check(self.L.monsel_job_cancel(self.ctx, ctypes.c_uint64(job_id)), "monsel_job_cancel")

    def wait_status(self, job_id, timeout_s=None):
        """monsel_job_wait's status (0 = completed, MONSEL_E_CANCELLED = -8, ...)."""
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        return self.L.monsel_job_wait(self.ctx, ctypes.c_uint64(job_id), ctypes.c_int64(us))

    def release(self, job_id):
        self._keep.pop(job_id, None)
        _check(self.L.monsel_job_release(self.ctx, ctypes.c_uint64(job_id)), "monsel_job_release")

    def poll(self, job_id):
        done = ctypes.c_int(0)
        _check(self.L.monsel_job_poll(self.ctx, ctypes.c_uint64(job_id), ctypes.byref(done)), "monsel_job_poll")
        return bool(done.value)

    def footprint(self, n_pixels, per_pixel, beam_eV, nbins=50):
        """Device bytes a job of this shape holds while in flight."""
        c = np.zeros((1, 3))
        jd = JobDesc(beam_eV * 1.60217653e-19, 0.5e-9, int(n_pixels), int(per_pixel), c.ctypes.data_as(c_double_p),
                     0, int(nbins), 0, 50.0, 0)
        nb = ctypes.c_size_t(0)
        _check(self.L.monsel_job_footprint(self.ctx, ctypes.byref(jd), ctypes.byref(nb)), "monsel_job_footprint")
        return nb.value

    def run(self, scene_id, centers_m, per_pixel, beam_eV, sigma_m, seed, nbins=50, se_cutoff_eV=50.0,
            max_steps=0, static=False):
        """One job through monsel_run, or with static=True through
        monsel_run_static (no device allocation); returns its tallies and
        statistics as result() does."""
        c = np.ascontiguousarray(centers_m, dtype=np.float64).reshape(-1, 3)
        n = c.shape[0]
        jd = JobDesc(beam_eV * 1.60217653e-19, sigma_m, n, int(per_pixel), c.ctypes.data_as(c_double_p),
                     int(seed), int(nbins), int(max_steps), float(se_cutoff_eV), 0, 0)
        r, t = self._tally_buffers(n, int(nbins))
        st = JobStats()
        fn = self.L.monsel_run_static if static else self.L.monsel_run
        _check(fn(self.ctx, ctypes.c_uint64(scene_id), ctypes.byref(jd), ctypes.byref(t), ctypes.byref(st)),
               "monsel_run_static" if static else "monsel_run")
        r["stats"] = {f: getattr(st, f) for f, _ in JobStats._fields_}
        return r

    @staticmethod
    def _tally_buffers(n, nb):
        r = {k: np.zeros(n, np.uint64) for k in ("back", "fwd", "se")}
        r.update({k: np.zeros(n) for k in ("se_sq", "bse_sq", "back_energy")})
        r["back_energy_hist"] = np.zeros(nb + 2, np.uint64)
        r["fwd_energy_hist"] = np.zeros(nb + 2, np.uint64)
        r["elevation_hist"] = np.zeros(182, np.uint64)
        r["azimuth_hist"] = np.zeros(362, np.uint64)
        t = Tallies()
        for k in ("back", "fwd", "se", "back_energy_hist", "fwd_energy_hist", "elevation_hist", "azimuth_hist"):
            setattr(t, k, r[k].ctypes.data_as(c_u64_p))
        for k in ("se_sq", "bse_sq", "back_energy"):
            setattr(t, k, r[k].ctypes.data_as(c_double_p))
        return r, t

    def result(self, job_id, timeout_s=None):
        """Waits for the job, returns its tallies and statistics, releases it."""
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        _check(self.L.monsel_job_wait(self.ctx, ctypes.c_uint64(job_id), ctypes.c_int64(us)), "monsel_job_wait")
        n, nb, nlog = self._keep.pop(job_id)
        r, t = self._tally_buffers(n, nb)
        _check(self.L.monsel_job_read(self.ctx, ctypes.c_uint64(job_id), ctypes.byref(t)), "monsel_job_read")
        st = JobStats()
        _check(self.L.monsel_job_stats_get(self.ctx, ctypes.c_uint64(job_id), ctypes.byref(st)), "monsel_job_stats_get")
        r["stats"] = {f: getattr(st, f) for f, _ in JobStats._fields_}
        if nlog > 0:
            # the count first, then a buffer for the records kept
            nd = ctypes.c_int64(0)
            _check(self.L.monsel_job_read_detected(self.ctx, ctypes.c_uint64(job_id), None, ctypes.c_int64(0),
                                                   ctypes.byref(nd)), "monsel_job_read_detected")
            buf = np.zeros(min(nlog, nd.value), DETECTED_DTYPE)
            if len(buf):
                _check(self.L.monsel_job_read_detected(self.ctx, ctypes.c_uint64(job_id),
                                                       buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(buf)),
                                                       ctypes.byref(nd)), "monsel_job_read_detected")
            r["detected"] = buf
            r["n_detected"] = nd.value
        self.L.monsel_job_release(self.ctx, ctypes.c_uint64(job_id))
        return r
