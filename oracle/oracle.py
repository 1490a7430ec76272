"""TEST INFRASTRUCTURE -- ctypes front end of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  The oracle consumes the same monsel_scene_desc as the GPU
engine (passed by address from cudamonsel_amd.CompiledScene.desc_address()).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_u64_p = ctypes.POINTER(ctypes.c_uint64)


class JobDesc(ctypes.Structure):
    """monsel_job_desc (include/monsel_gpu.h)"""
    _fields_ = [
        ("beam_energy", ctypes.c_double),
        ("beam_width", ctypes.c_double),
        ("n_pixels", ctypes.c_int32),
        ("primaries_per_pixel", ctypes.c_int32),
        ("pixel_center", c_double_p),
        ("seed", ctypes.c_uint64),
        ("n_energy_bins", ctypes.c_int32),
        ("max_steps", ctypes.c_int32),
        ("se_cutoff_ev", ctypes.c_double),
        ("primary_offset", ctypes.c_uint64),
        ("log_detected", ctypes.c_int64),
    ]


class Tallies(ctypes.Structure):
    """monsel_tallies"""
    _fields_ = [
        ("back", c_u64_p), ("fwd", c_u64_p), ("se", c_u64_p),
        ("se_sq", c_double_p), ("bse_sq", c_double_p), ("back_energy", c_double_p),
        ("back_energy_hist", c_u64_p), ("fwd_energy_hist", c_u64_p),
        ("elevation_hist", c_u64_p), ("azimuth_hist", c_u64_p),
    ]


class JobStats(ctypes.Structure):
    """monsel_job_stats"""
    _fields_ = [
        ("primaries", ctypes.c_uint64), ("electrons", ctypes.c_uint64), ("steps", ctypes.c_uint64),
        ("escapes", ctypes.c_uint64), ("watchdog_drops", ctypes.c_uint64), ("queue_overflow", ctypes.c_uint64),
        ("generations", ctypes.c_int32), ("transport_launches", ctypes.c_int32),
        ("device_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("transport_ms", ctypes.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_scene_create.restype = ctypes.c_void_p
        L.oracle_scene_create.argtypes = [ctypes.c_void_p]
        L.oracle_scene_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_run_job.argtypes = [ctypes.c_void_p, ctypes.POINTER(JobDesc), ctypes.POINTER(Tallies),
                                     ctypes.POINTER(JobStats), ctypes.c_int]
        for name in ("oracle_probe_rates", "oracle_probe_csd", "oracle_probe_elastic_angle",
                     "oracle_probe_elastic_sigma", "oracle_probe_find_end_of_step", "oracle_probe_locate",
                     "oracle_probe_nu_table", "oracle_probe_scatter", "oracle_probe_barrier"):
            getattr(L, name).restype = ctypes.c_int
        L.oracle_u_d1.restype = ctypes.c_double
        L.oracle_u_d1.argtypes = [c_double_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_double]
        L.oracle_u_d2.restype = ctypes.c_double
        L.oracle_hist_bin.restype = ctypes.c_int
        L.oracle_hist_bin.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.oracle_e0_from_dispersion.restype = ctypes.c_double
        L.oracle_e0_from_dispersion.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oracle_barrier_T.restype = ctypes.c_double
        L.oracle_barrier_T.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.oracle_child_lineage.restype = ctypes.c_uint64
        L.oracle_child_lineage.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def uptr(a):
    return a.ctypes.data_as(c_u64_p)


class OracleScene:
    """Oracle copy of a scene.  `compiled` is a cudamonsel_amd.CompiledScene
    (kept referenced while the descriptor is read) or a descriptor address."""

    def __init__(self, compiled):
        self.compiled = compiled
        addr = compiled if isinstance(compiled, int) else compiled.desc_address()
        self.h = lib().oracle_scene_create(ctypes.c_void_p(addr))
        if not self.h:
            raise RuntimeError("oracle scene: " + lib().oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_scene_destroy(self.h)
            self.h = None

    def run(self, centers, per_pixel, beam_energy_J, beam_width_m, seed, nbins, se_cutoff_eV=50.0,
            threads=0, max_steps=0, primary_offset=0):
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        n = centers.shape[0]
        job = JobDesc(beam_energy_J, beam_width_m, n, per_pixel, dptr(centers), seed, nbins, max_steps,
                      se_cutoff_eV, primary_offset)
        out = dict(back=np.zeros(n, np.uint64), fwd=np.zeros(n, np.uint64), se=np.zeros(n, np.uint64),
                   seSq=np.zeros(n), bseSq=np.zeros(n), backEnergy=np.zeros(n),
                   backEnergyHist=np.zeros(nbins + 2, np.uint64), fwdEnergyHist=np.zeros(nbins + 2, np.uint64),
                   elevationHist=np.zeros(182, np.uint64), azimuthHist=np.zeros(362, np.uint64))
        t = Tallies(uptr(out["back"]), uptr(out["fwd"]), uptr(out["se"]), dptr(out["seSq"]), dptr(out["bseSq"]),
                    dptr(out["backEnergy"]), uptr(out["backEnergyHist"]), uptr(out["fwdEnergyHist"]),
                    uptr(out["elevationHist"]), uptr(out["azimuthHist"]))
        st = JobStats()
        rc = lib().oracle_run_job(self.h, ctypes.byref(job), ctypes.byref(t), ctypes.byref(st), threads)
        if rc != 0:
            raise RuntimeError("oracle_run_job failed")
        out["stats"] = {k: getattr(st, k) for k, _ in JobStats._fields_}
        out["perPixel"] = per_pixel
        return out

    def primaries(self, centers, per_pixel, beam_width_m, seed, n, primary_offset=0):
        """Start points (m, n x 3) and regions of the job's first n primaries
        (GaussianBeam::createElectron as run_trajectory draws it)."""
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        job = JobDesc(1.0, beam_width_m, centers.shape[0], per_pixel, dptr(centers), seed, 1, 0, 50.0,
                      primary_offset)
        out = np.zeros((n, 4))
        if lib().oracle_probe_primaries(ctypes.c_void_p(self.h), ctypes.byref(job), int(n), dptr(out)):
            raise ValueError("bad primaries probe arguments")
        return out[:, :3], out[:, 3].astype(int)

    # ---- probes (float64 in, float64 out; eV / nm) ----
    def rates(self, msm, energies_eV, n_mech):
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)
        out = np.zeros((e.size, n_mech + 1))
        lib().oracle_probe_rates(ctypes.c_void_p(self.h), int(msm), e.size, dptr(e), dptr(out))
        return out

    def csd(self, msm, len_nm, energies_eV):
        l = np.ascontiguousarray(len_nm, dtype=np.float64)
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)
        out = np.zeros(e.size)
        lib().oracle_probe_csd(ctypes.c_void_p(self.h), int(msm), e.size, dptr(l), dptr(e), dptr(out))
        return out

    def elastic_angle(self, msm, element, energies_eV, u):
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 2)
        out = np.zeros(e.size)
        rc = lib().oracle_probe_elastic_angle(ctypes.c_void_p(self.h), int(msm), int(element), e.size, dptr(e), dptr(uu),
                                              dptr(out))
        if rc:
            raise ValueError("bad elastic probe arguments")
        return out

    def elastic_sigma(self, msm, element, energies_eV):
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)
        out = np.zeros(e.size)
        rc = lib().oracle_probe_elastic_sigma(ctypes.c_void_p(self.h), int(msm), int(element), e.size, dptr(e), dptr(out))
        if rc:
            raise ValueError("bad elastic probe arguments")
        return out

    def find_end_of_step(self, region, p0_nm, p1_nm):
        a = np.ascontiguousarray(p0_nm, dtype=np.float64).reshape(-1, 3)
        b = np.ascontiguousarray(p1_nm, dtype=np.float64).reshape(-1, 3)
        n = a.shape[0]
        u = np.zeros(n)
        nxt = np.zeros(n, np.int32)
        nrm = np.zeros((n, 3))
        lib().oracle_probe_find_end_of_step(ctypes.c_void_p(self.h), int(region), n, dptr(a), dptr(b), dptr(u),
                                            nxt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dptr(nrm))
        return u, nxt, nrm

    def locate(self, pos_nm):
        p = np.ascontiguousarray(pos_nm, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.int32)
        lib().oracle_probe_locate(ctypes.c_void_p(self.h), p.shape[0], dptr(p),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def nu_table(self, table, order, x, dim):
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, dim)
        out = np.zeros(xx.shape[0])
        rc = lib().oracle_probe_nu_table(ctypes.c_void_p(self.h), int(table), int(order), xx.shape[0], dptr(xx), dptr(out))
        if rc:
            raise ValueError("bad table probe arguments")
        return out

    def scatter(self, msm, mech, prev_eV, E_eV, dirs, u):
        """One scatter event per row; u is (n, 6), consumed in order.  Returns
        a dict of arrays: dir (n,3), E, has, sE, sdir (n,3), trapped, draws."""
        pe = np.ascontiguousarray(prev_eV, dtype=np.float64)
        e = np.ascontiguousarray(E_eV, dtype=np.float64)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 6)
        n = e.size
        out = np.zeros((n, 11))
        rc = lib().oracle_probe_scatter(ctypes.c_void_p(self.h), int(msm), int(mech), n, dptr(pe), dptr(e), dptr(d),
                                        dptr(uu), dptr(out))
        if rc:
            raise ValueError("bad scatter probe arguments")
        return dict(dir=out[:, 0:3], E=out[:, 3], has=out[:, 4] > 0, sE=out[:, 5], sdir=out[:, 6:9],
                    trapped=out[:, 9] > 0, draws=out[:, 10].astype(int))

    def barrier(self, region_from, region_to, dirs, normals, E_eV, u):
        """One barrier crossing per row with the outward barrier normal given.
        Returns dict: dir (n,3), E, region, sign (+1 along the normal)."""
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        nb = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(E_eV, dtype=np.float64)
        uu = np.ascontiguousarray(u, dtype=np.float64)
        n = e.size
        out = np.zeros((n, 6))
        rc = lib().oracle_probe_barrier(ctypes.c_void_p(self.h), int(region_from), int(region_to), n, dptr(d),
                                        dptr(nb), dptr(e), dptr(uu), dptr(out))
        if rc:
            raise ValueError("bad barrier probe arguments")
        return dict(dir=out[:, 0:3], E=out[:, 3], region=out[:, 4].astype(int), sign=out[:, 5].astype(int))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def stream(seed, lineage, primary, n):
    out = np.zeros(n)
    lib().oracle_stream(ctypes.c_uint64(seed), ctypes.c_uint64(lineage), ctypes.c_uint32(primary), n, dptr(out))
    return out


def hist_bin(val, lo, hi, nbins, compat):
    return lib().oracle_hist_bin(val, lo, hi, nbins, 1 if compat else 0)


def u_d1(f, x0, xinc, order, x):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().oracle_u_d1(dptr(f), f.size, x0, xinc, order, x)
