/*
 * monsel_gpu.hip -- MI355X (gfx950) transport engine behind include/monsel_gpu.h.
 *
 * Replaces the host trajectory loop of the reference,
 * MonteCarloSS::runMultipleTrajectories / runTrajectory / takeStep
 * (gov/nist/microanalysis/NISTMonte/MonteCarloSS.cu:107-231), with
 * generation-batched GPU transport:
 *   generation 0  : one lane per primary (GaussianBeam::createElectron,
 *                   GaussianBeam.cu:52-65), claimed dynamically by each wave;
 *   generation k+1: the secondaries spawned in generation k, appended to a
 *                   device queue with one wave-aggregated atomic per spawn burst.
 * Each lane runs its electron to completion, so the depth-first secondary
 * stack of the reference (MonteCarloSS.cu:165-207) becomes a breadth-first
 * queue; per-electron Philox streams keep the results independent of that
 * order.  Tallies are integer counters and FP64 sums (BackscatterStats.cu:147-170).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/monsel_gpu.h"
#include "device_scene.h"
#include "transport_device.h"

using namespace monsel_dev;

/* ===================================================================== */
/* errors                                                                 */
/* ===================================================================== */

static thread_local std::string g_err;
/* text of the HIP call in progress, named when a C++ exception escapes a
 * runtime call (e.g. from an interposed profiling library) */
static thread_local const char* g_hip_call = "";

static int set_err(int st, const std::string& m) {
  g_err = m;
  return st;
}

#define HIP_TRY(expr)                                                                                        \
  do {                                                                                                       \
    g_hip_call = #expr;                                                                                      \
    hipError_t _e = (expr);                                                                                  \
    if (_e != hipSuccess)                                                                                    \
      throw MonselError(MONSEL_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));                \
  } while (0)

struct MonselError : std::exception {
  int status;
  std::string msg;
  MonselError(int s, std::string m) : status(s), msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

/* the C-ABI status and message for an exception: a MonselError's own, an
 * allocation failure as MONSEL_E_NOMEM, anything else as a device error
 * named by the HIP call in progress */
static void status_of(const std::exception& ex, int& st, std::string& msg) {
  if (auto* me = dynamic_cast<const MonselError*>(&ex)) {
    st = me->status;
    msg = me->msg;
  } else if (dynamic_cast<const std::bad_alloc*>(&ex)) {
    st = MONSEL_E_NOMEM;
    msg = "host memory exhausted";
  } else {
    st = MONSEL_E_DEVICE;
    msg = std::string(ex.what()) + " (in " + g_hip_call + ")";
  }
}

/* ===================================================================== */
/* device kernels                                                         */
/* ===================================================================== */

constexpr int kQF = 7; /* float fields per queued electron: p, d, E */
constexpr int kQU = 7; /* u32 fields: region, primary, lineage lo/hi, job slot, steps, children */
constexpr int kEF = 7; /* float fields per escape record: tally origin, d, E */
constexpr int kEU = 5; /* u32 fields: primary, job slot << 1 | through-the-chamber-wall flag, lineage lo/hi, steps */
constexpr int kStatSteps = 0, kStatElectrons = 1, kStatEscapes = 2, kStatWatchdog = 3, kStatOverflow = 4;
/* a job's counters are kept in stripes of 8 u64 (waves add to the stripe of
 * their index), so that waves of many blocks do not contend for one address;
 * the host sums the stripes */
constexpr int kStatStripes = 64;

__device__ __forceinline__ unsigned long long* stat_stripe(unsigned long long* st) {
  return st + 8 * (((blockIdx.x * (blockDim.x >> 6)) + (threadIdx.x >> 6)) % kStatStripes);
}
constexpr int kMaxJobs = 1024; /* job-table slots of one context */
/* Back-escape energies are summed as integers in units of 2^-20 eV, so that
 * the sums are exact and independent of the order of the atomics. */
constexpr double kEnergyScale = 1048576.0;

/* One job's device record (a slot of the context's job table).  Every
 * electron carries its job's slot; jobs of one batch share the transport
 * launches and differ only in these records. */
struct JobRec {
  uint64_t lin0;          /* lineage of the job's primaries (from its seed) */
  float E0;               /* eV */
  float width;            /* nm */
  const double* centers;  /* m, [n_pixels][3] */
  int n_pixels;
  int ppp;
  uint32_t prim_base;     /* random-stream index of the job's primary 0 */
  int root;               /* the job's chamber region in the batch scene */
  int max_steps;
  int nbins;
  int maxSEbin;
  int killed;             /* set by the host: drop the job's electrons */
  double E0eV;
  double chamber_c[3];    /* nm */
  double chamber_r;       /* nm */
  uint32_t* prim_counts;  /* [primaries]: back-escapes, low 16 bits SE, high 16 bits BSE */
  uint32_t* prim_fwd;     /* [primaries]: forward escapes */
  unsigned long long* prim_energy; /* [primaries]: summed back-escape energy, units of kEnergyQuantum eV */
  unsigned long long* hist; /* back energy [nbins+2], fwd energy [nbins+2], elevation [182], azimuth [362] */
  unsigned long long* stats; /* [8] */
  float* wd_log;           /* optional: state of the first kWatchdogLog watchdog drops */
  unsigned int* wd_count;
  monsel_detected* det_log; /* optional: detected-electron records (det_max of them) */
  unsigned long long* det_count;
  unsigned long long det_max;
};
constexpr int kWatchdogLog = 16, kWatchdogFields = 10;

/* One transport launch ("generation"): new primaries described by
 * segments, then the queued electrons of the previous launch. */
struct GenDev {
  const JobRec* jobs;
  const uint32_t* seg_end;   /* [n_seg] running item count at the end of each segment */
  const uint32_t* seg_job;   /* [n_seg] job slot */
  const uint32_t* seg_first; /* [n_seg] first job-relative primary index */
  int n_seg;
  uint32_t n_prim;           /* new-primary items (= seg_end[n_seg-1]) */
  const float* in_f;
  const uint32_t* in_u;
  uint32_t in_cap;
  uint32_t in_n;
  unsigned int* claim;
  float* out_f;
  uint32_t* out_u;
  uint32_t out_cap;
  unsigned int* out_count;
  /* escape records, tallied by k_tally_escapes after the launch */
  float* esc_f;
  uint32_t* esc_u;
  uint32_t esc_cap;
  unsigned int* esc_count;
  int steps_per_launch;      /* regroup interval: steps before an electron returns to the queue (0 = none) */
  int geo_batch;             /* lanes of a wave that wait for findEndOfStep before the wave runs it */
  int handoff;               /* new secondaries go to idle lanes of their wave first */
  int tail_lanes;            /* tail compaction threshold (0 = off) */
};

/* launches with fewer items than this run their tail out (compaction would
 * only pass a few electrons from launch to launch) */
constexpr uint32_t kTailMinItems = 65536;
/* items a wave claims at a time (kept in reserve for its idle lanes);
 * 4 to 24 measure alike on the line scan, 128 and more starve other waves;
 * with the claim threshold below, 32 and 64 are 4 % and 8 % slower */
#ifndef MONSEL_CLAIM_CHUNK
#define MONSEL_CLAIM_CHUNK 16
#endif
constexpr uint32_t kClaimChunk = MONSEL_CLAIM_CHUNK;
/* idle lanes a wave gathers before it takes new work while other lanes still
 * step: a claim (a new primary, or a queued electron and its containment
 * check) costs the wave about as much for several lanes as for one, and a
 * wave claiming whenever one lane ends paid it every few steps (on the line
 * scan 16 measured best: 1 and 32 are 7 % and 3 % slower, 4 is 2 % slower) */
#ifndef MONSEL_CLAIM_MIN
#define MONSEL_CLAIM_MIN 16
#endif
constexpr int kClaimMin = MONSEL_CLAIM_MIN;
constexpr int kDefaultTailLanes = 32;

/* per-wave LDS table of the handoff: lane of the k-th spawning lane */
__device__ __forceinline__ int* lane_map() {
  __shared__ int l_giver[(kLaneBlock / 64) * 64];
  return l_giver + (threadIdx.x & ~63u);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long wave_ballot(bool p) { return __ballot(p); }

/* this lane's bit of a ballot.  Waterfall loops test membership this way
 * rather than by comparing with the readfirstlane value, which would let the
 * compiler substitute the per-lane value for the uniform one in the body. */
__device__ __forceinline__ bool lane_in(unsigned long long m) { return (m >> lane_id()) & 1ull; }

/* Histogram::bin with edges min + i*delta (Histogram.cu:73-79,
 * Algorithm.cu:5-14); compat = the reference's insertion-point semantics. */
__device__ __forceinline__ int hist_bin(double val, double min, double max, int nbins, bool compat) {
  const double delta = (max - min) / nbins;
  int l = 0, r = nbins;
  while (l <= r) {
    int mid = l + (r - l) / 2;
    double edge = mid == 0 ? min : min + mid * delta;
    if (edge == val) return mid;
    if (edge < val) l = mid + 1;
    else r = mid - 1;
  }
  if (compat) return l > nbins ? nbins : l;
  return l - 1;
}

/* Offsets of the four histograms of a job (Histogram.cu:25 slot layout:
 * 0 underrange, 1..n, n+1 overrange). */
__host__ __device__ __forceinline__ int hist_len(int nbins) { return 2 * (nbins + 2) + 182 + 362; }

/* BackscatterEvent -> BackscatterStats::actionPerformed (BackscatterStats.cu:147-170)
 * at the chamber wall point reached from (px,py,pz) along (dx,dy,dz).
 * Histogram counts go to `lh` (a block-private LDS copy) when given. */
__device__ __forceinline__ void tally_escape(const DevScene& S, const JobRec& J, uint32_t primary, float E, float px,
                                             float py, float pz, float dx, float dy, float dz, bool fromRoot,
                                             unsigned int* lh, uint64_t lineage, uint32_t steps) {
  double x = px, y = py, z = pz;
  if (fromRoot) {
    double cx = x - J.chamber_c[0], cy = y - J.chamber_c[1], cz = z - J.chamber_c[2];
    double ddx = dx, ddy = dy, ddz = dz;
    double dn = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    ddx /= dn;
    ddy /= dn;
    ddz /= dn;
    double R = J.chamber_r;
    double b = cx * ddx + cy * ddy + cz * ddz;
    double c = cx * cx + cy * cy + cz * cz - R * R;
    double disc = b * b - c;
    double t = -b + sqrt(disc > 0 ? disc : 0.);
    x += t * ddx;
    y += t * ddy;
    z += t * ddz;
  }
  const bool compat = (S.compat & MONSEL_COMPAT_HISTOGRAM_BIN) != 0;
  const double PI = 3.14159265358979323846;
  const double elevation = (PI / 2) - atan2(z, sqrt(x * x + y * y));
  double azimuth = atan2(y, x);
  if (azimuth < 0.0) azimuth = (2.0 * PI) + azimuth;
  const int nb2 = J.nbins + 2;
  const double kEeV = (double)E;
  const int bin = hist_bin(kEeV, 0.0, J.E0eV, J.nbins, compat);
  const bool fwd = elevation < (PI / 2.0);
  const int o_elev = 2 * nb2 + hist_bin(elevation, 0.0, PI, 180, compat) + 1;
  const int o_azim = 2 * nb2 + 182 + hist_bin(azimuth, 0.0, 2.0 * PI, 360, compat) + 1;
  const int o_e = (fwd ? nb2 : 0) + bin + 1;
  if (lh) {
    atomicAdd(&lh[o_elev], 1u);
    atomicAdd(&lh[o_azim], 1u);
    atomicAdd(&lh[o_e], 1u);
  } else {
    atomicAdd(&J.hist[o_elev], 1ull);
    atomicAdd(&J.hist[o_azim], 1ull);
    atomicAdd(&J.hist[o_e], 1ull);
  }
  if (J.det_log) {
    /* BackscatterStats.cu:166-168: the detected electron's Datum */
    const unsigned long long k = atomicAdd(J.det_count, 1ull);
    if (k < J.det_max) {
      monsel_detected* d = J.det_log + k;
      const double dn = sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
      d->electron = lineage;
      d->primary = primary;
      d->step = steps;
      d->energy_ev = (double)E;
      d->position[0] = x * 1e-9;
      d->position[1] = y * 1e-9;
      d->position[2] = z * 1e-9;
      d->theta = acos(fmin(1.0, fmax(-1.0, (double)dz / dn)));
      d->phi = atan2((double)dy, (double)dx);
    }
  }
  const uint32_t rel = primary - J.prim_base;
  if (fwd) {
    atomicAdd(&J.prim_fwd[rel], 1u);
  } else {
    const bool se = bin >= 0 && bin < J.maxSEbin;
    atomicAdd(&J.prim_counts[rel], se ? 1u : 0x10000u);
    atomicAdd(&J.prim_energy[rel], (unsigned long long)__double2ll_rn((double)E * kEnergyScale));
  }
}

enum StepResult { STEP_CONTINUE = 0, STEP_DEAD = 1, STEP_ESCAPED = 2, STEP_WATCHDOG = 3 };

/* ExpQMBarrierSM.cu:98-219: the crossing from the electron's region into
 * `next`, given the outward barrier normal (nbx, nby, nbz) resolved from the
 * struck shape.  Moves the electron a nudge along +nb (returns +1) or -nb
 * (returns -1); u1 decides a quantum-mechanical transmission. */
__device__ __forceinline__ int barrier_core(const DevScene& S, const MONSEL_CS DMSM& msm, const MONSEL_CS DMSM& nmsm, Elec& e, int next,
                                            float nbx, float nby, float nbz, float u1) {
  const float deltaU = -msm.cb + nmsm.cb;
  const bool compat = (S.compat & MONSEL_COMPAT_BARRIER_NORMAL) != 0;
  if (compat) {
    nbx = e.dx; nby = e.dy; nbz = e.dz; /* ExpQMBarrierSM.cu:111 */
  } else {
    float m2 = nbx * nbx + nby * nby + nbz * nbz;
    if (m2 > 0.f) {
      float inv = __builtin_amdgcn_rsqf(m2);
      nbx *= inv; nby *= inv; nbz *= inv;
    } else {
      nbx = e.dx; nby = e.dy; nbz = e.dz;
    }
  }
  const float cosalpha = compat ? 1.f : (e.dx * nbx + e.dy * nby + e.dz * nbz);
  const float dd = nudge_nm(e.px, e.py, e.pz);
  if (cosalpha <= 0.f) {
    e.px -= dd * nbx; e.py -= dd * nby; e.pz -= dd * nbz;
    return -1;
  }
  if (deltaU == 0.f) {
    e.px += dd * nbx; e.py += dd * nby; e.pz += dd * nbz;
    e.region = next;
    return 1;
  }
  const float kE0 = e.E;
  const float perpE = kE0 <= 0.f ? 0.f : cosalpha * cosalpha * kE0;
  bool transmits;
  float rootPerpE = 0.f, rootDiff = 0.f;
  if (perpE == 0.f || perpE <= deltaU) transmits = false;
  else {
    rootPerpE = fsqrt(perpE);
    rootDiff = fsqrt(perpE - deltaU);
    if (msm.barrier_kind == MONSEL_BARRIER_EXPQM_CLASSICAL) transmits = true;
    else {
      float tp;
      if (msm.lambda == 0.f) { /* ExpQMBarrierSM.cu:30-34 */
        float ks = rootPerpE + rootDiff;
        tp = (4.f * rootPerpE * rootDiff) / (ks * ks);
      } else { /* ExpQMBarrierSM.cu:36-48 */
        float k1 = msm.lambdaFactor * rootPerpE;
        if (k1 > 50.f) tp = 1.f;
        else {
          float k2 = msm.lambdaFactor * rootDiff;
          float ratio = sinhf(k1 - k2) / sinhf(k1 + k2);
          tp = 1.f - ratio * ratio;
        }
      }
      transmits = u1 < tp;
    }
  }
  if (transmits) {
    float factor = cosalpha * ((rootDiff / rootPerpE) - 1.f);
    float nx = e.dx + factor * nbx, ny = e.dy + factor * nby, nz = e.dz + factor * nbz;
    float inv = __builtin_amdgcn_rsqf(nx * nx + ny * ny + nz * nz);
    e.dx = nx * inv; e.dy = ny * inv; e.dz = nz * inv;
    e.E = kE0 - deltaU;
    e.region = next;
    e.px += dd * nbx; e.py += dd * nby; e.pz += dd * nbz;
    return 1;
  }
  float two = 2.f * cosalpha;
  float nx = e.dx - two * nbx, ny = e.dy - two * nby, nz = e.dz - two * nbz;
  float inv = __builtin_amdgcn_rsqf(nx * nx + ny * ny + nz * nz);
  e.dx = nx * inv; e.dy = ny * inv; e.dz = nz * inv;
  e.px -= dd * nbx; e.py -= dd * nby; e.pz -= dd * nbz;
  return -1;
}

/* Per-lane rates of the step in progress (mechanism k of the region's scatter
 * model), kept between the step's start and its end. */
__device__ __forceinline__ float* lane_rates() {
  __shared__ float l_rate[kMaxMech * kLaneBlock];
  return l_rate + threadIdx.x;
}

/* Per-lane terms sigma_i * atoms_i of the elastic mechanism's rate (the
 * region model's first elastic mechanism), kept between the step's start and
 * its scatter so that the element choice reuses them.  The choice needs all
 * terms but the last, so two slots serve materials of up to 3 elements
 * (larger ones recompute the terms at the scatter). */
constexpr int kKeptTerms = 2;
__device__ __forceinline__ float* lane_esig() {
  __shared__ float l_esig[kKeptTerms * kLaneBlock];
  return l_esig + threadIdx.x;
}

/* Per-lane anchor of the safety bound: e.safe bounds the distance to the
 * surfaces from this point, so from the electron's position it is at least
 * e.safe - |p - anchor| (triangle inequality).  A random walk drifts from
 * its anchor more slowly than its path grows, so the bound outlasts one
 * shrunk by every step. */
__device__ __forceinline__ float* lane_anchor() {
  __shared__ float l_anchor[3 * kLaneBlock];
  return l_anchor + threadIdx.x;
}
__device__ __forceinline__ void set_anchor(float x, float y, float z) {
  float* a = lane_anchor();
  a[0] = x;
  a[kLaneBlock] = y;
  a[2 * kLaneBlock] = z;
}
__device__ __forceinline__ float from_anchor(float x, float y, float z) {
  const float* a = lane_anchor();
  const float dx = x - a[0], dy = y - a[kLaneBlock], dz = z - a[2 * kLaneBlock];
  return fsqrt(dx * dx + dy * dy + dz * dz);
}

/* What the start of a step leaves for its end. */
struct StepStart {
  float freePath;
  float total;   /* summed mechanism rates (0 in a null material) */
  float u1, u2, u3;
  bool geo;      /* the step may reach a boundary: findEndOfStep must run */
};

/* MonteCarloSS::takeStep (MonteCarloSS.cu:107-163), first half: the
 * mechanism rates and the free path (MONSEL_MaterialScatterModel::
 * randomMeanPathLength, :64-91).  The electron is located (its region
 * contains it).  A free path that ends at least a nudge short of the
 * safety bound (anchored at lane_anchor, less the distance moved since)
 * cannot reach a boundary; every other step needs the geometry. */
__device__ __forceinline__ void begin_step_in(const DevScene& S, int region, Elec& e, StepStart& st) {
  PT_BEGIN(t_ra);
  const MONSEL_CS DMSM& msm = S.msms[S.regions[region].msm];
  float u[4];
  PT_BEGIN(t_un);
  step_uniforms_lo(e, kPhiloxKey0, kPhiloxKey1, u);
  PT_END(PH_UNIFORMS, t_un);
  float* rates = lane_rates();
  float total = 0.f;
  float freePath;
  const float maxFreePath = 2.f * S.chamber_r;
  if (msm.is_null) {
    freePath = 1.0e9f; /* NullMaterialScatterModel.cu:32-35: 1 m */
  } else {
    PT_BEGIN(t_rl);
    for (int k = 0; k < msm.n_mech; ++k) {
      const MONSEL_CS DMech& m = S.mechs[msm.mech0 + k];
      float rk;
      if (k == msm.elastic_k) {
        /* elastic_rate, keeping its terms */
        float* esig = lane_esig();
        const bool keep = m.nelem <= kKeptTerms + 1;
        rk = 0.f;
        for (int i = 0; i < m.nelem; ++i) {
          const MONSEL_CS DElem& el = S.elems[m.elem0 + i];
          const float c = elem_sigma(S, el, e.E) * el.natoms;
          if (keep && i < kKeptTerms) esig[i * kLaneBlock] = c;
          rk += c;
        }
      } else {
        rk = mech_rate(S, m, e.E);
      }
      rates[k * kLaneBlock] = rk;
      total += rk;
    }
    PT_END(PH_RATE_LOOP, t_rl);
    if (total != 0.f) {
      float fp = -flog(u[0]) / total;
      freePath = fp > maxFreePath ? maxFreePath : fp;
    } else {
      freePath = maxFreePath;
    }
  }
  st.freePath = freePath;
  st.total = total;
  st.u1 = u[1];
  st.u2 = u[2];
  st.u3 = u[3];
  const float margin = nudge_nm(e.px, e.py, e.pz);
  const float current = e.safe > 0.f ? e.safe - from_anchor(e.px, e.py, e.pz) : 0.f;  /* bound here */
  st.geo = !(freePath + margin <= current);
  PT_END(PH_RATES, t_ra);
}

/* The lanes of a wave take their turn by region (one pass per region present,
 * normally one): within a pass the region index is wave-uniform, so the
 * scene records it selects are read with scalar loads. */
__device__ __forceinline__ void begin_step(const DevScene& S, Elec& e, StepStart& st) {
  PT_UNIFORM(PH_RATES_UNI, e.region);
  for (;;) {
    const int r0 = __builtin_amdgcn_readfirstlane(e.region);
    if (lane_in(wave_ballot(e.region == r0))) {
      begin_step_in(S, r0, e, st);
      break;
    }
  }
}

/* The check that opens MonteCarloSS::takeStep (MonteCarloSS.cu:107-115):
 * re-locates an electron whose region (r == e.region, passed separately
 * where it is wave-uniform) does not contain it.  Returns false when no
 * region does. */
__device__ __forceinline__ bool locate(const DevScene& S, const JobRec* jobs, Elec& e, int r) {
  PT_BEGIN(t_co);
  e.safe = 0.f;
  if (r < 0 || !shape_contains(S, S.regions[r].sh, e.px, e.py, e.pz))
    e.region = containing_sub_region(S, jobs[e.job].root, e.px, e.py, e.pz);
  PT_END(PH_CONTAIN, t_co);
  return e.region >= 0;
}

/* Second half of takeStep: the end of the step (findEndOfStep when st.geo),
 * the move, slowing down, and scatter, barrier crossing or escape.  Ends
 * with the containment check of the next step's start wherever the electron
 * may have left its region (a boundary step, or one that ended within a
 * nudge of a surface). */
template <bool TAB>
__device__ __forceinline__ int finish_step_in(const DevScene& S, const JobRec* jobs, const int region, Elec& e,
                                              Spawn& sp, int maxSteps, const StepStart& st) {
  sp.has = false;
  const MONSEL_CS DMSM& msm = S.msms[S.regions[region].msm];
  const float freePath = st.freePath;
  const float px = e.px, py = e.py, pz = e.pz;
  StepEnd se;
  if (st.geo) {
    PT_BEGIN(t_ge);
    se = find_end_of_step(S, region, px, py, pz, e.dx, e.dy, e.dz, freePath);
    PT_END(PH_GEOM, t_ge);
  } else {
    se.t = freePath;
    se.next = region;
    se.boundary = false;
    se.winner = -1;
  }
  float len = se.t;
  e.px = px + len * e.dx;
  e.py = py + len * e.dy;
  e.pz = pz + len * e.dz;
  /* the safety bound: a geometry step's is its end point's; a step within
   * the bound keeps the anchored one.  Where findEndOfStep found no boundary
   * for some lane, the wave computes the end point's bound; the pass costs
   * the wave the same whichever lanes take part, so every lane of it that
   * crossed no boundary refreshes its bound (and keeps the better one). */
  if (st.geo || se.boundary) e.safe = 0.f;
  if (!msm.is_null && wave_ballot(st.geo && !se.boundary) != 0ull && !se.boundary) {
    PT_BEGIN(t_sf);
    const float fresh = region_safety(S, region, e.px, e.py, e.pz);
    const float kept = e.safe > 0.f ? e.safe - from_anchor(e.px, e.py, e.pz) : 0.f;
    if (fresh > kept) {
      e.safe = fresh;
      set_anchor(e.px, e.py, e.pz);
    }
    PT_END(PH_SAFETY, t_sf);
  }
  if (se.boundary && se.next == region) {
    /* a grazing crossing: the point a nudge beyond the boundary is back in
     * this region (a corner clipped within the nudge, or FP32 rounding
     * between the crossing walk and the containment test).  The step still
     * ends there as in RegionBase::findEndOfStep, but the electron is carried
     * past the feature so that the next step cannot stop at it again. */
    const float dd = nudge_nm(e.px, e.py, e.pz);
    e.px += dd * e.dx;
    e.py += dd * e.dy;
    e.pz += dd * e.dz;
  }
  float prevE = e.E;
  PT_BEGIN(t_cs);
  if (!msm.is_null) e.E += csd_compute(msm.csd, len, e.E);
  PT_END(PH_CSD, t_cs);
  e.steps++;
  if ((int)e.steps > maxSteps || !(e.E == e.E)) return STEP_WATCHDOG;
  if (e.E < msm.minE) return STEP_DEAD;
  if (se.next == region) {
    /* MONSEL_MaterialScatterModel::scatter (:93-134), rates at prevE */
    if (!msm.is_null && st.total != 0.f) {
      PT_BEGIN(t_se);
      const float* rates = lane_rates();
      const float r = st.u1 * st.total;
      /* first mechanism whose cumulative rate reaches r (MONSEL_MaterialScatterModel.cu:129-131);
       * the running sum repeats the order of `total` */
      int idx = 0;
      float rate = rates[0];
      float cum = rate;
      while (idx < msm.n_mech - 1 && cum < r) {
        ++idx;
        rate = rates[idx * kLaneBlock];
        cum += rate;
      }
      bool trapped = false;
      float u[kStepUniforms];
      u[0] = 0.f;
      u[1] = st.u1;
      u[2] = st.u2;
      u[3] = st.u3;
      step_uniforms_hi(e, kPhiloxKey0, kPhiloxKey1, u);
      PT_END(PH_SELECT, t_se);
      sp = mech_scatter<TAB>(S, S.mechs[msm.mech0 + idx], e, prevE, rate, trapped, u,
                             idx == msm.elastic_k && S.mechs[msm.mech0 + idx].nelem <= kKeptTerms + 1 ? lane_esig()
                                                                                               : nullptr);
      if (sp.has) {
        /* trackSecondaryElectron (MonteCarloSS.cu:165-174) */
        if (!(sp.E > msm.minE)) sp.has = false;
      }
      if (trapped || e.E < msm.minE) return STEP_DEAD;
    }
    if (st.geo && !(e.safe > nudge_nm(e.px, e.py, e.pz)) && !locate(S, jobs, e, region)) return STEP_DEAD;
    return STEP_CONTINUE;
  }
  if (se.next >= 0) {
    /* ExpQMBarrierSM::barrierScatter (ExpQMBarrierSM.cu:50-219) */
    PT_BEGIN(t_ba);
    int next = se.next;
    if (msm.is_null || msm.barrier_kind == MONSEL_BARRIER_NONE) {
      e.region = next;
    } else {
      /* ExpQMBarrierSM.cu:77-96: outward normal of the struck shape, negated
       * when entering a child region (unused in compat mode) */
      float nbx = 0.f, nby = 0.f, nbz = 0.f;
      if (!(S.compat & MONSEL_COMPAT_BARRIER_NORMAL)) {
        if (is_containing_region(S, region, next)) {
          int struck = next;
          while (S.regions[struck].parent != region) struck = S.regions[struck].parent;
          if (struck == se.winner) {
            nbx = -se.winNx; nby = -se.winNy; nbz = -se.winNz;
          } else {
            ShapeHit h = shape_first(S, S.regions[struck].sh, px, py, pz, e.dx, e.dy, e.dz, freePath);
            nbx = -h.nx; nby = -h.ny; nbz = -h.nz;
          }
        } else {
          nbx = se.ownNx; nby = se.ownNy; nbz = se.ownNz;
        }
      }
      barrier_core(S, msm, S.msms[S.regions[next].msm], e, next, nbx, nby, nbz, st.u1);
    }
    /* MonteCarloSS.cu:146-148: nudge along the new direction; the region is
     * re-located by the check that opens the next step */
    float dd = nudge_nm(e.px, e.py, e.pz);
    e.px += dd * e.dx;
    e.py += dd * e.dy;
    e.pz += dd * e.dz;
    PT_END(PH_BARRIER, t_ba);
    if (!locate(S, jobs, e, e.region)) return STEP_DEAD;
    return STEP_CONTINUE;
  }
  /* left the chamber: BackscatterEvent.  The escape is recorded for
   * k_tally_escapes: through the chamber wall the tally point is recomputed
   * in FP64 from the step's start, so that start replaces the position and
   * e.region carries the flag. */
  const bool fromRoot = S.regions[region].parent < 0 && se.winner < 0;
  if (fromRoot) {
    e.px = px;
    e.py = py;
    e.pz = pz;
  }
  e.region = fromRoot ? 1 : 0;
  return STEP_ESCAPED;
}

/* finish_step_in by region, as begin_step */
template <bool TAB>
__device__ __forceinline__ int finish_step(const DevScene& S, const JobRec* jobs, Elec& e, Spawn& sp, int maxSteps,
                                           const StepStart& st) {
  PT_UNIFORM(PH_GEOM_UNI, e.region);
  int res;
  for (;;) {
    const int r0 = __builtin_amdgcn_readfirstlane(e.region);
    if (lane_in(wave_ballot(e.region == r0))) {
      res = finish_step_in<TAB>(S, jobs, r0, e, sp, maxSteps, st);
      break;
    }
  }
  return res;
}

/* GaussianBeam::createElectron (GaussianBeam.cu:52-65) for primary `idx` of
 * job slot `job`. */
__device__ __forceinline__ void make_primary(const DevScene& S, const JobRec& J, uint32_t job, uint32_t idx, Elec& e) {
  const uint32_t pix = idx / (uint32_t)J.ppp;
  e.primary = J.prim_base + idx;
  e.lin_lo = (uint32_t)J.lin0;
  e.lin_hi = (uint32_t)(J.lin0 >> 32);
  e.job = job;
  e.nchild = 0;
  e.steps = 0;
  uint32_t a0, a1, a2, a3;
  philox4x32_10(kPrimaryBlock, e.lin_lo, e.lin_hi, e.primary, kPhiloxKey0, kPhiloxKey1, a0, a1, a2, a3);
  const float u1 = u01(a0), u2 = u01(a1);
  float r = fsqrt(-2.f * flog(u1)) * J.width;
  float s, c;
  __sincosf(kTwoPI * u2, &s, &c);
  const double* cen = J.centers + 3 * (size_t)pix;
  e.px = (float)(cen[0] * 1e9) + r * c;
  e.py = (float)(cen[1] * 1e9) + r * s;
  e.pz = (float)(cen[2] * 1e9);
  e.dx = 0.f;
  e.dy = 0.f;
  e.dz = 1.f;
  e.E = J.E0;
  e.safe = 0.f;
  e.region = containing_sub_region(S, J.root, e.px, e.py, e.pz);
}

/* Appends electrons to the next launch's queue: one atomic per wave. */
__device__ __forceinline__ void queue_append(const GenDev& G, bool has, float px, float py, float pz, float dx, float dy,
                                             float dz, float E, int region, uint32_t primary, uint32_t lin_lo,
                                             uint32_t lin_hi, uint32_t job, uint32_t steps, uint32_t nchild,
                                             unsigned long long lanemask_lt, uint32_t& nover) {
  const unsigned long long m = wave_ballot(has);
  if (!m) return;
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(G.out_count, (unsigned int)__popcll(m));
  base = __shfl(base, leader);
  if (!has) return;
  const uint32_t slot = base + (uint32_t)__popcll(m & lanemask_lt);
  if (slot >= G.out_cap) {
    nover++;
    return;
  }
  const uint32_t cap = G.out_cap;
  G.out_f[0 * cap + slot] = px;
  G.out_f[1 * cap + slot] = py;
  G.out_f[2 * cap + slot] = pz;
  G.out_f[3 * cap + slot] = dx;
  G.out_f[4 * cap + slot] = dy;
  G.out_f[5 * cap + slot] = dz;
  G.out_f[6 * cap + slot] = E;
  G.out_u[0 * cap + slot] = (uint32_t)region;
  G.out_u[1 * cap + slot] = primary;
  G.out_u[2 * cap + slot] = lin_lo;
  G.out_u[3 * cap + slot] = lin_hi;
  G.out_u[4 * cap + slot] = job;
  G.out_u[5 * cap + slot] = steps;
  G.out_u[6 * cap + slot] = nchild;
}

/* adds one lane's counters to its job's statistics (per-lane atomics) */
__device__ __forceinline__ void flush_lane_stats(const GenDev& G, uint32_t job, uint32_t& nsteps, uint32_t& nelec,
                                                 uint32_t& nesc, uint32_t& nwd, uint32_t& nover) {
  unsigned long long* st = stat_stripe(G.jobs[job].stats);
  if (nsteps) atomicAdd(&st[kStatSteps], (unsigned long long)nsteps);
  if (nelec) atomicAdd(&st[kStatElectrons], (unsigned long long)nelec);
  if (nesc) atomicAdd(&st[kStatEscapes], (unsigned long long)nesc);
  if (nwd) atomicAdd(&st[kStatWatchdog], (unsigned long long)nwd);
  if (nover) atomicAdd(&st[kStatOverflow], (unsigned long long)nover);
  nsteps = nelec = nesc = nwd = nover = 0;
}

/* waves per SIMD the register allocation of k_transport targets */
#ifndef MONSEL_TRANSPORT_WAVES_PER_EU
#define MONSEL_TRANSPORT_WAVES_PER_EU 4
#endif
template <bool TAB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MONSEL_TRANSPORT_WAVES_PER_EU)))
k_transport(DevScene S, GenDev G) {
  Elec e;
  StepStart st;
  bool alive = false, done = false, pending = false;
  uint32_t r_base = 0, r_left = 0;     /* wave-uniform reserve of claimed item indices */
  int maxSteps = 0;
  uint32_t steps0 = 0;                 /* step count when the lane took the electron */
  uint32_t cur_job = 0xFFFFFFFFu;      /* job the lane's counters belong to */
  uint32_t nsteps = 0, nelec = 0, nesc = 0, nwd = 0, nover = 0;
  const uint32_t n_items = G.n_prim + G.in_n;
  const unsigned long long lanemask_lt = (1ull << lane_id()) - 1ull;
  stage_mott_sigma(S);
  PT_INIT();
  for (;;) {
    PT_BEGIN(t_cl);
    /* claim work for idle lanes (once kClaimMin of them wait, or no lane
     * steps): from the wave's reserve of claimed items, topped up a chunk
     * at a time by one atomic per wave */
    bool need = !alive && !done;
    unsigned long long m = wave_ballot(need);
    if (m && __popcll(m) < kClaimMin && wave_ballot(alive) != 0ull) m = 0ull;
    if (m) {
      const uint32_t cnt = (uint32_t)__popcll(m);
      uint32_t fresh = 0, want = 0;
      if (cnt > r_left) {
        want = max(cnt - r_left, (uint32_t)kClaimChunk);
        const int leader = __ffsll((long long)m) - 1;
        if (lane_id() == leader) fresh = atomicAdd(G.claim, want);
        fresh = __shfl(fresh, leader);
      }
      const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt);
      const uint32_t idx = rank < r_left ? r_base + rank : fresh + (rank - r_left);
      if (cnt > r_left) {
        r_base = fresh + (cnt - r_left);
        r_left = want - (cnt - r_left);
      } else {
        r_base += cnt;
        r_left -= cnt;
      }
      if (need) {
        if (idx < n_items) {
          if (idx < G.n_prim) {
            /* segment holding item idx (binary search over the running ends) */
            int lo = 0, hi = G.n_seg - 1;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (G.seg_end[mid] > idx) hi = mid;
              else lo = mid + 1;
            }
            const uint32_t start = lo > 0 ? G.seg_end[lo - 1] : 0u;
            const uint32_t job = G.seg_job[lo];
            make_primary(S, G.jobs[job], job, G.seg_first[lo] + (idx - start), e);
          } else {
            const uint32_t q = idx - G.n_prim;
            const uint32_t cap = G.in_cap;
            e.px = G.in_f[0 * cap + q];
            e.py = G.in_f[1 * cap + q];
            e.pz = G.in_f[2 * cap + q];
            e.dx = G.in_f[3 * cap + q];
            e.dy = G.in_f[4 * cap + q];
            e.dz = G.in_f[5 * cap + q];
            e.E = G.in_f[6 * cap + q];
            e.region = (int)G.in_u[0 * cap + q];
            e.primary = G.in_u[1 * cap + q];
            e.lin_lo = G.in_u[2 * cap + q];
            e.lin_hi = G.in_u[3 * cap + q];
            e.job = G.in_u[4 * cap + q];
            e.steps = G.in_u[5 * cap + q];
            e.nchild = G.in_u[6 * cap + q];
            e.safe = 0.f;
          }
          if (e.job != cur_job) {
            if (cur_job != 0xFFFFFFFFu) flush_lane_stats(G, cur_job, nsteps, nelec, nesc, nwd, nover);
            cur_job = e.job;
          }
          const JobRec& J = G.jobs[e.job];
          if (J.killed) {
            nover++;
          } else {
            maxSteps = J.max_steps;
            steps0 = e.steps;
            /* a queued electron is located again (the check opening its next step) */
            alive = idx < G.n_prim ? e.region >= 0 : locate(S, G.jobs, e, e.region);
          }
        } else {
          done = true;
        }
      }
    }
    /* the wave leaves only once every lane has run past the launch's items: a
     * claim round can hand out nothing but dead electrons (queued ones of a
     * job the host stopped, or primaries that start outside the chamber),
     * and the items after them must still be claimed */
    if (wave_ballot(alive || !done) == 0ull) break;
    PT_END(PH_CLAIM, t_cl);
    if (wave_ballot(alive) == 0ull) continue;
    /* start a step on every lane without one in progress; a step that may
     * reach a boundary waits for the wave's next findEndOfStep pass, which
     * runs once geo_batch lanes wait or no lane could move without it */
    if (alive && !pending) {
      begin_step(S, e, st);
      pending = true;
    }
    const unsigned long long m_geo = wave_ballot(alive && st.geo);
    const unsigned long long m_free = wave_ballot(alive && !st.geo);
    const bool run_geo = m_free == 0ull || __popcll(m_geo) >= G.geo_batch;
    Spawn sp;
    sp.has = false;
    int res = STEP_CONTINUE;
    const bool stepped = alive && (!st.geo || run_geo);
    if (stepped) {
      PT_BEGIN(t_st);
      res = finish_step<TAB>(S, G.jobs, e, sp, maxSteps, st);
      pending = false;
      nsteps++;
      /* an electron is counted once, at its first step (it may be queued
       * again before taking one) */
      if (e.steps == 1u) nelec++;
      PT_END(PH_STEP, t_st);
    }
    PT_BEGIN(t_ta);
    /* a new secondary starts where its parent's step ended, in the region
     * the parent is in (trackSecondaryElectron, MonteCarloSS.cu:165-174);
     * its state is set aside before any lane takes a new electron below */
    uint32_t clo = 0, chi = 0;
    if (sp.has) {
      const uint64_t lin = child_lineage(((uint64_t)e.lin_hi << 32) | e.lin_lo, e.nchild++);
      clo = (uint32_t)lin;
      chi = (uint32_t)(lin >> 32);
    }
    const float cpx = e.px, cpy = e.py, cpz = e.pz;
    const float csafe = sp.has && e.safe > 0.f ? fmaxf(0.f, e.safe - from_anchor(cpx, cpy, cpz)) : 0.f;
    const int creg = e.region;
    const uint32_t cprim = e.primary, cjob = e.job;
    /* regroup: an electron past its step budget for this launch continues in the next one */
    const bool requeue = stepped && res == STEP_CONTINUE && G.steps_per_launch > 0 &&
                         (int)(e.steps - steps0) >= G.steps_per_launch;
    queue_append(G, requeue, e.px, e.py, e.pz, e.dx, e.dy, e.dz, e.E, e.region, e.primary, e.lin_lo, e.lin_hi, e.job,
                 e.steps, e.nchild, lanemask_lt, nover);
    /* escape records: one atomic per wave */
    const bool esc = stepped && res == STEP_ESCAPED;
    unsigned long long em = wave_ballot(esc);
    if (em) {
      int leader = __ffsll((long long)em) - 1;
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(G.esc_count, (unsigned int)__popcll(em));
      base = __shfl(base, leader);
      if (esc) {
        uint32_t slot = base + (uint32_t)__popcll(em & lanemask_lt);
        if (slot < G.esc_cap) {
          const uint32_t cap = G.esc_cap;
          G.esc_f[0 * cap + slot] = e.px;
          G.esc_f[1 * cap + slot] = e.py;
          G.esc_f[2 * cap + slot] = e.pz;
          G.esc_f[3 * cap + slot] = e.dx;
          G.esc_f[4 * cap + slot] = e.dy;
          G.esc_f[5 * cap + slot] = e.dz;
          G.esc_f[6 * cap + slot] = e.E;
          G.esc_u[0 * cap + slot] = e.primary;
          G.esc_u[1 * cap + slot] = (e.job << 1) | (uint32_t)e.region;
          G.esc_u[2 * cap + slot] = e.lin_lo;
          G.esc_u[3 * cap + slot] = e.lin_hi;
          G.esc_u[4 * cap + slot] = e.steps;
        } else {
          nover++;
        }
      }
    }
    const bool ends = stepped && (res != STEP_CONTINUE || requeue);
    if (ends) {
      PT_END_ELECTRON(e.steps - steps0);
      alive = false;
      if (res == STEP_ESCAPED) nesc++;
      if (res == STEP_WATCHDOG) {
        nwd++;
        const JobRec& J = G.jobs[e.job];
        if (J.wd_log) {
          const unsigned int k = atomicAdd(J.wd_count, 1u);
          if (k < (unsigned int)kWatchdogLog) {
            float* w = J.wd_log + (size_t)k * kWatchdogFields;
            w[0] = e.px; w[1] = e.py; w[2] = e.pz;
            w[3] = e.dx; w[4] = e.dy; w[5] = e.dz;
            w[6] = e.E; w[7] = (float)e.region; w[8] = (float)e.steps; w[9] = (float)e.primary;
          }
        }
      }
    }
    /* new secondaries: the k-th spawning lane hands its secondary to the
     * k-th lane without an electron; the rest go to the queue for the next
     * launch.  An electron's stream depends only on its lineage, so which
     * lane runs it changes nothing. */
    bool queue_spawn = sp.has;
    if (G.handoff) {
      /* a secondary whose parent left the chamber (region -1) goes to the
       * queue, whose containment check at the claim ends it */
      const bool give = sp.has && creg >= 0;
      const unsigned long long mg = wave_ballot(give);
      const unsigned long long mt = wave_ballot(!alive);
      if (mg && mt) {
        const int n = min(__popcll(mg), __popcll(mt));
        const int grank = __popcll(mg & lanemask_lt);
        const int trank = __popcll(mt & lanemask_lt);
        int* giver = lane_map();
        if (give && grank < n) giver[grank] = lane_id();
        __builtin_amdgcn_wave_barrier();
        const bool take = !alive && trank < n;
        const int src = take ? giver[trank] : lane_id();
        const float tpx = __shfl(cpx, src), tpy = __shfl(cpy, src), tpz = __shfl(cpz, src);
        const float tdx = __shfl(sp.dx, src), tdy = __shfl(sp.dy, src), tdz = __shfl(sp.dz, src);
        const float tE = __shfl(sp.E, src), tsafe = __shfl(csafe, src);
        const int treg = __shfl(creg, src), tmax = __shfl(maxSteps, src);
        const uint32_t tprim = __shfl(cprim, src), tlo = __shfl(clo, src), thi = __shfl(chi, src);
        const uint32_t tjob = __shfl(cjob, src);
        if (take) {
          e.px = tpx; e.py = tpy; e.pz = tpz;
          e.dx = tdx; e.dy = tdy; e.dz = tdz;
          e.E = tE;
          e.region = treg;
          e.primary = tprim;
          e.lin_lo = tlo;
          e.lin_hi = thi;
          e.job = tjob;
          e.steps = 0;
          e.nchild = 0;
          e.safe = tsafe;
          set_anchor(tpx, tpy, tpz);
          if (e.job != cur_job) {
            if (cur_job != 0xFFFFFFFFu) flush_lane_stats(G, cur_job, nsteps, nelec, nesc, nwd, nover);
            cur_job = e.job;
          }
          maxSteps = tmax;
          steps0 = 0;
          alive = true;
        }
        if (give && grank < n) queue_spawn = false;
      }
    }
    queue_append(G, queue_spawn, cpx, cpy, cpz, sp.dx, sp.dy, sp.dz, sp.E, creg, cprim, clo, chi, cjob, 0u, 0u,
                 lanemask_lt, nover);
    /* tail compaction: once the launch's own items are all claimed, a wave
     * down to tail_lanes live lanes queues them for the next launch (which
     * packs them into full waves) and leaves.  A step begun but not finished
     * is redone there from the same state and draws. */
    if (G.tail_lanes > 0) {
      const unsigned long long ma = wave_ballot(alive);
      if (ma && __popcll(ma) <= G.tail_lanes && wave_ballot(done) != 0ull) {
        queue_append(G, alive, e.px, e.py, e.pz, e.dx, e.dy, e.dz, e.E, e.region, e.primary, e.lin_lo, e.lin_hi,
                     e.job, e.steps, e.nchild, lanemask_lt, nover);
        alive = false;
        pending = false;
      }
    }
    PT_END(PH_TAIL, t_ta);
  }
  PT_FLUSH();
  /* flush counters: reduced across the wave when every lane's counters
   * belong to one job, per lane otherwise */
  const unsigned long long valid = wave_ballot(cur_job != 0xFFFFFFFFu);
  if (!valid) return;
  const uint32_t j0 = __shfl(cur_job, __ffsll((long long)valid) - 1);
  if (wave_ballot(cur_job != 0xFFFFFFFFu && cur_job != j0) == 0ull) {
    unsigned long long wsteps = nsteps, welec = nelec, wesc = nesc, wwd = nwd, wover = nover;
    for (int off = 32; off > 0; off >>= 1) {
      wsteps += __shfl_down(wsteps, off);
      welec += __shfl_down(welec, off);
      wesc += __shfl_down(wesc, off);
      wwd += __shfl_down(wwd, off);
      wover += __shfl_down(wover, off);
    }
    if (lane_id() == 0) {
      unsigned long long* st = stat_stripe(G.jobs[j0].stats);
      atomicAdd(&st[kStatSteps], wsteps);
      atomicAdd(&st[kStatElectrons], welec);
      if (wesc) atomicAdd(&st[kStatEscapes], wesc);
      if (wwd) atomicAdd(&st[kStatWatchdog], wwd);
      if (wover) atomicAdd(&st[kStatOverflow], wover);
    }
  } else if (cur_job != 0xFFFFFFFFu) {
    flush_lane_stats(G, cur_job, nsteps, nelec, nesc, nwd, nover);
  }
}

/* Job slots whose histograms a tally launch keeps in LDS, with their word
 * offsets in the dynamic shared memory (off[k] .. off[k+1]). */
constexpr int kLdsJobs = 16;
struct LdsJobs {
  int n;
  int slot[kLdsJobs];
  int off[kLdsJobs + 1];
};

/* Tallies of the escape records of one transport launch (grid-stride over
 * the record count the launch left in *count).  The histograms of the jobs
 * in `lj` accumulate in block-private LDS copies that are added to the
 * jobs' histograms at the end; other jobs' records use global atomics. */
__global__ void __launch_bounds__(256) k_tally_escapes(DevScene S, const JobRec* jobs, const float* ef,
                                                       const uint32_t* eu, uint32_t cap, const unsigned int* count,
                                                       LdsJobs lj) {
  extern __shared__ unsigned int lh[];
  const int L = lj.off[lj.n];
  for (int i = threadIdx.x; i < L; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  const uint32_t n = min(*count, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t w = eu[cap + i];
    const int job = (int)(w >> 1);
    unsigned int* own = nullptr;
    for (int k = 0; k < lj.n; ++k)
      if (lj.slot[k] == job) own = lh + lj.off[k];
    tally_escape(S, jobs[job], eu[i], ef[6 * cap + i], ef[i], ef[cap + i], ef[2 * cap + i], ef[3 * cap + i],
                 ef[4 * cap + i], ef[5 * cap + i], (w & 1u) != 0u, own,
                 ((uint64_t)eu[3 * cap + i] << 32) | eu[2 * cap + i], eu[4 * cap + i]);
  }
  __syncthreads();
  for (int k = 0; k < lj.n; ++k) {
    unsigned long long* h = jobs[lj.slot[k]].hist;
    const int base = lj.off[k], len = lj.off[k + 1] - base;
    for (int i = threadIdx.x; i < len; i += blockDim.x)
      if (lh[base + i]) atomicAdd(&h[i], (unsigned long long)lh[base + i]);
  }
}

/* Electrons per job slot in a launch's output queue (block-private LDS
 * counts of up to kMaxJobs slots, then one atomic per slot and block). */
__global__ void __launch_bounds__(256) k_count_jobs(const uint32_t* job_field, const unsigned int* count, uint32_t cap,
                                                    unsigned int* per_job, int n_slots) {
  __shared__ unsigned int lc[kMaxJobs];
  for (int i = threadIdx.x; i < n_slots; i += blockDim.x) lc[i] = 0u;
  __syncthreads();
  const uint32_t n = min(*count, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t j = job_field[i];
    if (j < (uint32_t)n_slots) atomicAdd(&lc[j], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_slots; i += blockDim.x)
    if (lc[i]) atomicAdd(&per_job[i], lc[i]);
}

/* Counting sort of a launch's output queue by (region, energy class), so
 * that the waves of the next launch start on electrons of one material and
 * one energy regime (the cross-section and slowing-down branches switch at
 * 50 eV and 45 eV).  The order inside a bucket is arbitrary: results do not
 * depend on it (each electron carries its own random stream). */
constexpr int kSortKeys = 4096;
constexpr float kSortSplitEV = 50.f;

__device__ __forceinline__ uint32_t sort_key(const uint32_t* in_u, const float* in_f, uint32_t cap, uint32_t i,
                                             int n_keys) {
  const uint32_t k = 2u * in_u[i] + (in_f[6 * (size_t)cap + i] > kSortSplitEV ? 1u : 0u);
  return k < (uint32_t)n_keys ? k : (uint32_t)n_keys - 1u;
}

/* The sort runs in three passes over chunks of the queue, one chunk per
 * block: k_sort_hist writes each chunk's count per key into a key-major
 * table (table[key * n_blocks + block]), k_sort_scan turns the table into
 * exclusive offsets (so that the entries of one key are grouped, chunk by
 * chunk), and k_sort_place writes each chunk's entries of a key to the
 * contiguous range its table entry names.  No global atomics: a chunk's
 * entries of one key land in one run of slots, written by one block. */
constexpr uint32_t kSortChunk = 1024;           /* entries per block at least */
constexpr uint32_t kSortTable = 256u * 1024u;   /* table entries: n_keys x blocks at most */

/* one LDS atomic per distinct key of the wave; returns the lane's rank among
 * the wave's lanes of its key plus the key's base from the LDS counter */
__device__ __forceinline__ uint32_t wave_key_slot(unsigned int* lc, bool valid, uint32_t key) {
  const unsigned long long lanemask_lt = (1ull << lane_id()) - 1ull;
  uint32_t slot = 0;
  unsigned long long left = wave_ballot(valid);
  while (left) {
    const uint32_t k0 = __shfl(key, __ffsll((long long)left) - 1);
    const unsigned long long same = wave_ballot(valid && key == k0);
    const int leader = __ffsll((long long)same) - 1;
    uint32_t b = 0;
    if (lane_id() == leader) b = atomicAdd(&lc[k0], (unsigned int)__popcll(same));
    b = __shfl(b, leader);
    if (valid && key == k0) slot = b + (uint32_t)__popcll(same & lanemask_lt);
    left &= ~same;
  }
  return slot;
}

__global__ void __launch_bounds__(256) k_sort_hist(const float* __restrict__ in_f, const uint32_t* __restrict__ in_u,
                                                   uint32_t cap, uint32_t n, uint32_t chunk, int n_keys,
                                                   unsigned int* __restrict__ table) {
  __shared__ unsigned int lc[kSortKeys];
  for (int i = threadIdx.x; i < n_keys; i += blockDim.x) lc[i] = 0u;
  __syncthreads();
  const uint32_t first = blockIdx.x * chunk;
  const uint32_t last = min(n, first + chunk);
  /* uniform trip count per wave so that the ballots see every lane */
  for (uint32_t base = first + (threadIdx.x & ~63u); base < last; base += blockDim.x) {
    const uint32_t i = base + lane_id();
    const bool valid = i < last;
    const uint32_t key = valid ? sort_key(in_u, in_f, cap, i, n_keys) : 0u;
    (void)wave_key_slot(lc, valid, key);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n_keys; k += blockDim.x) table[(size_t)k * gridDim.x + blockIdx.x] = lc[k];
}

/* Exclusive prefix sum of table[0..total) in place, in tiles of
 * kScanTile entries (256 threads x 4): k_sort_tile_sums sums each tile,
 * k_sort_scan_tiles scans the tile sums (at most kSortTable / kScanTile =
 * 256 of them) in every block and then its own tile. */
constexpr uint32_t kScanTile = 1024;

/* inclusive scan of one value per thread of a 256-thread block */
__device__ __forceinline__ unsigned int block_scan_256(unsigned int v, unsigned int* l) {
  const int t = threadIdx.x;
  l[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const unsigned int a = t >= off ? l[t - off] : 0u;
    __syncthreads();
    l[t] += a;
    __syncthreads();
  }
  const unsigned int r = l[t];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_sort_tile_sums(const unsigned int* __restrict__ table, uint32_t total,
                                                        unsigned int* __restrict__ sums) {
  __shared__ unsigned int l[256];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4u;
  unsigned int s = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4u; ++k)
    if (base + k < total) s += table[base + k];
  s = block_scan_256(s, l);
  if (threadIdx.x == 255) sums[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_sort_scan_tiles(unsigned int* __restrict__ table, uint32_t total,
                                                         const unsigned int* __restrict__ sums) {
  __shared__ unsigned int l[256];
  /* this tile's offset: the sum of the tile sums before it */
  const unsigned int mine = threadIdx.x < blockIdx.x ? sums[threadIdx.x] : 0u;
  const unsigned int before = block_scan_256(mine, l);
  __shared__ unsigned int tile_off;
  if (threadIdx.x == 255) tile_off = before;
  __syncthreads();
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4u;
  unsigned int v[4];
  unsigned int s = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4u; ++k) {
    v[k] = base + k < total ? table[base + k] : 0u;
    s += v[k];
  }
  unsigned int run = tile_off + block_scan_256(s, l) - s;
#pragma unroll
  for (uint32_t k = 0; k < 4u; ++k)
    if (base + k < total) {
      table[base + k] = run;
      run += v[k];
    }
}

__global__ void __launch_bounds__(256) k_sort_place(const float* __restrict__ in_f, const uint32_t* __restrict__ in_u,
                                                    float* __restrict__ out_f, uint32_t* __restrict__ out_u,
                                                    uint32_t cap, uint32_t n, uint32_t chunk, int n_keys,
                                                    const unsigned int* __restrict__ table) {
  __shared__ unsigned int lc[kSortKeys];
  for (int k = threadIdx.x; k < n_keys; k += blockDim.x) lc[k] = table[(size_t)k * gridDim.x + blockIdx.x];
  __syncthreads();
  const uint32_t first = blockIdx.x * chunk;
  const uint32_t last = min(n, first + chunk);
  for (uint32_t base = first + (threadIdx.x & ~63u); base < last; base += blockDim.x) {
    const uint32_t i = base + lane_id();
    const bool valid = i < last;
    const uint32_t key = valid ? sort_key(in_u, in_f, cap, i, n_keys) : 0u;
    const uint32_t slot = wave_key_slot(lc, valid, key);
    if (valid) {
#pragma unroll
      for (int f = 0; f < kQF; ++f) out_f[(size_t)f * cap + slot] = in_f[(size_t)f * cap + i];
#pragma unroll
      for (int f = 0; f < kQU; ++f) out_u[(size_t)f * cap + slot] = in_u[(size_t)f * cap + i];
    }
  }
}

/* Per-primary tallies of one job -> per-pixel sums (exact u64 counts,
 * second moments and fixed-point energy).  px holds back, fwd, se, se^2,
 * bse^2, energy ([6][n_pixels]).  Primaries of one pixel are consecutive, so a wave whose
 * 64 primaries share a pixel reduces in registers and issues one atomic per
 * sum; a wave that straddles pixels falls back to per-lane atomics. */
__global__ void k_reduce_primaries(const JobRec* jobs, int slot, uint32_t n, unsigned long long* px) {
  const JobRec& J = jobs[slot];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const uint32_t iv = valid ? i : n - 1u;
  const uint32_t c = valid ? J.prim_counts[i] : 0u;
  unsigned long long nse = c & 0xFFFFu, nbse = c >> 16;
  unsigned long long vf = valid ? J.prim_fwd[i] : 0u;
  unsigned long long ve = valid ? J.prim_energy[i] : 0ull;
  const uint32_t pix = iv / (uint32_t)J.ppp;
  unsigned long long vb = nse + nbse, vs = nse, vss = nse * nse, vbb = nbse * nbse;
  const uint32_t np = (uint32_t)J.n_pixels;
  const uint32_t pix0 = __shfl(pix, 0);
  if (wave_ballot(pix != pix0) == 0ull) {
    for (int off = 32; off > 0; off >>= 1) {
      vb += __shfl_xor(vb, off);
      vf += __shfl_xor(vf, off);
      vs += __shfl_xor(vs, off);
      vss += __shfl_xor(vss, off);
      vbb += __shfl_xor(vbb, off);
      ve += __shfl_xor(ve, off);
    }
    if (lane_id() == 0) {
      if (vb) atomicAdd(&px[pix0], vb);
      if (vf) atomicAdd(&px[np + pix0], vf);
      if (vs) atomicAdd(&px[2 * np + pix0], vs);
      if (vss) atomicAdd(&px[3 * np + pix0], vss);
      if (vbb) atomicAdd(&px[4 * np + pix0], vbb);
      if (ve) atomicAdd(&px[5 * np + pix0], ve);
    }
  } else {
    if (vb) atomicAdd(&px[pix], vb);
    if (vf) atomicAdd(&px[np + pix], vf);
    if (vs) atomicAdd(&px[2 * np + pix], vs);
    if (vss) atomicAdd(&px[3 * np + pix], vss);
    if (vbb) atomicAdd(&px[4 * np + pix], vbb);
    if (ve) atomicAdd(&px[5 * np + pix], ve);
  }
}

/* Zeroes n u64 words (clears device buffers without hipMemsetAsync). */
__global__ void k_zero_u64(unsigned long long* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

/* Per-pixel sums of squares and fixed-point energies (u64, pixel tallies 3..5
 * of k_reduce_primaries) as the doubles of monsel_tallies: se_sq, bse_sq and
 * back_energy (eV), each [np]. */
__global__ void k_pixels_to_double(const unsigned long long* px, int np, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * np) return;
  const double v = (double)px[3 * (size_t)np + i];
  out[i] = i >= 2 * np ? v / kEnergyScale : v;
}

/* ---- probe kernels ---------------------------------------------------- */

__global__ void k_probe_rates(DevScene S, int msm, int n, const float* E, float* out) {
  stage_mott_sigma(S);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MONSEL_CS DMSM& m = S.msms[msm];
  float tot = 0.f;
  for (int k = 0; k < m.n_mech; ++k) {
    float r = mech_rate(S, S.mechs[m.mech0 + k], E[i]);
    out[i * (m.n_mech + 1) + k] = r;
    tot += r;
  }
  out[i * (m.n_mech + 1) + m.n_mech] = tot;
}

__global__ void k_probe_elastic(DevScene S, int mech, int element, int n, const float* E, const float* u,
                                float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MONSEL_CS DMech& m = S.mechs[mech];
  const MONSEL_CS DElem& el = S.elems[m.elem0 + element];
  float EeV = E[i], r1 = u[2 * i], r2 = u[2 * i + 1];
  float ca;
  if (el.kind == MONSEL_MECH_BROWNING_ELASTIC || (el.kind == MONSEL_MECH_NISTMOTT_ELASTIC && EeV < el.extrap))
    ca = browning_cos(el, EeV, r1, r2);
  else if (el.kind == MONSEL_MECH_RUTHERFORD_ELASTIC || EeV >= kMaxNISTMott)
    ca = rutherford_cos(el, EeV, r1);
  else
    ca = mott_cos(S, el.mott, EeV, r1);
  out[i] = acosf(fminf(1.f, fmaxf(-1.f, ca)));
}

/* One mechanism scatter with injected uniforms u[6i..6i+5] (slots 2..7 of
 * the step's uniforms); out[11i..]: dir (3), E, secondary flag, secondary
 * E, secondary dir (3), trapped flag, 0. */
__global__ void k_probe_scatter(DevScene S, int mech, int n, const float* prevE, const float* E, const float* dir,
                                const float* u, float* out) {
  stage_mott_sigma(S);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MONSEL_CS DMech& m = S.mechs[mech];
  Elec e;
  e.px = e.py = e.pz = 0.f;
  e.dx = dir[3 * i];
  e.dy = dir[3 * i + 1];
  e.dz = dir[3 * i + 2];
  e.E = E[i];
  float uu[kStepUniforms];
  uu[0] = uu[1] = 0.5f;
#pragma unroll
  for (int k = 0; k < 6; ++k) uu[2 + k] = u[6 * i + k];
  bool trapped = false;
  const float rate = mech_rate(S, m, prevE[i]);
  Spawn sp = mech_scatter(S, m, e, prevE[i], rate, trapped, uu);
  float* o = out + 11 * (size_t)i;
  o[0] = e.dx;
  o[1] = e.dy;
  o[2] = e.dz;
  o[3] = e.E;
  o[4] = sp.has ? 1.f : 0.f;
  o[5] = sp.has ? sp.E : 0.f;
  o[6] = sp.has ? sp.dx : 0.f;
  o[7] = sp.has ? sp.dy : 0.f;
  o[8] = sp.has ? sp.dz : 0.f;
  o[9] = trapped ? 1.f : 0.f;
  o[10] = 0.f;
}

/* One barrier crossing from region `from` into `to` with the outward barrier
 * normal given; out[6i..]: dir (3), E, region after, nudge sign. */
__global__ void k_probe_barrier(DevScene S, int from, int to, int n, const float* dir, const float* nb, const float* E,
                                const float* u, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Elec e;
  e.px = e.py = e.pz = 0.f;
  e.dx = dir[3 * i];
  e.dy = dir[3 * i + 1];
  e.dz = dir[3 * i + 2];
  e.E = E[i];
  e.region = from;
  int sign = barrier_core(S, S.msms[S.regions[from].msm], S.msms[S.regions[to].msm], e, to, nb[3 * i], nb[3 * i + 1],
                          nb[3 * i + 2], u[i]);
  float* o = out + 6 * (size_t)i;
  o[0] = e.dx;
  o[1] = e.dy;
  o[2] = e.dz;
  o[3] = e.E;
  o[4] = (float)e.region;
  o[5] = (float)sign;
}

__global__ void k_probe_csd(DevScene S, int msm, int n, const float* len, const float* E, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = csd_compute(S.msms[msm].csd, len[i], E[i]);
}

__global__ void k_probe_step_end(DevScene S, int region, int n, const float* p0, const float* p1, float* out_u,
                                 int* out_next, float* out_n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ax = p0[3 * i], ay = p0[3 * i + 1], az = p0[3 * i + 2];
  float bx = p1[3 * i] - ax, by = p1[3 * i + 1] - ay, bz = p1[3 * i + 2] - az;
  float L = sqrtf(bx * bx + by * by + bz * bz);
  float inv = 1.f / L;
  StepEnd se = find_end_of_step(S, region, ax, ay, az, bx * inv, by * inv, bz * inv, L);
  out_u[i] = se.boundary ? se.t / L : kInf;
  out_next[i] = se.next;
  out_n[3 * i] = se.boundary ? se.winNx : 0.f;
  out_n[3 * i + 1] = se.boundary ? se.winNy : 0.f;
  out_n[3 * i + 2] = se.boundary ? se.winNz : 0.f;
}

__global__ void k_probe_safety(DevScene S, int region, int n, const float* p, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = region_safety(S, region, p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__global__ void k_probe_locate(DevScene S, int n, const float* p, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = containing_sub_region(S, 0, p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__global__ void k_probe_philox(int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out, float* out_u) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o[4];
  philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i], key[2 * i + 1], o[0], o[1],
                o[2], o[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    out[4 * i + k] = o[k];
    if (out_u) out_u[4 * i + k] = u01(o[k]);
  }
}

__global__ void k_probe_hist_bin(int n, const double* val, double lo, double hi, int nbins, int compat, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = hist_bin(val[i], lo, hi, nbins, compat != 0);
}

__global__ void k_probe_primaries(DevScene S, const JobRec* J, int n, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Elec e;
  make_primary(S, *J, 0u, (uint32_t)i, e);
  out[4 * i] = e.px;
  out[4 * i + 1] = e.py;
  out[4 * i + 2] = e.pz;
  out[4 * i + 3] = (float)e.region;
}

__global__ void k_probe_nu(DevScene S, int table, int order, int n, const float* x, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int dim = S.nu[table].dim;
  double xd[4] = {0., 0., 0., 0.};
  for (int k = 0; k < dim; ++k) xd[k] = x[(size_t)i * dim + k];
  out[i] = (float)nu_interp(S, table, order, xd, (S.compat & MONSEL_COMPAT_D3_ABSCISSAE) != 0);
}

/* ===================================================================== */
/* host: scene compilation                                                */
/* ===================================================================== */

namespace {

const double kE = 1.60217653e-19;
const double kAvogadro = 6.0221415e23;
const double kBohr = 5.291772083e-11;
const double kKB = 1.3806505e-23;
const double kHbar = 1.05457266e-34;
const double kMe = 9.1093826e-31;

/* device buffer */
struct DBuf {
  void* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t bytes) {
    if (bytes <= n) return;
    if (p) HIP_TRY(hipFree(p));
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw MonselError(MONSEL_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    n = bytes;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

/* pinned host buffer: copies to and from it run on the stream without the
 * host waiting (pageable copies are staged and wait) */
struct HBuf {
  void* p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() { reset(); }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  /* grows without keeping the contents; only when no copy uses it */
  void ensure(size_t bytes) {
    if (bytes <= n) return;
    reset();
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      throw MonselError(MONSEL_E_NOMEM, "hipHostMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    n = bytes;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

struct HostScene {
  std::vector<DRegion> regions;
  std::vector<int> children;
  std::vector<DShape> shapes;
  std::vector<DPrim> prims;
  std::vector<float4> planes;
  std::vector<DMSM> msms;
  std::vector<DMech> mechs;
  std::vector<DElem> elems;
  std::vector<float> spwem, x1;
  std::vector<DNUTable> nu;
  std::vector<float> nudata;
  std::vector<double> nudata64;
  uint32_t compat = 0;
  /* bound on tracked secondaries per eV of primary energy (queue reservation) */
  double secondaries_per_eV = 0.;
  float chamber_r = 1e8f;
  float chamber_c[3] = {0, 0, 0};
};

struct SceneDev {
  DBuf blob;
  DevScene dev{};
  size_t bytes = 0;
  int step_prims = 0; /* the most primitives one findEndOfStep call tests */
  HostScene host;     /* kept to assemble batch scenes */
};

/* The most primitives one findEndOfStep call tests in the scene: a region's
 * own shape and its children's, the chamber's sphere aside (an electron
 * meets it once, on its way out). */
int max_step_prims(const HostScene& H) {
  auto prims = [&](int region) {
    const DShape& sh = H.shapes[H.regions[region].shape];
    return sh.is_sphere ? 0 : sh.n_prims;
  };
  int most = 0;
  for (int r = 0; r < (int)H.regions.size(); ++r) {
    int n = prims(r);
    for (int c = 0; c < H.regions[r].n_children; ++c) n += prims(H.children[H.regions[r].first_child + c]);
    most = std::max(most, n);
  }
  return most;
}

/* Lanes a wave gathers for findEndOfStep when the tuning leaves it open (0):
 * gathering pays where a call tests several primitives (line and image
 * scenes, 3 or more: 11 to 14 % more primaries/s with 8) and costs where it
 * tests one or two (flat and step-edge yield jobs: 2 to 4 % slower).  A
 * launch of few items (a small job alone) leaves most of the device idle and
 * lasts as long as its longest cascade, which waiting only lengthens. */
int geometry_batch_for(int tuned, int stepPrims, uint64_t items) {
  if (tuned > 0) return tuned;
  return stepPrims >= 3 && items >= kTailMinItems ? 8 : 1;
}

/* New primaries per launch when the tuning leaves it open (0): 262,144, so
 * that a large job runs as a chain of launches whose queues are sorted in
 * between (the 256 x 256 x 300 image: 3.5 % more primaries/s than in one
 * launch of all its primaries; the 512-job batch alike within 1 %). */
constexpr uint64_t kDefaultPrimariesPerLaunch = 262144;
uint64_t primaries_per_launch_for(int tuned, uint64_t cap) {
  return tuned > 0 ? (uint64_t)tuned : std::min<uint64_t>(cap, kDefaultPrimariesPerLaunch);
}

/* JoyLuoNieminenCSD::init (JoyLuoNieminenCSD.cu:42-82) in SI, then eV/nm. */
DCSD make_csd(const monsel_material_desc& m, double breakE) {
  DCSD c;
  memset(&c, 0, sizeof(c));
  c.active = 1;
  c.nce = m.n_elements;
  if (c.nce == 0) return c;
  double bh = -m.energy_cb_bottom;
  double bhplus1eV = bh + kE;
  if (breakE < bhplus1eV) breakE = bhplus1eV;
  double gamma = 0.;
  double recipJ[MONSEL_MAX_ELEMENTS], beta[MONSEL_MAX_ELEMENTS], coef[MONSEL_MAX_ELEMENTS];
  for (int i = 0; i < c.nce; ++i) {
    double z = m.Z[i];
    double J = m.Z[i] < 13 ? kE * (11.5 * z) : kE * (9.76 * z + 58.8 * pow(z, -0.19));
    recipJ[i] = 1.166 / J;
    beta[i] = 1. - recipJ[i] * bhplus1eV;
    coef[i] = 2.01507E-28 * m.density * m.weight_fraction[i] * m.Z[i] / m.atomic_weight[i];
    gamma += coef[i] * log(recipJ[i] * breakE + beta[i]);
  }
  gamma /= pow(breakE, 3.5);
  for (int i = 0; i < c.nce; ++i) {
    c.recipJ[i] = (float)(recipJ[i] * kE);
    c.beta[i] = (float)beta[i];
    c.coef[i] = (float)(coef[i] * 1e-9 / (kE * kE));
  }
  c.breakE = (float)(breakE / kE);
  c.gamma = (float)(gamma * 1e-9 * pow(kE, 1.5));
  c.minE = (float)(bh / kE);
  return c;
}

/* collect the convex leaves of a CSG tree (deduplicated by shape index) */
void collect_leaves(const monsel_scene_desc* d, int s, std::vector<int>& leaves, int depth) {
  if (depth > 64) throw MonselError(MONSEL_E_INVALID, "shape tree too deep or cyclic");
  const monsel_shape_desc& sh = d->shapes[s];
  if (sh.kind == MONSEL_SHAPE_INTERSECTION || sh.kind == MONSEL_SHAPE_UNION) {
    if (sh.a < 0 || sh.b < 0 || sh.a >= d->n_shapes || sh.b >= d->n_shapes)
      throw MonselError(MONSEL_E_INVALID, "CSG child index out of range");
    collect_leaves(d, sh.a, leaves, depth + 1);
    collect_leaves(d, sh.b, leaves, depth + 1);
    return;
  }
  if (std::find(leaves.begin(), leaves.end(), s) == leaves.end()) leaves.push_back(s);
}

bool eval_tree(const monsel_scene_desc* d, int s, const std::vector<int>& leaves, uint32_t bits) {
  const monsel_shape_desc& sh = d->shapes[s];
  if (sh.kind == MONSEL_SHAPE_INTERSECTION)
    return eval_tree(d, sh.a, leaves, bits) && eval_tree(d, sh.b, leaves, bits);
  if (sh.kind == MONSEL_SHAPE_UNION) return eval_tree(d, sh.a, leaves, bits) || eval_tree(d, sh.b, leaves, bits);
  int k = (int)(std::find(leaves.begin(), leaves.end(), s) - leaves.begin());
  return (bits >> k) & 1u;
}

HostScene compile_scene(const monsel_scene_desc* d) {
  if (!d || d->abi_version != MONSEL_ABI_VERSION) throw MonselError(MONSEL_E_INVALID, "bad scene descriptor or ABI version");
  if (d->n_regions < 1 || !d->regions) throw MonselError(MONSEL_E_INVALID, "scene has no regions");
  if (d->regions[0].parent != -1) throw MonselError(MONSEL_E_INVALID, "region 0 must be the chamber");
  HostScene H;
  H.compat = d->compat_flags;
  double R = d->chamber_radius > 0 ? d->chamber_radius : 0.1;
  H.chamber_r = (float)(R * 1e9);
  /* Mott tables */
  std::map<int, int> mottIdx;
  for (int i = 0; i < d->n_mott; ++i) {
    mottIdx[d->mott[i].Z] = i;
    for (int j = 0; j < kMottNE; ++j) H.spwem.push_back((float)d->mott[i].spwem[j]);
    for (int j = 0; j < kMottNE * kMottNQ; ++j) H.x1.push_back((float)d->mott[i].x1[j]);
  }
  /* NU tables (raw units) and their domains / ranges */
  std::vector<std::array<double, 2>> dom0(d->n_nu_tables), range(d->n_nu_tables);
  for (int i = 0; i < d->n_nu_tables; ++i) {
    const monsel_nu_table& t = d->nu_tables[i];
    if (t.dim < 1 || t.dim > 4) throw MonselError(MONSEL_E_INVALID, "NU table dimension must be 1..4");
    DNUTable dt;
    memset(&dt, 0, sizeof(dt));
    dt.dim = t.dim;
    size_t tot = 1;
    for (int k = 0; k < t.dim; ++k) {
      if (t.n[k] < 2) throw MonselError(MONSEL_E_INVALID, "NU table axis too short");
      dt.n[k] = t.n[k];
      dt.xoff[k] = (int)H.nudata.size();
      for (int j = 0; j < t.n[k]; ++j) {
        H.nudata.push_back((float)t.x[k][j]);
        H.nudata64.push_back(t.x[k][j]);
      }
      tot *= (size_t)t.n[k];
    }
    dt.voff = (int)H.nudata.size();
    double r0 = INFINITY, r1 = -INFINITY;
    for (size_t j = 0; j < tot; ++j) {
      double v = t.values[j];
      H.nudata.push_back((float)v);
      H.nudata64.push_back(v);
      /* NUTableInterpolation.cu:169-172 */
      if (v < r0) r0 = v;
      else if (v > r1) r1 = v;
    }
    double a = t.x[0][0], b = t.x[0][t.n[0] - 1];
    dom0[i] = {std::min(a, b), std::max(a, b)};
    range[i] = {r0, r1};
    H.nu.push_back(dt);
  }
  /* materials -> MSMs */
  for (int i = 0; i < d->n_msm; ++i) {
    const monsel_msm_desc& md = d->msms[i];
    if (md.material < 0 || md.material >= d->n_materials) throw MonselError(MONSEL_E_INVALID, "MSM material index");
    if (md.n_mech < 0 || md.n_mech > MONSEL_MAX_MECH) throw MonselError(MONSEL_E_INVALID, "MSM mechanism count");
    const monsel_material_desc& mat = d->materials[md.material];
    if (mat.n_elements < 0 || mat.n_elements > MONSEL_MAX_ELEMENTS) throw MonselError(MONSEL_E_INVALID, "element count");
    DMSM m;
    memset(&m, 0, sizeof(m));
    m.is_null = md.is_null;
    m.n_mech = md.is_null ? 0 : md.n_mech;
    m.mech0 = (int)H.mechs.size();
    m.barrier_kind = md.barrier_kind;
    m.elastic_k = -1;
    m.minE = (float)(md.min_e_for_tracking / kE);
    m.cb = (float)(mat.energy_cb_bottom / kE);
    m.lambda = (float)(md.barrier_lambda * 1e9);
    m.lambdaFactor = (float)((M_PI * md.barrier_lambda * sqrt(kMe / 2.)) / kHbar * sqrt(kE));
    if (md.csd_kind == MONSEL_CSD_JOY_LUO_NIEMINEN) m.csd = make_csd(mat, md.csd_break_e);
    double eFermi = -mat.energy_cb_bottom - mat.workfunction;
    for (int k = 0; k < m.n_mech; ++k) {
      const monsel_mech_desc& e = md.mech[k];
      DMech dm;
      memset(&dm, 0, sizeof(dm));
      dm.kind = e.kind;
      dm.method = e.method;
      for (int q = 0; q < 4; ++q) dm.table[q] = e.table[q];
      switch (e.kind) {
        case MONSEL_MECH_NISTMOTT_ELASTIC:
        case MONSEL_MECH_BROWNING_ELASTIC:
        case MONSEL_MECH_RUTHERFORD_ELASTIC: {
          if (m.elastic_k < 0) m.elastic_k = k;
          dm.elem0 = (int)H.elems.size();
          dm.nelem = mat.n_elements;
          for (int j = 0; j < mat.n_elements; ++j) {
            DElem el;
            memset(&el, 0, sizeof(el));
            int Z = mat.Z[j];
            el.kind = e.kind;
            el.Z = Z;
            el.method = e.method < 1 ? 1 : e.method;
            el.z = (float)Z;
            el.zp17 = (float)pow((double)Z, 1.7);
            el.zp2 = (float)pow((double)Z, 2.0);
            el.zp3 = (float)pow((double)Z, 3.0);
            el.rz13 = (float)pow((double)Z, 1.0 / 3.0);
            el.rz23 = (float)pow((double)Z, 2.0 / 3.0);
            /* SelectableElasticSM.cu:85 scalefactor, :72 densityNa; atoms/nm^3 */
            el.natoms = (float)((1000. * mat.weight_fraction[j]) / mat.atomic_weight[j] * mat.density * kAvogadro * 1e-27);
            el.mott = -1;
            if (e.kind == MONSEL_MECH_NISTMOTT_ELASTIC) {
              auto it = mottIdx.find(Z);
              if (it == mottIdx.end()) throw MonselError(MONSEL_E_INVALID, "missing NIST Mott table for Z=" + std::to_string(Z));
              el.mott = it->second;
              /* NISTMottRS.cu:76-91 in FP64 */
              double extrap = (el.method == 1 ? 50.0 : 100.0);
              const double* sp = d->mott[it->second].spwem;
              double DL50 = log(50.0 * kE), PARAM = (log(20000.0 * kE) - DL50) / 60.0;
              double red = (log(extrap * kE) - DL50) / PARAM;
              int i0 = (int)red - 1;
              i0 = i0 < 0 ? 0 : (i0 > kMottNE - 4 ? kMottNE - 4 : i0);
              double t = red - i0;
              double l0 = -(t - 1) * (t - 2) * (t - 3) / 6., l1 = t * (t - 2) * (t - 3) / 2.;
              double l2 = -t * (t - 1) * (t - 3) / 2., l3 = t * (t - 1) * (t - 2) / 6.;
              double sig = kBohr * kBohr * (l0 * sp[i0] + l1 * sp[i0 + 1] + l2 * sp[i0 + 2] + l3 * sp[i0 + 3]);
              double e_keV = extrap * 1e-3;
              double re = sqrt(e_keV);
              double zp17 = pow((double)Z, 1.7), zp2 = pow((double)Z, 2.0);
              double brown = 3.0e-22 * zp17 / (e_keV + 0.005 * zp17 * re + 0.0007 * zp2 / re);
              el.extrap = (float)extrap;
              el.mottAtMin = (float)(sig * 1e18);
              el.sfBrowning = (float)(sig / brown);
            }
            H.elems.push_back(el);
          }
          break;
        }
        case MONSEL_MECH_FITTED_INELASTIC: {
          /* FittedInelSM.cu:20-25: one SE per E_SEgen of energy lost, and an
           * SE (born at E_SEgen + E_F, the rate threshold) emits none; 1.5x
           * covers the Poisson spread of a cascade */
          H.secondaries_per_eV = std::max(H.secondaries_per_eV, 1.5 * kE / e.p[0]);
          double seE = e.p[0] + eFermi;
          dm.p[0] = (float)(seE / kE);
          dm.p[1] = (float)(seE / kE);
          dm.p[2] = (float)(kE / e.p[0]);
          if ((int)e.p[2] == MONSEL_CSD_JOY_LUO_NIEMINEN) dm.csd = make_csd(mat, e.p[1]);
          break;
        }
        case MONSEL_MECH_POLARON_TRAP: {
          double cutoff = -log(10. / e.p[0]) / e.p[1];
          dm.p[0] = (float)(e.p[0] * 1e-9);
          dm.p[1] = (float)(e.p[1] * kE);
          dm.p[2] = (float)(cutoff / kE);
          break;
        }
        case MONSEL_MECH_PHONON: {
          double occ = 0.5 * (1. + (1. / (exp(e.p[1] / (kKB * e.p[2])) - 1.)));
          double epsRatio = (e.p[3] - e.p[4]) / e.p[3] / e.p[4];
          dm.p[0] = (float)(e.p[1] / kE);
          dm.p[1] = (float)((e.p[0] * occ * epsRatio) / kBohr * 1e-9);
          break;
        }
        case MONSEL_MECH_TABULATED_INELASTIC: {
          /* SEs may cascade: one per 8 eV of primary energy at most */
          H.secondaries_per_eV = std::max(H.secondaries_per_eV, 1.0 / 8.0);
          int method = (e.method != 2 && e.method != 3) ? 1 : e.method;
          dm.method = method;
          /* table dimensions and the cubic (order 3) windows the lookups use */
          const int wantDim[4] = {1, 2, 3, 2};
          for (int q = 0; q < (method >= 2 ? 4 : 3); ++q) {
            if (e.table[q] < 0 || e.table[q] >= d->n_nu_tables)
              throw MonselError(MONSEL_E_INVALID, "tabulated mechanism table index");
            const monsel_nu_table& t = d->nu_tables[e.table[q]];
            if (t.dim != wantDim[q]) throw MonselError(MONSEL_E_INVALID, "tabulated mechanism table dimension");
            for (int k = 0; k < t.dim; ++k)
              if (t.n[k] < (q == 0 ? 2 : 4)) throw MonselError(MONSEL_E_INVALID, "table axis shorter than the interpolation order");
          }
          double gap = mat.bandgap;
          dm.p[TP_OFFSET] = (float)(e.p[0] / kE);
          dm.p[TP_MINEGEN] = (float)(e.p[1] / kE);
          dm.p[TP_RATEMULT] = (float)e.p[2];
          dm.p[TP_E0DISP] = (float)e.p[3];
          dm.p[TP_CB] = (float)(mat.energy_cb_bottom / kE);
          dm.p[TP_BANDGAP] = (float)(gap / kE);
          dm.p[TP_GAP] = (float)(gap / kE);
          dm.p[TP_OFFSETFERMI] = (float)((eFermi + e.p[0]) / kE);
          dm.p[TP_BEREF] = (float)((gap > 0. ? -gap : eFermi) / kE);
          /* TabulatedInelasticSM.cu:491-497: domains in J, kept in J */
          double lo = dom0[e.table[1]][0], hi = dom0[e.table[1]][1];
          lo = std::max(lo, dom0[e.table[2]][0]);
          hi = std::min(hi, dom0[e.table[2]][1]);
          dm.p[TP_EIDOM0] = (float)lo;
          dm.p[TP_EIDOM1] = (float)hi;
          dm.p[TP_IIMFPDOM0] = (float)dom0[e.table[0]][0];
          dm.p[TP_IIMFPDOM1] = (float)dom0[e.table[0]][1];
          if (method >= 2) {
            dm.p[TP_SE0RANGE0] = (float)(range[e.table[3]][0] / kE);
            dm.p[TP_SE0RANGE1] = (float)(range[e.table[3]][1] / kE);
          }
          if (mat.n_core > kMaxCore) throw MonselError(MONSEL_E_UNSUPPORTED, "too many core levels");
          dm.p[TP_NCORE] = (float)mat.n_core;
          for (int c = 0; c < mat.n_core; ++c) dm.p[TP_CORE0 + c] = (float)mat.core_energy[c];
          /* the same slots in J (FP64) for the scatter; TabulatedInelasticSM.cu:461-498 */
          dm.q[TP_OFFSET] = e.p[0];
          dm.q[TP_MINEGEN] = e.p[1];
          dm.q[TP_RATEMULT] = e.p[2];
          dm.q[TP_E0DISP] = e.p[3];
          dm.q[TP_CB] = mat.energy_cb_bottom;
          dm.q[TP_BANDGAP] = gap;
          dm.q[TP_GAP] = gap;
          dm.q[TP_OFFSETFERMI] = eFermi + e.p[0];
          dm.q[TP_BEREF] = gap > 0. ? -gap : eFermi;
          dm.q[TP_EIDOM0] = lo;
          dm.q[TP_EIDOM1] = hi;
          dm.q[TP_IIMFPDOM0] = dom0[e.table[0]][0];
          dm.q[TP_IIMFPDOM1] = dom0[e.table[0]][1];
          if (method >= 2) {
            dm.q[TP_SE0RANGE0] = range[e.table[3]][0];
            dm.q[TP_SE0RANGE1] = range[e.table[3]][1];
          }
          dm.q[TP_NCORE] = mat.n_core;
          for (int c = 0; c < mat.n_core; ++c) dm.q[TP_CORE0 + c] = mat.core_energy[c];
          break;
        }
        default:
          throw MonselError(MONSEL_E_UNSUPPORTED, "unknown mechanism kind " + std::to_string(e.kind));
      }
      H.mechs.push_back(dm);
    }
    H.msms.push_back(m);
  }
  /* region shapes -> primitives + truth tables */
  std::map<int, int> shapeIdx;
  for (int r = 0; r < d->n_regions; ++r) {
    int s = d->regions[r].shape;
    if (s < 0 || s >= d->n_shapes) throw MonselError(MONSEL_E_INVALID, "region shape index");
    if (shapeIdx.count(s)) continue;
    std::vector<int> leaves;
    collect_leaves(d, s, leaves, 0);
    if ((int)leaves.size() > kMaxPrims)
      throw MonselError(MONSEL_E_UNSUPPORTED, "region shape has more than 8 convex parts");
    DShape ds;
    memset(&ds, 0, sizeof(ds));
    ds.first_prim = (int)H.prims.size();
    ds.n_prims = (int)leaves.size();
    for (int leaf : leaves) {
      const monsel_shape_desc& sh = d->shapes[leaf];
      DPrim p;
      memset(&p, 0, sizeof(p));
      if (sh.kind == MONSEL_SHAPE_PLANES) {
        if (sh.n_planes < 1 || sh.first_plane < 0 || sh.first_plane + sh.n_planes > d->n_planes)
          throw MonselError(MONSEL_E_INVALID, "plane range");
        if (sh.n_planes > kMaxPlanes)
          throw MonselError(MONSEL_E_UNSUPPORTED, "multi-plane shape with more than " + std::to_string(kMaxPlanes) +
                                                      " planes");
        p.kind = PRIM_PLANES;
        p.first_plane = (int)H.planes.size();
        p.n_planes = sh.n_planes;
        double maxn = 0.;
        for (int k = 0; k < sh.n_planes; ++k) {
          const monsel_plane_desc& pl = d->planes[sh.first_plane + k];
          double n[3] = {pl.normal[0], pl.normal[1], pl.normal[2]};
          double b = (n[0] * pl.point[0] + n[1] * pl.point[1] + n[2] * pl.point[2]) * 1e9;
          H.planes.push_back(make_float4((float)n[0], (float)n[1], (float)n[2], (float)b));
          if (k < kInlinePlanes) p.pl[k] = H.planes.back();
          maxn = std::max(maxn, sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
        }
        /* min over the planes of 1/|n|, rounded down (distance bounds) */
        p.sdscale = maxn > 0. ? std::nextafter((float)(1.0 / maxn), 0.f) : 0.f;
      } else if (sh.kind == MONSEL_SHAPE_CYLINDER) {
        p.kind = PRIM_CYLINDER;
        p.sdscale = 1.f;
        double ax[3], L2 = 0;
        for (int k = 0; k < 3; ++k) {
          ax[k] = (sh.end1[k] - sh.end0[k]) * 1e9;
          L2 += ax[k] * ax[k];
        }
        double L = sqrt(L2);
        if (!(L > 0)) throw MonselError(MONSEL_E_INVALID, "degenerate cylinder");
        for (int k = 0; k < 3; ++k) {
          p.e0[k] = (float)(sh.end0[k] * 1e9);
          p.ax[k] = (float)(ax[k] / L);
          p.axv[k] = (float)ax[k];
        }
        p.len = (float)L;
        p.len2 = (float)L2;
        p.r2 = (float)(sh.radius * 1e9 * sh.radius * 1e9);
      } else if (sh.kind == MONSEL_SHAPE_SPHERE) {
        p.kind = PRIM_SPHERE;
        p.sdscale = 1.f;
        for (int k = 0; k < 3; ++k) p.e0[k] = (float)(sh.center[k] * 1e9);
        p.r2 = (float)(sh.radius * 1e9 * sh.radius * 1e9);
        ds.is_sphere = 1;
        if (r == 0) {
          H.chamber_r = (float)(sh.radius * 1e9);
          for (int k = 0; k < 3; ++k) H.chamber_c[k] = p.e0[k];
        }
      } else {
        throw MonselError(MONSEL_E_UNSUPPORTED, "unknown shape kind");
      }
      H.prims.push_back(p);
    }
    for (uint32_t bits = 0; bits < (1u << ds.n_prims); ++bits)
      if (eval_tree(d, s, leaves, bits)) ds.truth[bits >> 5] |= 1u << (bits & 31u);
    shapeIdx[s] = (int)H.shapes.size();
    H.shapes.push_back(ds);
  }
  /* regions */
  std::vector<std::vector<int>> kids(d->n_regions);
  for (int r = 0; r < d->n_regions; ++r) {
    int par = d->regions[r].parent;
    if (r > 0 && (par < 0 || par >= r))
      throw MonselError(MONSEL_E_INVALID, "regions must be listed parents-first");
    if (d->regions[r].msm < 0 || d->regions[r].msm >= d->n_msm) throw MonselError(MONSEL_E_INVALID, "region MSM index");
    if (par >= 0) kids[par].push_back(r);
  }
  for (int r = 0; r < d->n_regions; ++r) {
    DRegion dr;
    memset(&dr, 0, sizeof(dr));
    dr.parent = d->regions[r].parent;
    dr.shape = shapeIdx[d->regions[r].shape];
    dr.msm = d->regions[r].msm;
    dr.first_child = (int)H.children.size();
    dr.n_children = (int)kids[r].size();
    dr.depth = dr.parent < 0 ? 0 : H.regions[dr.parent].depth + 1;
    dr.sh = H.shapes[dr.shape];
    for (int c : kids[r]) H.children.push_back(c);
    H.regions.push_back(dr);
  }
  if (H.children.empty()) H.children.push_back(0);
  return H;
}

template <class T>
size_t append(std::vector<char>& blob, const std::vector<T>& v) {
  size_t off = (blob.size() + 255) & ~(size_t)255;
  blob.resize(off + v.size() * sizeof(T));
  if (!v.empty()) memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

size_t host_scene_bytes(const HostScene& H) {
  std::vector<char> b;
  size_t total = 0;
  auto add = [&](size_t bytes) { total = ((total + 255) & ~(size_t)255) + bytes; };
  add(H.regions.size() * sizeof(DRegion));
  add(H.children.size() * sizeof(int));
  add(H.shapes.size() * sizeof(DShape));
  add(H.prims.size() * sizeof(DPrim));
  add(H.planes.size() * sizeof(float4));
  add(H.msms.size() * sizeof(DMSM));
  add(H.mechs.size() * sizeof(DMech));
  add(H.elems.size() * sizeof(DElem));
  add(H.spwem.size() * sizeof(float));
  add(H.x1.size() * sizeof(float));
  add(H.nu.size() * sizeof(DNUTable));
  add(H.nudata.size() * sizeof(float));
  add(H.nudata64.size() * sizeof(double));
  return total;
}

void upload_scene(const HostScene& H, SceneDev& sd) {
  std::vector<char> blob;
  size_t o_reg = append(blob, H.regions), o_ch = append(blob, H.children), o_sh = append(blob, H.shapes);
  size_t o_pr = append(blob, H.prims), o_pl = append(blob, H.planes), o_msm = append(blob, H.msms);
  size_t o_me = append(blob, H.mechs), o_el = append(blob, H.elems), o_sp = append(blob, H.spwem);
  size_t o_x1 = append(blob, H.x1), o_nu = append(blob, H.nu), o_nd = append(blob, H.nudata);
  size_t o_nd64 = append(blob, H.nudata64);
  sd.blob.ensure(blob.size() + 256);
  HIP_TRY(hipMemcpy(sd.blob.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
  char* b = sd.blob.as<char>();
  DevScene& S = sd.dev;
  S.n_regions = (int)H.regions.size();
  S.n_shapes = (int)H.shapes.size();
  S.n_prims = (int)H.prims.size();
  S.n_planes = (int)H.planes.size();
  S.n_msm = (int)H.msms.size();
  S.n_mech = (int)H.mechs.size();
  S.n_elem = (int)H.elems.size();
  S.n_mott = (int)(H.spwem.size() / kMottNE);
  S.n_nu = (int)H.nu.size();
  S.has_tabulated = 0;
  for (const DMech& m : H.mechs)
    if (m.kind == MONSEL_MECH_TABULATED_INELASTIC) S.has_tabulated = 1;
  S.compat = H.compat;
  S.chamber_r = H.chamber_r;
  for (int k = 0; k < 3; ++k) S.chamber_c[k] = H.chamber_c[k];
  S.regions = (decltype(S.regions))(b + o_reg);
  S.children = (decltype(S.children))(b + o_ch);
  S.shapes = (decltype(S.shapes))(b + o_sh);
  S.prims = (decltype(S.prims))(b + o_pr);
  S.planes = (decltype(S.planes))(b + o_pl);
  S.msms = (decltype(S.msms))(b + o_msm);
  S.mechs = (decltype(S.mechs))(b + o_me);
  S.elems = (decltype(S.elems))(b + o_el);
  S.mott_spwem = (decltype(S.mott_spwem))(b + o_sp);
  S.mott_x1 = (decltype(S.mott_x1))(b + o_x1);
  S.nu = (decltype(S.nu))(b + o_nu);
  S.nu_data = (decltype(S.nu_data))(b + o_nd);
  S.nu_data64 = (decltype(S.nu_data64))(b + o_nd64);
  sd.bytes = blob.size();
  sd.step_prims = max_step_prims(H);
}

/* One device scene holding several scenes side by side: every index is
 * shifted by the arrays that precede it; roots[k] is the chamber region of
 * part k.  Electrons never leave their own part, so a batch of jobs on
 * different scenes runs in the same launches. */
HostScene merge_scenes(const std::vector<const HostScene*>& parts, std::vector<int>& roots) {
  HostScene M;
  M.chamber_r = 0.f;
  roots.clear();
  for (const HostScene* P : parts) {
    const int o_reg = (int)M.regions.size(), o_ch = (int)M.children.size(), o_sh = (int)M.shapes.size();
    const int o_pr = (int)M.prims.size(), o_pl = (int)M.planes.size(), o_msm = (int)M.msms.size();
    const int o_me = (int)M.mechs.size(), o_el = (int)M.elems.size(), o_mott = (int)(M.spwem.size() / kMottNE);
    const int o_nu = (int)M.nu.size(), o_nd = (int)M.nudata.size();
    roots.push_back(o_reg);
    for (DRegion r : P->regions) {
      if (r.parent >= 0) r.parent += o_reg;
      r.shape += o_sh;
      r.sh.first_prim += o_pr;
      r.msm += o_msm;
      r.first_child += o_ch;
      M.regions.push_back(r);
    }
    for (int c : P->children) M.children.push_back(c + o_reg);
    for (DShape sh : P->shapes) {
      sh.first_prim += o_pr;
      M.shapes.push_back(sh);
    }
    for (DPrim pr : P->prims) {
      if (pr.kind == PRIM_PLANES) pr.first_plane += o_pl;
      M.prims.push_back(pr);
    }
    M.planes.insert(M.planes.end(), P->planes.begin(), P->planes.end());
    for (DMSM m : P->msms) {
      m.mech0 += o_me;
      M.msms.push_back(m);
    }
    for (DMech m : P->mechs) {
      m.elem0 += o_el;
      if (m.kind == MONSEL_MECH_TABULATED_INELASTIC)
        for (int q = 0; q < 4; ++q)
          if (m.table[q] >= 0) m.table[q] += o_nu;
      M.mechs.push_back(m);
    }
    for (DElem el : P->elems) {
      if (el.mott >= 0) el.mott += o_mott;
      M.elems.push_back(el);
    }
    M.spwem.insert(M.spwem.end(), P->spwem.begin(), P->spwem.end());
    M.x1.insert(M.x1.end(), P->x1.begin(), P->x1.end());
    for (DNUTable t : P->nu) {
      for (int k = 0; k < t.dim; ++k) t.xoff[k] += o_nd;
      t.voff += o_nd;
      M.nu.push_back(t);
    }
    M.nudata.insert(M.nudata.end(), P->nudata.begin(), P->nudata.end());
    M.nudata64.insert(M.nudata64.end(), P->nudata64.begin(), P->nudata64.end());
    M.compat = P->compat;
    M.chamber_r = std::max(M.chamber_r, P->chamber_r);
  }
  return M;
}

/* ===================================================================== */
/* allocation-free jobs: one static device array per GPU                  */
/* ===================================================================== */

/* Every device buffer of monsel_run_static beyond the caller's inputs and
 * outputs (job record, launch counters, per-primary and per-pixel tallies,
 * histograms, pixel centres, both generation queues and the escape records)
 * is carved out of this array.  Each GPU the library runs on holds its own
 * copy from the code object's load until the process ends.  Host code never
 * passes `scratch` itself (on the host it is a stand-in, not a device
 * address): it uses the address hipGetSymbolAddress gives per device. */
constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
constexpr int kMaxStaticDevices = 64;
static std::mutex g_static_mu[kMaxStaticDevices]; /* one call at a time per GPU */
static void* g_static_ptr[kMaxStaticDevices];     /* device address of `scratch` per GPU */
static hipError_t g_static_status[kMaxStaticDevices]; /* of the lookup, when it failed */
static std::mutex g_static_init_mu;

/* Resolves the device address of `scratch` on `device`, which must be the
 * current device (monsel_create calls this once per context).  A failed
 * lookup is kept and reported by monsel_run_static; the rest of the context
 * does not need the array. */
static void static_scratch_resolve(int device) {
  if (device < 0 || device >= kMaxStaticDevices) return; /* monsel_run_static reports it */
  std::lock_guard<std::mutex> lk(g_static_init_mu);
  if (g_static_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_static_status[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    (void)hipGetLastError(); /* the failure belongs to monsel_run_static, not to later calls */
    return;
  }
  g_static_ptr[device] = p;
}

/* ===================================================================== */
/* host: jobs                                                             */
/* ===================================================================== */

struct Job {
  monsel_job_id id = 0;
  monsel_scene_id scene = 0;
  monsel_job_desc desc{};
  std::vector<double> centers;
  /* results */
  std::vector<uint64_t> back, fwd, se;
  std::vector<double> se_sq, bse_sq, back_energy;
  std::vector<uint64_t> hback, hfwd, helev, hazim;
  monsel_job_stats stats{};
  int status = MONSEL_E_NOTREADY;
  std::string error;
  bool done = false;
  std::chrono::steady_clock::time_point t_submit, t_start;
  /* set by monsel_job_cancel (any thread), acted on by the worker */
  std::atomic<bool> cancel{false};
  /* its scene, held from submission to the job's end */
  std::shared_ptr<SceneDev> sd;
  /* engine state (worker thread only) */
  bool cancelling = false;    /* stopped: drains its queued electrons, then ends */
  int slot = -1;
  uint64_t total = 0, launched = 0;
  uint32_t K = 0;             /* queue entries reserved per launched primary */
  uint64_t reserved = 0;      /* queue entries reserved for its electrons in flight */
  uint32_t outstanding = 0;   /* its electrons in the current queue */
  int generations = 0;
  /* its tallies and pixel centres in its slot's buffer (monsel_ctx_s::slotbuf):
   * [stats | histograms | per-pixel sums | per-primary counts] cleared at
   * admission, then the centres */
  size_t o_hist = 0, o_px = 0, o_pc = 0, o_pf = 0, o_pe = 0, o_cen = 0, zero_bytes = 0;
  char* base = nullptr;
  DBuf wd, det;
  std::vector<monsel_detected> detected; /* the kept records, after completion */
  uint64_t n_detected = 0;
};

}  // namespace

struct monsel_ctx_s {
  int device = 0;
  int n_cu = 0;
  size_t hbm = 0;
  std::string name;
  monsel_tuning tuning{};
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  std::map<monsel_scene_id, std::shared_ptr<SceneDev>> scenes;
  std::map<monsel_job_id, std::shared_ptr<Job>> jobs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  /* engine (worker thread only) */
  DBuf q[2], esc, counts, segs, jobtab, sort_table;
  uint32_t qcap = 0;          /* entries per queue buffer */
  int cur = 0;                /* queue holding the next launch's input */
  uint32_t in_count = 0;      /* electrons in q[cur] */
  /* cumulative over the context's transport launches (read under mu) */
  uint64_t n_launches = 0, n_steps = 0;
  double transport_ms = 0.;
  unsigned int* host_count = nullptr; /* pinned: out, esc, per-job counts */
  uint32_t* host_segs = nullptr;      /* pinned: segment staging */
  HBuf pin_count, pin_segs;
  std::vector<std::shared_ptr<Job>> active;
  std::vector<int> free_slots;
  /* host copy of the device job table (pinned: a job's record is copied to
   * the device on the stream while the host goes on) */
  HBuf pin_jobrec;
  JobRec* jobrec = nullptr;
  /* pinned staging of each slot's pixel centres, and of the read-back of
   * the jobs that complete after one launch */
  HBuf pin_cen[kMaxJobs];
  HBuf pin_readback;
  /* device memory of each job slot, kept from job to job (grown as needed,
   * given back when a job left it above kSlotKeepBytes): admitting a job
   * then allocates nothing */
  DBuf slotbuf[kMaxJobs];
  SceneDev batch;             /* the scenes of the active jobs, side by side */
  std::vector<std::shared_ptr<SceneDev>> batch_parts;
  std::vector<int> batch_roots;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
  int blocks_per_cu = 0;
};

namespace {

uint32_t secondaries_per_primary(double E0eV) {
  double k = E0eV / 8.0;
  if (k < 16) k = 16;
  if (k > 4096) k = 4096;
  return (uint32_t)k;
}

uint64_t splitmix64_host(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* lineage of a job's primaries: the seed enters every electron's stream here */
uint64_t job_lineage(uint64_t seed) { return splitmix64_host(seed ^ 0x5EED5EED5EED5EEDull); }

constexpr size_t kQueueEntryBytes = (size_t)(kQF + kQU) * 4;
constexpr size_t kEscEntryBytes = (size_t)(kEF + kEU) * 4;
/* largest queue capacity whose field offsets (up to 7 x capacity) fit u32 */
constexpr uint32_t kMaxQueueCapacity = 0xFFFFFFFFu / 8u;

/* Queue entries for the device budget: both generation buffers plus the
 * escape records (tuning.queue_capacity overrides). */
uint32_t queue_capacity(monsel_ctx_s* C) {
  if (C->tuning.queue_capacity > 0) return (uint32_t)C->tuning.queue_capacity;
  size_t budget = std::min<size_t>((size_t)6 << 30, C->hbm / 8);
  uint64_t cap = budget / (2 * kQueueEntryBytes + kEscEntryBytes);
  return (uint32_t)std::min<uint64_t>(cap, kMaxQueueCapacity);
}

/* Rebuilds the batch scene from the scenes of the active jobs. Only between
 * launches (the stream is idle). */
void rebuild_batch_scene(monsel_ctx_s* C) {
  std::vector<const HostScene*> parts;
  for (auto& p : C->batch_parts) parts.push_back(&p->host);
  HostScene M = merge_scenes(parts, C->batch_roots);
  SceneDev fresh;
  upload_scene(M, fresh);
  C->batch.blob = std::move(fresh.blob);
  C->batch.dev = fresh.dev;
  C->batch.bytes = fresh.bytes;
  C->batch.step_prims = fresh.step_prims;
}

void finish_job(monsel_ctx_s* C, const std::shared_ptr<Job>& job, int st, const std::string& msg) {
  {
    std::lock_guard<std::mutex> lk(C->mu);
    job->status = st;
    job->error = msg;
    job->stats.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - job->t_submit).count();
    job->done = true;
  }
  C->cv_done.notify_all();
}

/* Admits a job: slot, device buffers, its scene in the batch scene. */
void setup_job(monsel_ctx_s* C, Job& job) {
  const monsel_job_desc& jd = job.desc;
  const int np = jd.n_pixels;
  const int nbins = jd.n_energy_bins > 0 ? jd.n_energy_bins : 1;
  const double E0eV = jd.beam_energy / kE;
  job.total = (uint64_t)jd.n_pixels * (uint64_t)jd.primaries_per_pixel;
  /* queue entries reserved per primary: itself plus its secondaries, bounded
   * by the scene's inelastic mechanisms */
  job.K = (uint32_t)std::min<double>(4096., 16. + E0eV * job.sd->host.secondaries_per_eV);
  job.launched = job.reserved = 0;
  job.outstanding = 0;
  job.generations = 0;
  job.t_start = std::chrono::steady_clock::now();
  int part = -1;
  for (size_t k = 0; k < C->batch_parts.size(); ++k)
    if (C->batch_parts[k] == job.sd) part = (int)k;
  if (part < 0) {
    C->batch_parts.push_back(job.sd);
    try {
      rebuild_batch_scene(C);
    } catch (...) {
      /* the batch scene on the device is still the one without it */
      C->batch_parts.pop_back();
      throw;
    }
    part = (int)C->batch_parts.size() - 1;
  }
  job.slot = C->free_slots.back();
  C->free_slots.pop_back();
  const size_t n = (size_t)std::max<uint64_t>(job.total, 1);
  /* one buffer per slot: stats, histograms and per-pixel sums first (read
   * back together), then the per-primary counts, all cleared by one fill,
   * then the pixel centres */
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  take(8 * 8 * kStatStripes);
  job.o_hist = take((size_t)hist_len(nbins) * 8);
  job.o_px = take((size_t)np * 6 * 8);
  job.o_pc = take(n * 4);
  job.o_pf = take(n * 4);
  job.o_pe = take(n * 8);
  job.zero_bytes = off;
  job.o_cen = take((size_t)np * 3 * 8);
  C->slotbuf[job.slot].ensure(off);
  job.base = C->slotbuf[job.slot].as<char>();
  HIP_TRY(hipMemsetAsync(job.base, 0, job.zero_bytes, C->stream));
  /* the slot's staging is free: its previous job's copies ended before that
   * job completed */
  HBuf& pc = C->pin_cen[job.slot];
  pc.ensure((size_t)np * 3 * 8);
  memcpy(pc.p, job.centers.data(), (size_t)np * 3 * 8);
  HIP_TRY(hipMemcpyAsync(job.base + job.o_cen, pc.p, (size_t)np * 3 * 8, hipMemcpyHostToDevice, C->stream));
  const HostScene& H = job.sd->host;
  JobRec& R = C->jobrec[job.slot];
  memset(&R, 0, sizeof(R));
  R.lin0 = job_lineage(jd.seed);
  R.E0 = (float)E0eV;
  R.width = (float)(jd.beam_width * 1e9);
  R.centers = (const double*)(job.base + job.o_cen);
  R.n_pixels = np;
  R.ppp = jd.primaries_per_pixel;
  R.prim_base = (uint32_t)jd.primary_offset;
  R.root = C->batch_roots[part];
  R.max_steps = jd.max_steps > 0 ? jd.max_steps : 1000000;
  R.nbins = nbins;
  R.maxSEbin = (int)(jd.se_cutoff_ev / (E0eV / nbins));
  R.killed = 0;
  R.E0eV = E0eV;
  for (int k = 0; k < 3; ++k) R.chamber_c[k] = H.chamber_c[k];
  R.chamber_r = H.chamber_r;
  R.prim_counts = (uint32_t*)(job.base + job.o_pc);
  R.prim_fwd = (uint32_t*)(job.base + job.o_pf);
  R.prim_energy = (unsigned long long*)(job.base + job.o_pe);
  R.hist = (unsigned long long*)(job.base + job.o_hist);
  R.stats = (unsigned long long*)job.base;
  /* MONSEL_DEBUG_WATCHDOG=1: keep the state of the first watchdog drops and
   * print it when the job completes */
  if (getenv("MONSEL_DEBUG_WATCHDOG")) {
    job.wd.ensure(kWatchdogLog * kWatchdogFields * 4 + 16);
    HIP_TRY(hipMemsetAsync(job.wd.p, 0, kWatchdogLog * kWatchdogFields * 4 + 16, C->stream));
    R.wd_log = job.wd.as<float>();
    R.wd_count = (unsigned int*)(job.wd.as<float>() + kWatchdogLog * kWatchdogFields);
  }
  if (jd.log_detected > 0) {
    const size_t bytes = 16 + (size_t)jd.log_detected * sizeof(monsel_detected);
    job.det.ensure(bytes);
    HIP_TRY(hipMemsetAsync(job.det.p, 0, 16, C->stream));
    R.det_count = job.det.as<unsigned long long>();
    R.det_log = (monsel_detected*)(job.det.as<char>() + 16);
    R.det_max = (unsigned long long)jd.log_detected;
  }
  /* the copies' sources (pinned: the slot's centre staging, its record in
   * C->jobrec) are not written again until the job ends, and the next
   * launch orders after them on the stream: no synchronisation here */
  HIP_TRY(hipMemcpyAsync(C->jobtab.as<JobRec>() + job.slot, &R, sizeof(JobRec), hipMemcpyHostToDevice, C->stream));
}

/* bytes of a job's read-back: stats, histograms and per-pixel sums */
size_t readback_bytes(const Job& job) { return job.o_px + (size_t)job.desc.n_pixels * 6 * 8; }

/* Per-pixel reduction of a job whose electrons are all done, and the
 * read-back of its tallies into `dst` (pinned, readback_bytes long), on the
 * stream: the jobs that complete after one launch are read back together,
 * with one synchronisation. */
void complete_enqueue(monsel_ctx_s* C, Job& job, char* dst) {
  if (job.total > 0) {
    const int rb = 256;
    g_hip_call = "launch k_reduce_primaries";
    hipLaunchKernelGGL(k_reduce_primaries, dim3((unsigned)((job.total + rb - 1) / rb)), dim3(rb), 0, C->stream,
                       (const JobRec*)C->jobtab.as<JobRec>(), job.slot, (uint32_t)job.total,
                       (unsigned long long*)(job.base + job.o_px));
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(dst, job.base, readback_bytes(job), hipMemcpyDeviceToHost, C->stream));
}

/* A completed job's results from its read-back (after the stream has
 * passed complete_enqueue). */
void complete_collect(monsel_ctx_s* C, Job& job, const char* src) {
  const monsel_job_desc& jd = job.desc;
  const int np = jd.n_pixels;
  const int nbins = C->jobrec[job.slot].nbins;
  const unsigned long long* all = (const unsigned long long*)src;
  std::vector<unsigned long long> st(8, 0ull);
  const unsigned long long* sts = all;
  const std::vector<unsigned long long> h(all + job.o_hist / 8, all + job.o_hist / 8 + hist_len(nbins));
  const std::vector<unsigned long long> px(all + job.o_px / 8, all + job.o_px / 8 + (size_t)np * 6);
  for (int k = 0; k < 8 * kStatStripes; ++k) st[k % 8] += sts[k];
  job.back.assign(px.begin(), px.begin() + np);
  job.fwd.assign(px.begin() + np, px.begin() + 2 * np);
  job.se.assign(px.begin() + 2 * np, px.begin() + 3 * np);
  job.se_sq.resize(np);
  job.bse_sq.resize(np);
  job.back_energy.resize(np);
  for (int p = 0; p < np; ++p) {
    job.se_sq[p] = (double)px[3 * (size_t)np + p];
    job.bse_sq[p] = (double)px[4 * (size_t)np + p];
    job.back_energy[p] = (double)px[5 * (size_t)np + p] / kEnergyScale;
  }
  const int nb2 = nbins + 2;
  job.hback.assign(h.begin(), h.begin() + nb2);
  job.hfwd.assign(h.begin() + nb2, h.begin() + 2 * nb2);
  job.helev.assign(h.begin() + 2 * nb2, h.begin() + 2 * nb2 + 182);
  job.hazim.assign(h.begin() + 2 * nb2 + 182, h.begin() + 2 * nb2 + 182 + 362);
  if (job.det.p) {
    unsigned long long nd = 0;
    HIP_TRY(hipMemcpy(&nd, job.det.p, 8, hipMemcpyDeviceToHost));
    job.n_detected = nd;
    const size_t kept = (size_t)std::min<unsigned long long>(nd, (unsigned long long)jd.log_detected);
    job.detected.resize(kept);
    if (kept)
      HIP_TRY(hipMemcpy(job.detected.data(), job.det.as<char>() + 16, kept * sizeof(monsel_detected),
                        hipMemcpyDeviceToHost));
  }
  job.stats.primaries = job.total;
  job.stats.steps = st[kStatSteps];
  {
    std::lock_guard<std::mutex> lk(C->mu);
    C->n_steps += st[kStatSteps];
  }
  job.stats.electrons = st[kStatElectrons];
  job.stats.escapes = st[kStatEscapes];
  job.stats.watchdog_drops = st[kStatWatchdog];
  job.stats.queue_overflow = st[kStatOverflow];
  job.stats.generations = job.generations;
  job.stats.device_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - job.t_start).count();
  if (job.wd.p) {
    std::vector<float> w(kWatchdogLog * kWatchdogFields + 4);
    HIP_TRY(hipMemcpy(w.data(), job.wd.p, w.size() * 4, hipMemcpyDeviceToHost));
    unsigned int n = 0;
    memcpy(&n, w.data() + kWatchdogLog * kWatchdogFields, 4);
    fprintf(stderr, "[monsel] job %llu: %u watchdog drops\n", (unsigned long long)job.id, n);
    for (unsigned int k = 0; k < std::min<unsigned int>(n, kWatchdogLog); ++k) {
      const float* v = w.data() + k * kWatchdogFields;
      fprintf(stderr, "[monsel]   p=(%.6g %.6g %.6g) nm d=(%.6g %.6g %.6g) E=%.6g eV region=%d steps=%.0f primary=%.0f\n",
              v[0], v[1], v[2], v[3], v[4], v[5], v[6], (int)v[7], v[8], v[9]);
    }
  }
}

/* a slot keeps its buffer for the next job up to this size */
constexpr size_t kSlotKeepBytes = (size_t)64 << 20;

void release_job(monsel_ctx_s* C, Job& job) {
  if (job.slot >= 0) {
    if (C->slotbuf[job.slot].n > kSlotKeepBytes) C->slotbuf[job.slot].reset();
    /* its pinned centre staging likewise (the job's copies have ended: it
     * completed, or its setup failed and the stream was synchronised) */
    if (C->pin_cen[job.slot].n > kSlotKeepBytes) C->pin_cen[job.slot].reset();
    C->free_slots.push_back(job.slot);
  }
  job.slot = -1;
  job.base = nullptr;
  job.wd.reset();
  job.det.reset();
  job.sd.reset();
}

/* Marks a job's device record so that the transport kernel drops its
 * electrons when it claims them (counted as lost in its statistics). */
void stop_job(monsel_ctx_s* C, Job& job) {
  if (C->jobrec[job.slot].killed) return;
  C->jobrec[job.slot].killed = 1;
  HIP_TRY(hipMemcpyAsync(&C->jobtab.as<JobRec>()[job.slot].killed, &C->jobrec[job.slot].killed, sizeof(int),
                         hipMemcpyHostToDevice, C->stream));
  HIP_TRY(hipStreamSynchronize(C->stream));
}

/* One launch of the transport kernel over new primaries of the active jobs
 * and every electron queued by the previous launch, then its tallies. */
void run_generation(monsel_ctx_s* C) {
  const uint32_t cap = C->qcap;
  const int maxGen = C->tuning.max_generations > 0 ? C->tuning.max_generations : 100000;
  const uint64_t perLaunch = primaries_per_launch_for(C->tuning.primaries_per_launch, cap);
  /* jobs with electrons in flight before this launch */
  std::vector<Job*> inflight;
  uint64_t reservedTotal = 0;
  for (auto& j : C->active) reservedTotal += j->reserved;
  /* new primaries within the queue reservation: oldest job first, or (with
   * admission_order 1) the job with the fewest primaries left first */
  std::vector<Job*> order;
  for (auto& jp : C->active) order.push_back(jp.get());
  if (C->tuning.admission_order == 1)
    std::stable_sort(order.begin(), order.end(),
                     [](const Job* a, const Job* b) { return a->total - a->launched < b->total - b->launched; });
  int nseg = 0;
  uint64_t nprim = 0;
  for (Job* jp : order) {
    Job& j = *jp;
    if (j.launched >= j.total || nseg >= kMaxJobs) continue;
    const uint64_t avail = cap > reservedTotal ? cap - reservedTotal : 0;
    uint64_t n = std::min<uint64_t>(j.total - j.launched, avail / j.K);
    n = std::min<uint64_t>(n, perLaunch - nprim);
    if (n == 0) continue;
    C->host_segs[nseg] = (uint32_t)(nprim + n);
    C->host_segs[kMaxJobs + nseg] = (uint32_t)j.slot;
    C->host_segs[2 * kMaxJobs + nseg] = (uint32_t)j.launched;
    ++nseg;
    nprim += n;
    j.launched += n;
    j.reserved += n * j.K;
    reservedTotal += n * j.K;
  }
  for (auto& jp : C->active)
    if (jp->outstanding > 0 || jp->reserved > 0) inflight.push_back(jp.get());
  if (nprim == 0 && C->in_count == 0) {
    /* nothing to run: jobs without primaries complete; a job whose primaries
     * cannot be admitted into an empty queue is too large for it */
    for (auto& jp : C->active)
      if (jp->launched < jp->total) throw MonselError(MONSEL_E_NOMEM, "queue capacity below one primary's reservation");
    return;
  }
  unsigned int* counts = C->counts.as<unsigned int>();
  /* counts[0] claim, [1] queued electrons, [2] escapes, [4..] per-job queue counts */
  HIP_TRY(hipMemsetAsync(counts, 0, (4 + kMaxJobs) * sizeof(unsigned int), C->stream));
  if (nseg > 0)
    HIP_TRY(hipMemcpyAsync(C->segs.p, C->host_segs, 3 * kMaxJobs * 4, hipMemcpyHostToDevice, C->stream));
  const int cur = C->cur;
  GenDev G;
  memset(&G, 0, sizeof(G));
  G.jobs = C->jobtab.as<JobRec>();
  G.seg_end = C->segs.as<uint32_t>();
  G.seg_job = C->segs.as<uint32_t>() + kMaxJobs;
  G.seg_first = C->segs.as<uint32_t>() + 2 * kMaxJobs;
  G.n_seg = nseg;
  G.n_prim = (uint32_t)nprim;
  G.in_f = C->q[cur].as<float>();
  G.in_u = (const uint32_t*)(C->q[cur].as<float>() + (size_t)kQF * cap);
  G.in_cap = cap;
  G.in_n = C->in_count;
  G.claim = counts;
  G.out_f = C->q[cur ^ 1].as<float>();
  G.out_u = (uint32_t*)(C->q[cur ^ 1].as<float>() + (size_t)kQF * cap);
  G.out_cap = cap;
  G.out_count = counts + 1;
  G.esc_f = C->esc.as<float>();
  G.esc_u = (uint32_t*)(C->esc.as<float>() + (size_t)kEF * cap);
  G.esc_cap = cap;
  G.esc_count = counts + 2;
  G.steps_per_launch = C->tuning.steps_per_launch;
  G.geo_batch = geometry_batch_for(C->tuning.geometry_batch, C->batch.step_prims, nprim + C->in_count);
  G.handoff = C->tuning.spawn_handoff >= 0 ? 1 : 0;
  G.tail_lanes = C->tuning.tail_requeue < 0 || nprim + C->in_count < kTailMinItems
                     ? 0
                     : (C->tuning.tail_requeue > 0 ? C->tuning.tail_requeue : kDefaultTailLanes);
  const int block = C->tuning.block_size > 0 ? C->tuning.block_size : 256;
  const int grid = C->n_cu * (C->blocks_per_cu > 0 ? C->blocks_per_cu : 4);
  /* Event markers time the kernels; MONSEL_NO_EVENT_TIMING=1 leaves them out
   * (hardware-counter profiling sessions reject the marker packets). */
  const bool timing = !getenv("MONSEL_NO_EVENT_TIMING");
  if (timing) HIP_TRY(hipEventRecord(C->evk0, C->stream));
  g_hip_call = "launch k_transport";
  /* scenes without a tabulated mechanism run the kernel built without it */
  if (C->batch.dev.has_tabulated)
    hipLaunchKernelGGL(k_transport<true>, dim3(grid), dim3(block), 0, C->stream, C->batch.dev, G);
  else
    hipLaunchKernelGGL(k_transport<false>, dim3(grid), dim3(block), 0, C->stream, C->batch.dev, G);
  HIP_TRY(hipGetLastError());
  if (timing) HIP_TRY(hipEventRecord(C->evk1, C->stream));
  /* escape tallies: the histograms go through LDS when one job is in flight */
  LdsJobs lj;
  memset(&lj, 0, sizeof(lj));
  for (Job* j : inflight) {
    const int len = hist_len(C->jobrec[j->slot].nbins);
    if (lj.n == kLdsJobs || (size_t)(lj.off[lj.n] + len) * 4 > 64 * 1024) break;
    lj.slot[lj.n] = j->slot;
    lj.off[lj.n + 1] = lj.off[lj.n] + len;
    ++lj.n;
  }
  const size_t lds_bytes = (size_t)lj.off[lj.n] * 4;
  g_hip_call = "launch k_tally_escapes";
  hipLaunchKernelGGL(k_tally_escapes, dim3(C->n_cu * 4), dim3(256), lds_bytes, C->stream, C->batch.dev,
                     (const JobRec*)G.jobs, (const float*)G.esc_f, (const uint32_t*)G.esc_u, cap,
                     (const unsigned int*)(counts + 2), lj);
  HIP_TRY(hipGetLastError());
  int n_slots = 0;
  for (Job* j : inflight) n_slots = std::max(n_slots, j->slot + 1);
  if (inflight.size() > 1) {
    g_hip_call = "launch k_count_jobs";
    hipLaunchKernelGGL(k_count_jobs, dim3(C->n_cu * 2), dim3(256), 0, C->stream,
                       (const uint32_t*)(G.out_u + 4 * (size_t)cap), (const unsigned int*)(counts + 1), cap,
                       counts + 4, n_slots);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(C->host_count, counts, (4 + std::max(n_slots, 1)) * sizeof(unsigned int),
                         hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipStreamSynchronize(C->stream));
  /* every item of the launch was claimed (waves claim in chunks, so the
   * counter ends at or beyond the item count) */
  if ((uint64_t)C->host_count[0] < nprim + (uint64_t)G.in_n)
    throw MonselError(MONSEL_E_DEVICE, "transport launch ended with unclaimed items (" +
                                           std::to_string(C->host_count[0]) + " of " +
                                           std::to_string(nprim + G.in_n) + ")");
  const uint32_t outc = C->host_count[1];
  C->in_count = outc > cap ? cap : outc;
  C->cur = cur ^ 1;
  /* regroup the queue by (region, energy class) for the next launch */
  const int n_keys = 2 * C->batch.dev.n_regions;
  if (C->tuning.sort_queue >= 0 && C->in_count >= 65536 && n_keys <= kSortKeys) {
    const uint32_t n = C->in_count;
    /* chunks of kSortChunk entries, larger when the key x chunk table would
     * exceed kSortTable entries (then at least kSortTable / n_keys chunks) */
    uint32_t chunk = kSortChunk;
    uint64_t nb = ((uint64_t)n + chunk - 1) / chunk;
    if (nb * (uint64_t)n_keys > kSortTable) {
      const uint64_t maxb = kSortTable / (uint64_t)n_keys;
      chunk = (uint32_t)((((uint64_t)n + maxb - 1) / maxb + 255) / 256 * 256);
      nb = ((uint64_t)n + chunk - 1) / chunk;
    }
    unsigned int* table = C->sort_table.as<unsigned int>();
    const float* sf = C->q[C->cur].as<float>();
    const uint32_t* su = (const uint32_t*)(sf + (size_t)kQF * cap);
    float* df = C->q[C->cur ^ 1].as<float>();
    uint32_t* du = (uint32_t*)(df + (size_t)kQF * cap);
    g_hip_call = "launch k_sort_hist";
    hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)nb), dim3(256), 0, C->stream, sf, su, cap, n, chunk, n_keys, table);
    HIP_TRY(hipGetLastError());
    const uint32_t total = (uint32_t)(nb * (uint64_t)n_keys);
    const uint32_t tiles = (total + kScanTile - 1) / kScanTile;   /* <= 256 */
    unsigned int* sums = table + kSortTable;
    g_hip_call = "launch k_sort_tile_sums";
    hipLaunchKernelGGL(k_sort_tile_sums, dim3(tiles), dim3(256), 0, C->stream, (const unsigned int*)table, total, sums);
    HIP_TRY(hipGetLastError());
    g_hip_call = "launch k_sort_scan_tiles";
    hipLaunchKernelGGL(k_sort_scan_tiles, dim3(tiles), dim3(256), 0, C->stream, table, total,
                       (const unsigned int*)sums);
    HIP_TRY(hipGetLastError());
    g_hip_call = "launch k_sort_place";
    hipLaunchKernelGGL(k_sort_place, dim3((unsigned)nb), dim3(256), 0, C->stream, sf, su, df, du, cap, n, chunk,
                       n_keys, (const unsigned int*)table);
    HIP_TRY(hipGetLastError());
    C->cur ^= 1;
  }
  float kms = 0.f;
  if (timing) HIP_TRY(hipEventElapsedTime(&kms, C->evk0, C->evk1));
  {
    std::lock_guard<std::mutex> lk(C->mu);
    C->n_launches += 1;
    C->transport_ms += kms;
  }
  for (Job* j : inflight) {
    j->outstanding = inflight.size() > 1 ? C->host_count[4 + j->slot] : C->in_count;
    j->generations += 1;
    j->stats.transport_ms += kms;
    j->stats.transport_launches += 1;
    /* an electron still queued has at most K descendants (its energy is below
     * the beam energy), so the reservation shrinks with the queue */
    j->reserved = std::min<uint64_t>(j->reserved, (uint64_t)j->outstanding * j->K);
    /* generation limit: the job's remaining electrons are dropped and counted */
    if (j->generations > maxGen) stop_job(C, *j);
  }
}

void worker_main(monsel_ctx_s* C) {
  (void)hipSetDevice(C->device);
  const int maxActive = std::min(kMaxJobs, C->tuning.max_jobs_in_flight > 0 ? C->tuning.max_jobs_in_flight : kMaxJobs);
  for (;;) {
    std::vector<std::shared_ptr<Job>> admitted;
    {
      std::unique_lock<std::mutex> lk(C->mu);
      if (C->active.empty())
        C->cv_work.wait(lk, [&] { return C->stop || !C->queue.empty(); });
      if (C->stop && C->queue.empty() && C->active.empty()) return;
      while (!C->queue.empty() && (int)(C->active.size() + admitted.size()) < maxActive) {
        auto j = C->queue.front();
        /* one batch scene holds scenes of one compatibility mode: a job of
         * another mode waits until the jobs in flight are done */
        const uint32_t mode = j->sd->host.compat;
        if (!admitted.empty() && admitted[0]->sd->host.compat != mode) break;
        if (admitted.empty() && !C->active.empty() && C->active[0]->sd->host.compat != mode) break;
        C->queue.pop_front();
        admitted.push_back(j);
      }
    }
    /* jobs taken out of `active` to complete (finished ones are reset) */
    std::vector<std::shared_ptr<Job>> ended;
    try {
      if (C->active.empty() && C->in_count == 0 && !C->batch_parts.empty()) {
        /* idle engine: start a fresh batch scene */
        C->batch_parts.clear();
        C->batch.blob.reset();
        C->batch.dev = DevScene{};
      }
      for (auto& j : admitted) {
        try {
          setup_job(C, *j);
          C->active.push_back(j);
        } catch (const std::exception& ex) {
          /* copies it enqueued read the slot's pinned staging: they end
           * before the slot is given to another job */
          int st;
          std::string msg;
          status_of(ex, st, msg);
          (void)hipStreamSynchronize(C->stream);
          release_job(C, *j);
          finish_job(C, j, st, msg);
        }
      }
      if (C->active.empty()) continue;
      /* cancelled jobs: no further primaries; their queued electrons are
       * dropped when the next launch claims them */
      for (auto& j : C->active)
        if (j->cancel.load() && !j->cancelling) {
          j->cancelling = true;
          j->launched = j->total;
          stop_job(C, *j);
        }
      run_generation(C);
      /* completed jobs: all primaries launched and none of their electrons
       * queued; their reductions and read-backs go on the stream together */
      for (size_t k = 0; k < C->active.size();) {
        if (C->active[k]->launched >= C->active[k]->total && C->active[k]->outstanding == 0) {
          ended.push_back(C->active[k]);
          C->active.erase(C->active.begin() + k);
        } else {
          ++k;
        }
      }
      std::vector<size_t> at(ended.size() + 1, 0);
      for (size_t k = 0; k < ended.size(); ++k)
        at[k + 1] = at[k] + (ended[k]->cancelling ? 0 : (readback_bytes(*ended[k]) + 255) & ~(size_t)255);
      if (at.back() > 0) {
        C->pin_readback.ensure(at.back());
        for (size_t k = 0; k < ended.size(); ++k)
          if (!ended[k]->cancelling) complete_enqueue(C, *ended[k], C->pin_readback.as<char>() + at[k]);
        HIP_TRY(hipStreamSynchronize(C->stream));
      }
      for (size_t k = 0; k < ended.size(); ++k) {
        std::shared_ptr<Job> j = ended[k];
        if (!j->cancelling) complete_collect(C, *j, C->pin_readback.as<char>() + at[k]);
        release_job(C, *j);
        ended[k].reset();
        if (j->cancelling) finish_job(C, j, MONSEL_E_CANCELLED, "job cancelled");
        else finish_job(C, j, MONSEL_OK, "");
      }
      ended.clear();
    } catch (const std::exception& ex) {
      /* a device failure ends every job in flight */
      int st;
      std::string msg;
      status_of(ex, st, msg);
      (void)hipStreamSynchronize(C->stream);
      for (auto& j : ended)
        if (j) C->active.push_back(j);
      for (auto& j : C->active) {
        release_job(C, *j);
        finish_job(C, j, st, msg);
      }
      C->active.clear();
      C->in_count = 0;
    }
  }
}

std::shared_ptr<SceneDev> find_scene(monsel_ctx_s* C, monsel_scene_id id) {
  std::lock_guard<std::mutex> lk(C->mu);
  auto it = C->scenes.find(id);
  return it == C->scenes.end() ? nullptr : it->second;
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& ex) {
    int st;
    std::string msg;
    status_of(ex, st, msg);
    return set_err(st, msg);
  }
}

}  // namespace

/* ===================================================================== */
/* C-ABI                                                                  */
/* ===================================================================== */

extern "C" {

const char* monsel_last_error(void) { return g_err.c_str(); }

int monsel_abi_version(void) { return (int)MONSEL_ABI_VERSION; }

int monsel_create(int device_ordinal, const monsel_tuning* tuning, monsel_ctx* out) {
  if (!out) return set_err(MONSEL_E_INVALID, "out is NULL");
  return guarded([&] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw MonselError(MONSEL_E_DEVICE, "no HIP device");
    if (device_ordinal < 0 || device_ordinal >= count) throw MonselError(MONSEL_E_INVALID, "device ordinal out of range");
    std::unique_ptr<monsel_ctx_s> C(new monsel_ctx_s());
    C->device = device_ordinal;
    if (tuning) C->tuning = *tuning;
    HIP_TRY(hipSetDevice(device_ordinal));
    static_scratch_resolve(device_ordinal);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_ordinal));
    C->n_cu = prop.multiProcessorCount;
    C->hbm = prop.totalGlobalMem;
    C->name = prop.gcnArchName;
    HIP_TRY(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    C->pin_count.ensure((4 + kMaxJobs) * sizeof(unsigned int));
    C->host_count = C->pin_count.as<unsigned int>();
    C->pin_segs.ensure(3 * kMaxJobs * sizeof(uint32_t));
    C->host_segs = C->pin_segs.as<uint32_t>();
    C->pin_jobrec.ensure(kMaxJobs * sizeof(JobRec));
    C->jobrec = C->pin_jobrec.as<JobRec>();
    if (C->tuning.max_jobs_in_flight < 0 || C->tuning.primaries_per_launch < 0 || C->tuning.steps_per_launch < 0 ||
        C->tuning.queue_capacity < 0 || C->tuning.admission_order < 0 || C->tuning.admission_order > 1 ||
        C->tuning.sort_queue < -1 || C->tuning.sort_queue > 0 || C->tuning.geometry_batch < 0 ||
        C->tuning.geometry_batch > 64 || C->tuning.spawn_handoff < -1 || C->tuning.spawn_handoff > 0 ||
        C->tuning.tail_requeue < -1 || C->tuning.tail_requeue > 64)
      throw MonselError(MONSEL_E_INVALID, "tuning value out of range");
    C->jobtab.ensure(kMaxJobs * sizeof(JobRec));
    C->counts.ensure((4 + kMaxJobs) * sizeof(unsigned int));
    C->sort_table.ensure((kSortTable + kSortTable / kScanTile) * sizeof(unsigned int));
    C->segs.ensure(3 * kMaxJobs * sizeof(uint32_t));
    for (int k = kMaxJobs - 1; k >= 0; --k) C->free_slots.push_back(k);
    C->qcap = queue_capacity(C.get());
    if (C->qcap < 65536) throw MonselError(MONSEL_E_INVALID, "queue_capacity below 65536 entries");
    /* device code addresses the SoA queues and escape records as
     * field * capacity + slot in 32-bit arithmetic (7 fields) */
    if (C->qcap > kMaxQueueCapacity)
      throw MonselError(MONSEL_E_INVALID, "queue_capacity above " + std::to_string(kMaxQueueCapacity) + " entries");
    C->q[0].ensure((size_t)C->qcap * kQueueEntryBytes);
    C->q[1].ensure((size_t)C->qcap * kQueueEntryBytes);
    C->esc.ensure((size_t)C->qcap * kEscEntryBytes);
    HIP_TRY(hipEventCreate(&C->ev0));
    HIP_TRY(hipEventCreate(&C->ev1));
    HIP_TRY(hipEventCreate(&C->evk0));
    HIP_TRY(hipEventCreate(&C->evk1));
    int block = C->tuning.block_size > 0 ? C->tuning.block_size : 256;
    if (block % 64 != 0 || block > kLaneBlock)
      throw MonselError(MONSEL_E_INVALID, "block_size must be a multiple of 64, <= 256");
    if (C->tuning.blocks_per_cu > 0) C->blocks_per_cu = C->tuning.blocks_per_cu;
    else {
      int occ_t = 0, occ_f = 0;
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_t, k_transport<true>, block, 0));
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, k_transport<false>, block, 0));
      const int occ = std::min(occ_t, occ_f);
      C->blocks_per_cu = occ > 0 ? occ : 1;
    }
    monsel_ctx_s* raw = C.get();
    raw->worker = std::thread(worker_main, raw);
    *out = C.release();
    return MONSEL_OK;
  });
}

int monsel_destroy(monsel_ctx ctx) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->stop = true;
  }
  ctx->cv_work.notify_all();
  if (ctx->worker.joinable()) {
    try {
      ctx->worker.join();
    } catch (const std::exception& e) {
      return set_err(MONSEL_E_DEVICE, std::string("joining the dispatcher: ") + e.what());
    }
  }
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->scenes.clear();
  ctx->jobs.clear();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->evk0) (void)hipEventDestroy(ctx->evk0);
  if (ctx->evk1) (void)hipEventDestroy(ctx->evk1);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MONSEL_OK;
}

int monsel_device_info(monsel_ctx ctx, int* n_cu, size_t* hbm_bytes, char* name, size_t name_len) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  if (n_cu) *n_cu = ctx->n_cu;
  if (hbm_bytes) *hbm_bytes = ctx->hbm;
  if (name && name_len) {
    strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  return MONSEL_OK;
}

int monsel_ctx_counters(monsel_ctx ctx, uint64_t* launches, uint64_t* steps, double* transport_ms) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (launches) *launches = ctx->n_launches;
  if (steps) *steps = ctx->n_steps;
  if (transport_ms) *transport_ms = ctx->transport_ms;
  return MONSEL_OK;
}

int monsel_scene_footprint(const monsel_scene_desc* desc, size_t* hbm_bytes) {
  if (!hbm_bytes) return set_err(MONSEL_E_INVALID, "hbm_bytes is NULL");
  return guarded([&] {
    HostScene H = compile_scene(desc);
    *hbm_bytes = host_scene_bytes(H);
    return MONSEL_OK;
  });
}

int monsel_scene_upload(monsel_ctx ctx, const monsel_scene_desc* desc, monsel_scene_id* out) {
  if (!ctx || !out) return set_err(MONSEL_E_INVALID, "NULL argument");
  return guarded([&] {
    HostScene H = compile_scene(desc);
    auto sd = std::make_shared<SceneDev>();
    HIP_TRY(hipSetDevice(ctx->device));
    upload_scene(H, *sd);
    sd->host = std::move(H);
    std::lock_guard<std::mutex> lk(ctx->mu);
    monsel_scene_id id = ctx->next_id++;
    ctx->scenes[id] = sd;
    *out = id;
    return MONSEL_OK;
  });
}

int monsel_scene_release(monsel_ctx ctx, monsel_scene_id scene) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (!ctx->scenes.erase(scene)) return set_err(MONSEL_E_NOTFOUND, "unknown scene");
  return MONSEL_OK;
}

int monsel_job_footprint(monsel_ctx ctx, const monsel_job_desc* job, size_t* hbm_bytes) {
  if (!ctx || !job || !hbm_bytes) return set_err(MONSEL_E_INVALID, "NULL argument");
  /* per-job buffers (the generation queues belong to the context: their
   * share is the job's queue reservation) */
  const uint64_t total = (uint64_t)job->n_pixels * (uint64_t)job->primaries_per_pixel;
  const uint32_t K = secondaries_per_primary(job->beam_energy / kE);
  const uint64_t resv = std::min<uint64_t>(total * K, ctx->qcap);
  const int nbins = job->n_energy_bins > 0 ? job->n_energy_bins : 1;
  *hbm_bytes = (size_t)(total * 16 + (uint64_t)job->n_pixels * (6 * 8 + 24) + (uint64_t)hist_len(nbins) * 8 + 64 +
                        resv * (2 * kQueueEntryBytes + kEscEntryBytes) +
                        (job->log_detected > 0 ? 16 + (uint64_t)job->log_detected * sizeof(monsel_detected) : 0));
  return MONSEL_OK;
}

static int validate_job(const monsel_job_desc* job) {
  if (!job) return set_err(MONSEL_E_INVALID, "NULL job");
  if (job->n_pixels <= 0 || !job->pixel_center) return set_err(MONSEL_E_INVALID, "job needs pixels");
  if (job->primaries_per_pixel <= 0) return set_err(MONSEL_E_INVALID, "primaries_per_pixel must be positive");
  if (!(job->beam_energy > 0)) return set_err(MONSEL_E_INVALID, "beam energy must be positive");
  if (job->beam_width < 0) return set_err(MONSEL_E_INVALID, "beam width must be >= 0");
  uint64_t total = (uint64_t)job->n_pixels * (uint64_t)job->primaries_per_pixel;
  if (total + job->primary_offset > 0xFFFFFFFFull) return set_err(MONSEL_E_INVALID, "more than 2^32 primaries in one job");
  if (job->n_energy_bins > (1 << 20)) return set_err(MONSEL_E_INVALID, "too many energy bins");
  if (job->log_detected < 0 || job->log_detected > ((int64_t)1 << 32))
    return set_err(MONSEL_E_INVALID, "log_detected must be in [0, 2^32]");
  return MONSEL_OK;
}

int monsel_job_submit(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, monsel_job_id* out) {
  if (!ctx || !out) return set_err(MONSEL_E_INVALID, "NULL argument");
  int v = validate_job(job);
  if (v) return v;
  auto sd = find_scene(ctx, scene);
  if (!sd) return set_err(MONSEL_E_NOTFOUND, "unknown scene");
  /* the job record and its copy of the pixel centres can fail to allocate */
  return guarded([&] {
    auto j = std::make_shared<Job>();
    j->scene = scene;
    j->sd = sd;   /* the job holds its scene: a scene released meanwhile lives until the job ends */
    j->desc = *job;
    j->centers.assign(job->pixel_center, job->pixel_center + 3 * (size_t)job->n_pixels);
    j->desc.pixel_center = j->centers.data();
    j->t_submit = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      j->id = ctx->next_id++;
      ctx->jobs[j->id] = j;
      ctx->queue.push_back(j);
      *out = j->id;
    }
    ctx->cv_work.notify_one();
    return MONSEL_OK;
  });
}

int monsel_job_poll(monsel_ctx ctx, monsel_job_id job, int* done) {
  if (!ctx || !done) return set_err(MONSEL_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->jobs.find(job);
  if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
  *done = it->second->done ? 1 : 0;
  return MONSEL_OK;
}

int monsel_job_wait(monsel_ctx ctx, monsel_job_id job, int64_t timeout_us) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  std::unique_lock<std::mutex> lk(ctx->mu);
  auto it = ctx->jobs.find(job);
  if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
  auto j = it->second;
  if (timeout_us < 0) ctx->cv_done.wait(lk, [&] { return j->done; });
  else if (!ctx->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return j->done; }))
    return set_err(MONSEL_E_TIMEOUT, "job not finished");
  if (j->status != MONSEL_OK) return set_err(j->status, j->error);
  return MONSEL_OK;
}

int monsel_job_read(monsel_ctx ctx, monsel_job_id job, monsel_tallies* out) {
  if (!ctx || !out) return set_err(MONSEL_E_INVALID, "NULL argument");
  std::shared_ptr<Job> j;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(job);
    if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
    j = it->second;
    if (!j->done) return set_err(MONSEL_E_NOTREADY, "job still running");
    if (j->status != MONSEL_OK) return set_err(j->status, j->error);
  }
  auto cp = [](auto* dst, const auto& src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
  };
  cp(out->back, j->back);
  cp(out->fwd, j->fwd);
  cp(out->se, j->se);
  cp(out->se_sq, j->se_sq);
  cp(out->bse_sq, j->bse_sq);
  cp(out->back_energy, j->back_energy);
  cp(out->back_energy_hist, j->hback);
  cp(out->fwd_energy_hist, j->hfwd);
  cp(out->elevation_hist, j->helev);
  cp(out->azimuth_hist, j->hazim);
  return MONSEL_OK;
}

int monsel_job_stats_get(monsel_ctx ctx, monsel_job_id job, monsel_job_stats* out) {
  if (!ctx || !out) return set_err(MONSEL_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->jobs.find(job);
  if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
  if (!it->second->done) return set_err(MONSEL_E_NOTREADY, "job still running");
  *out = it->second->stats;
  return MONSEL_OK;
}

int monsel_job_read_detected(monsel_ctx ctx, monsel_job_id job, monsel_detected* out, int64_t max,
                             int64_t* n_detected) {
  if (!ctx || !n_detected || (max > 0 && !out)) return set_err(MONSEL_E_INVALID, "NULL argument");
  std::shared_ptr<Job> j;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(job);
    if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
    j = it->second;
    if (!j->done) return set_err(MONSEL_E_NOTREADY, "job still running");
    if (j->status != MONSEL_OK) return set_err(j->status, j->error);
  }
  *n_detected = (int64_t)j->n_detected;
  const size_t k = std::min<size_t>(j->detected.size(), max > 0 ? (size_t)max : 0);
  if (k) std::copy(j->detected.begin(), j->detected.begin() + k, out);
  return MONSEL_OK;
}

int monsel_job_release(monsel_ctx ctx, monsel_job_id job) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->jobs.find(job);
  if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
  if (!it->second->done) return set_err(MONSEL_E_NOTREADY, "job still running");
  ctx->jobs.erase(it);
  return MONSEL_OK;
}

int monsel_job_cancel(monsel_ctx ctx, monsel_job_id job) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  std::shared_ptr<Job> waiting;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(job);
    if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
    auto j = it->second;
    if (j->done) return MONSEL_OK;
    auto q = std::find(ctx->queue.begin(), ctx->queue.end(), j);
    if (q != ctx->queue.end()) {
      /* not admitted yet: it holds nothing on the device */
      ctx->queue.erase(q);
      waiting = j;
    } else {
      j->cancel.store(true);
    }
  }
  if (waiting) finish_job(ctx, waiting, MONSEL_E_CANCELLED, "job cancelled");
  ctx->cv_work.notify_one();
  return MONSEL_OK;
}

int monsel_run(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, monsel_tallies* out,
               monsel_job_stats* stats) {
  monsel_job_id id;
  int st = monsel_job_submit(ctx, scene, job, &id);
  if (st) return st;
  st = monsel_job_wait(ctx, id, -1);
  if (st == MONSEL_OK && out) st = monsel_job_read(ctx, id, out);
  if (st == MONSEL_OK && stats) st = monsel_job_stats_get(ctx, id, stats);
  std::string keep = g_err;
  monsel_job_release(ctx, id);
  g_err = keep;
  return st;
}

/* Allocation-free variant of monsel_run: the same job on the same kernels,
 * run on the calling thread, with every device buffer beyond the caller's
 * inputs and outputs carved out of the static array `scratch`. */
int monsel_run_static(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, monsel_tallies* out,
                      monsel_job_stats* stats) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  const int v = validate_job(job);
  if (v) return v;
  if (job->log_detected > 0)
    return set_err(MONSEL_E_UNSUPPORTED, "monsel_run_static keeps no detected-electron log");
  if (ctx->device >= kMaxStaticDevices)
    return set_err(MONSEL_E_UNSUPPORTED, "monsel_run_static: device ordinal above " +
                                             std::to_string(kMaxStaticDevices - 1));
  std::shared_ptr<SceneDev> sd = find_scene(ctx, scene);
  if (!sd) return set_err(MONSEL_E_NOTFOUND, "unknown scene");
  return guarded([&] {
    const auto t0 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(g_static_mu[ctx->device]);
    char* const base = (char*)g_static_ptr[ctx->device];
    if (!base)
      throw MonselError(MONSEL_E_DEVICE, std::string("monsel_run_static: no device address for the static buffer: ") +
                                             hipGetErrorString(g_static_status[ctx->device]));
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t stream = ctx->stream;
    const monsel_job_desc& jd = *job;
    const int np = jd.n_pixels;
    const int nbins = jd.n_energy_bins > 0 ? jd.n_energy_bins : 1;
    const double E0eV = jd.beam_energy / kE;
    const uint64_t total = (uint64_t)np * (uint64_t)jd.primaries_per_pixel;
    const uint32_t K = (uint32_t)std::min<double>(4096., 16. + E0eV * sd->host.secondaries_per_eV);
    /* layout: fixed-size records, then the job's tallies (cleared), then the
     * pixel centres, then the queues take what is left */
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    const size_t o_rec = take(sizeof(JobRec)), o_cnt = take(8 * 4), o_seg = take(4 * 4);
    const size_t o_zero = off;
    const size_t o_st = take(8 * 8 * kStatStripes), o_hist = take((size_t)hist_len(nbins) * 8);
    const size_t o_px = take((size_t)np * 6 * 8);
    const size_t o_pc = take(total * 4), o_pf = take(total * 4), o_pe = take(total * 8);
    const size_t zero_end = off;
    const size_t o_dbl = take((size_t)np * 3 * 8), o_cen = take((size_t)np * 3 * 8);
    const size_t per_entry = 2 * kQueueEntryBytes + kEscEntryBytes;
    const uint64_t cap = off + 3 * 256 < kStaticScratchBytes
                             ? std::min<uint64_t>((kStaticScratchBytes - off - 3 * 256) / per_entry, kMaxQueueCapacity)
                             : 0;
    if (cap < 4096 || cap < K)
      throw MonselError(MONSEL_E_TOOLARGE,
                        "monsel_run_static: " + std::to_string(np) + " pixels x " +
                            std::to_string(jd.primaries_per_pixel) + " primaries need " + std::to_string(off) +
                            " bytes of tallies, leaving room for " + std::to_string(cap) +
                            " queued electrons (at least " + std::to_string(std::max<uint32_t>(K, 4096)) +
                            " needed) in its " + std::to_string(kStaticScratchBytes >> 20) + " MiB static buffer");
    const size_t o_q0 = take(cap * kQueueEntryBytes), o_q1 = take(cap * kQueueEntryBytes);
    const size_t o_esc = take(cap * kEscEntryBytes);
    if (off > kStaticScratchBytes) throw MonselError(MONSEL_E_DEVICE, "monsel_run_static: layout overflow");
    JobRec* d_rec = (JobRec*)(base + o_rec);
    unsigned int* d_cnt = (unsigned int*)(base + o_cnt);
    uint32_t* d_seg = (uint32_t*)(base + o_seg);
    unsigned long long* d_st = (unsigned long long*)(base + o_st);
    unsigned long long* d_hist = (unsigned long long*)(base + o_hist);
    unsigned long long* d_px = (unsigned long long*)(base + o_px);
    double* d_dbl = (double*)(base + o_dbl);
    float* d_q[2] = {(float*)(base + o_q0), (float*)(base + o_q1)};
    float* d_esc = (float*)(base + o_esc);
    /* the previous call's contents stay in the array: clear the tallies */
    const size_t nz = (zero_end - o_zero) / 8;
    g_hip_call = "launch k_zero_u64";
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<size_t>((nz + 255) / 256, (size_t)ctx->n_cu * 8)),
                       dim3(256), 0, stream, (unsigned long long*)(base + o_zero), nz);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(base + o_cen, jd.pixel_center, (size_t)np * 3 * 8, hipMemcpyHostToDevice, stream));
    const HostScene& H = sd->host;
    JobRec R;
    memset(&R, 0, sizeof(R));
    R.lin0 = job_lineage(jd.seed);
    R.E0 = (float)E0eV;
    R.width = (float)(jd.beam_width * 1e9);
    R.centers = (const double*)(base + o_cen);
    R.n_pixels = np;
    R.ppp = jd.primaries_per_pixel;
    R.prim_base = (uint32_t)jd.primary_offset;
    R.root = 0; /* the scene's own chamber region (the scene is not merged with others) */
    R.max_steps = jd.max_steps > 0 ? jd.max_steps : 1000000;
    R.nbins = nbins;
    R.maxSEbin = (int)(jd.se_cutoff_ev / (E0eV / nbins));
    R.E0eV = E0eV;
    for (int k = 0; k < 3; ++k) R.chamber_c[k] = H.chamber_c[k];
    R.chamber_r = H.chamber_r;
    R.prim_counts = (uint32_t*)(base + o_pc);
    R.prim_fwd = (uint32_t*)(base + o_pf);
    R.prim_energy = (unsigned long long*)(base + o_pe);
    R.hist = d_hist;
    R.stats = d_st;
    HIP_TRY(hipMemcpyAsync(d_rec, &R, sizeof(JobRec), hipMemcpyHostToDevice, stream));
    /* generations, as run_generation does for one job (no queue sort) */
    const monsel_tuning& T = ctx->tuning;
    const int maxGen = T.max_generations > 0 ? T.max_generations : 100000;
    const uint64_t perLaunch = primaries_per_launch_for(T.primaries_per_launch, cap);
    const int block = T.block_size > 0 ? T.block_size : 256;
    const int grid = ctx->n_cu * (ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : 4);
    uint64_t launched = 0, reserved = 0;
    uint32_t in_count = 0;
    int cur = 0, generations = 0;
    LdsJobs lj;
    memset(&lj, 0, sizeof(lj));
    if ((size_t)hist_len(nbins) * 4 <= 64 * 1024) {
      lj.n = 1;
      lj.off[1] = hist_len(nbins);
    }
    while (launched < total || in_count > 0) {
      const uint64_t avail = cap > reserved ? cap - reserved : 0;
      const uint64_t n = std::min<uint64_t>(std::min<uint64_t>(total - launched, avail / K), perLaunch);
      if (n == 0 && in_count == 0) throw MonselError(MONSEL_E_DEVICE, "monsel_run_static: no work admitted");
      const uint32_t segs[4] = {(uint32_t)n, 0u, (uint32_t)launched, 0u};
      HIP_TRY(hipMemcpyAsync(d_seg, segs, sizeof(segs), hipMemcpyHostToDevice, stream));
      g_hip_call = "launch k_zero_u64";
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, (unsigned long long*)d_cnt, (size_t)4);
      HIP_TRY(hipGetLastError());
      GenDev G;
      memset(&G, 0, sizeof(G));
      G.jobs = d_rec;
      G.seg_end = d_seg;
      G.seg_job = d_seg + 1;
      G.seg_first = d_seg + 2;
      G.n_seg = n > 0 ? 1 : 0;
      G.n_prim = (uint32_t)n;
      G.in_f = d_q[cur];
      G.in_u = (const uint32_t*)(d_q[cur] + (size_t)kQF * cap);
      G.in_cap = (uint32_t)cap;
      G.in_n = in_count;
      G.claim = d_cnt;
      G.out_f = d_q[cur ^ 1];
      G.out_u = (uint32_t*)(d_q[cur ^ 1] + (size_t)kQF * cap);
      G.out_cap = (uint32_t)cap;
      G.out_count = d_cnt + 1;
      G.esc_f = d_esc;
      G.esc_u = (uint32_t*)(d_esc + (size_t)kEF * cap);
      G.esc_cap = (uint32_t)cap;
      G.esc_count = d_cnt + 2;
      G.steps_per_launch = T.steps_per_launch;
      G.geo_batch = geometry_batch_for(T.geometry_batch, sd->step_prims, n + in_count);
      G.handoff = T.spawn_handoff >= 0 ? 1 : 0;
      G.tail_lanes = T.tail_requeue < 0 || n + in_count < kTailMinItems
                         ? 0
                         : (T.tail_requeue > 0 ? T.tail_requeue : kDefaultTailLanes);
      g_hip_call = "launch k_transport";
      if (sd->dev.has_tabulated)
        hipLaunchKernelGGL(k_transport<true>, dim3(grid), dim3(block), 0, stream, sd->dev, G);
      else
        hipLaunchKernelGGL(k_transport<false>, dim3(grid), dim3(block), 0, stream, sd->dev, G);
      HIP_TRY(hipGetLastError());
      g_hip_call = "launch k_tally_escapes";
      hipLaunchKernelGGL(k_tally_escapes, dim3(ctx->n_cu * 4), dim3(256), (size_t)lj.off[lj.n] * 4, stream, sd->dev,
                         (const JobRec*)d_rec, (const float*)G.esc_f, (const uint32_t*)G.esc_u, (uint32_t)cap,
                         (const unsigned int*)(d_cnt + 2), lj);
      HIP_TRY(hipGetLastError());
      unsigned int hc[4] = {0u, 0u, 0u, 0u};
      HIP_TRY(hipMemcpyAsync(hc, d_cnt, sizeof(hc), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      if ((uint64_t)hc[0] < n + (uint64_t)in_count)
        throw MonselError(MONSEL_E_DEVICE, "transport launch ended with unclaimed items");
      in_count = hc[1] > cap ? (uint32_t)cap : hc[1];
      cur ^= 1;
      launched += n;
      reserved = std::min<uint64_t>(reserved + n * K, (uint64_t)in_count * K);
      if (++generations > maxGen && !R.killed) {
        /* generation limit: the remaining electrons are dropped and counted */
        R.killed = 1;
        HIP_TRY(hipMemcpyAsync(&d_rec->killed, &R.killed, sizeof(int), hipMemcpyHostToDevice, stream));
      }
    }
    if (total > 0) {
      g_hip_call = "launch k_reduce_primaries";
      hipLaunchKernelGGL(k_reduce_primaries, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                         (const JobRec*)d_rec, 0, (uint32_t)total, d_px);
      HIP_TRY(hipGetLastError());
    }
    g_hip_call = "launch k_pixels_to_double";
    hipLaunchKernelGGL(k_pixels_to_double, dim3((unsigned)((3 * np + 255) / 256)), dim3(256), 0, stream,
                       (const unsigned long long*)d_px, np, d_dbl);
    HIP_TRY(hipGetLastError());
    unsigned long long sts[8 * kStatStripes];
    HIP_TRY(hipMemcpyAsync(sts, d_st, sizeof(sts), hipMemcpyDeviceToHost, stream));
    if (out) {
      auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (dst) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
      };
      const size_t nb2 = (size_t)nbins + 2;
      d2h(out->back, d_px, (size_t)np * 8);
      d2h(out->fwd, d_px + np, (size_t)np * 8);
      d2h(out->se, d_px + 2 * (size_t)np, (size_t)np * 8);
      d2h(out->se_sq, d_dbl, (size_t)np * 8);
      d2h(out->bse_sq, d_dbl + np, (size_t)np * 8);
      d2h(out->back_energy, d_dbl + 2 * (size_t)np, (size_t)np * 8);
      d2h(out->back_energy_hist, d_hist, nb2 * 8);
      d2h(out->fwd_energy_hist, d_hist + nb2, nb2 * 8);
      d2h(out->elevation_hist, d_hist + 2 * nb2, 182 * 8);
      d2h(out->azimuth_hist, d_hist + 2 * nb2 + 182, 362 * 8);
    }
    HIP_TRY(hipStreamSynchronize(stream));
    if (stats) {
      unsigned long long st[8] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
      for (int k = 0; k < 8 * kStatStripes; ++k) st[k % 8] += sts[k];
      memset(stats, 0, sizeof(*stats));
      stats->primaries = total;
      stats->steps = st[kStatSteps];
      stats->electrons = st[kStatElectrons];
      stats->escapes = st[kStatEscapes];
      stats->watchdog_drops = st[kStatWatchdog];
      stats->queue_overflow = st[kStatOverflow];
      stats->generations = generations;
      stats->transport_launches = generations;
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      stats->device_ms = ms;
      stats->total_ms = ms;
    }
    return MONSEL_OK;
  });
}

}  /* extern "C" */

/* ---- probes ------------------------------------------------------------ */

namespace {
template <class T>
struct Scratch {
  T* d = nullptr;
  explicit Scratch(size_t n) {
    if (n == 0) n = 1;
    HIP_TRY(hipMalloc(&d, n * sizeof(T)));
  }
  ~Scratch() {
    if (d) (void)hipFree(d);
  }
};

int probe_prologue(monsel_ctx ctx, monsel_scene_id scene, std::shared_ptr<SceneDev>& sd) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  sd = find_scene(ctx, scene);
  if (!sd) return set_err(MONSEL_E_NOTFOUND, "unknown scene");
  if (hipSetDevice(ctx->device) != hipSuccess) return set_err(MONSEL_E_DEVICE, "hipSetDevice failed");
  return MONSEL_OK;
}

inline unsigned grid_for(int n) { return (unsigned)((n + 255) / 256); }
}  // namespace

extern "C" {

int monsel_probe_rates(monsel_ctx ctx, monsel_scene_id scene, int msm, int n, const float* energies_ev, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (msm < 0 || msm >= sd->dev.n_msm || n < 0 || (n > 0 && (!energies_ev || !out)))
    return set_err(MONSEL_E_INVALID, "probe_rates arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    std::vector<DMSM> m(1);
    HIP_TRY(hipMemcpy(m.data(), sd->dev.msms + msm, sizeof(DMSM), hipMemcpyDeviceToHost));
    int w = m[0].n_mech + 1;
    Scratch<float> dE(n), dO((size_t)n * w);
    HIP_TRY(hipMemcpy(dE.d, energies_ev, 4 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_rates, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, msm, n, dE.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 4 * (size_t)n * w, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_elastic_angle(monsel_ctx ctx, monsel_scene_id scene, int msm, int element, int n,
                               const float* energies_ev, const float* u, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (msm < 0 || msm >= sd->dev.n_msm || n < 0 || (n > 0 && (!energies_ev || !u || !out)))
    return set_err(MONSEL_E_INVALID, "probe_elastic_angle arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    DMSM m;
    HIP_TRY(hipMemcpy(&m, sd->dev.msms + msm, sizeof(DMSM), hipMemcpyDeviceToHost));
    int mech = -1;
    for (int k = 0; k < m.n_mech; ++k) {
      DMech dm;
      HIP_TRY(hipMemcpy(&dm, sd->dev.mechs + m.mech0 + k, sizeof(DMech), hipMemcpyDeviceToHost));
      if (dm.kind == MONSEL_MECH_NISTMOTT_ELASTIC || dm.kind == MONSEL_MECH_BROWNING_ELASTIC ||
          dm.kind == MONSEL_MECH_RUTHERFORD_ELASTIC) {
        if (element < 0 || element >= dm.nelem) throw MonselError(MONSEL_E_INVALID, "element slot out of range");
        mech = m.mech0 + k;
      }
    }
    if (mech < 0) throw MonselError(MONSEL_E_INVALID, "MSM has no elastic mechanism");
    Scratch<float> dE(n), dU(2 * (size_t)n), dO(n);
    HIP_TRY(hipMemcpy(dE.d, energies_ev, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dU.d, u, 8 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_elastic, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, mech, element, n, dE.d, dU.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_csd(monsel_ctx ctx, monsel_scene_id scene, int msm, int n, const float* len_nm,
                     const float* energies_ev, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (msm < 0 || msm >= sd->dev.n_msm || n < 0 || (n > 0 && (!len_nm || !energies_ev || !out)))
    return set_err(MONSEL_E_INVALID, "probe_csd arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    Scratch<float> dL(n), dE(n), dO(n);
    HIP_TRY(hipMemcpy(dL.d, len_nm, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dE.d, energies_ev, 4 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_csd, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, msm, n, dL.d, dE.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_find_end_of_step(monsel_ctx ctx, monsel_scene_id scene, int region, int n, const float* p0,
                                  const float* p1, float* out_u, int* out_next, float* out_normal) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (region < 0 || region >= sd->dev.n_regions || n < 0 || (n > 0 && (!p0 || !p1 || !out_u || !out_next || !out_normal)))
    return set_err(MONSEL_E_INVALID, "probe_find_end_of_step arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    Scratch<float> dA(3 * (size_t)n), dB(3 * (size_t)n), dU(n), dN(3 * (size_t)n);
    Scratch<int> dX(n);
    HIP_TRY(hipMemcpy(dA.d, p0, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dB.d, p1, 12 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_step_end, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, region, n, dA.d, dB.d, dU.d,
                       dX.d, dN.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out_u, dU.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out_next, dX.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out_normal, dN.d, 12 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_safety(monsel_ctx ctx, monsel_scene_id scene, int region, int n, const float* pos, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (region < 0 || region >= sd->dev.n_regions || n < 0 || (n > 0 && (!pos || !out)))
    return set_err(MONSEL_E_INVALID, "probe_safety arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    Scratch<float> dP(3 * (size_t)n), dO(n);
    HIP_TRY(hipMemcpy(dP.d, pos, 12 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_safety, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, region, n, dP.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_locate(monsel_ctx ctx, monsel_scene_id scene, int n, const float* pos, int* out_region) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (n < 0 || (n > 0 && (!pos || !out_region))) return set_err(MONSEL_E_INVALID, "probe_locate arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    Scratch<float> dP(3 * (size_t)n);
    Scratch<int> dO(n);
    HIP_TRY(hipMemcpy(dP.d, pos, 12 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_locate, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, n, dP.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out_region, dO.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_nu_table(monsel_ctx ctx, monsel_scene_id scene, int table, int order, int n, const float* x,
                          float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (table < 0 || table >= sd->dev.n_nu || order < 1 || order > 3 || n < 0 || (n > 0 && (!x || !out)))
    return set_err(MONSEL_E_INVALID, "probe_nu_table arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    DNUTable t;
    HIP_TRY(hipMemcpy(&t, sd->dev.nu + table, sizeof(DNUTable), hipMemcpyDeviceToHost));
    for (int k = 0; k < t.dim; ++k)
      if (t.n[k] < order + 1) throw MonselError(MONSEL_E_INVALID, "table axis shorter than order + 1");
    Scratch<float> dX((size_t)n * t.dim), dO(n);
    HIP_TRY(hipMemcpy(dX.d, x, 4 * (size_t)n * t.dim, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_nu, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, table, order, n, dX.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_scatter(monsel_ctx ctx, monsel_scene_id scene, int msm, int mech, int n, const float* prev_ev,
                         const float* energies_ev, const float* dir, const float* u, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (msm < 0 || msm >= sd->dev.n_msm || n < 0 || (n > 0 && (!prev_ev || !energies_ev || !dir || !u || !out)))
    return set_err(MONSEL_E_INVALID, "probe_scatter arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    DMSM m;
    HIP_TRY(hipMemcpy(&m, sd->dev.msms + msm, sizeof(DMSM), hipMemcpyDeviceToHost));
    if (m.is_null || mech < 0 || mech >= m.n_mech) throw MonselError(MONSEL_E_INVALID, "mechanism slot out of range");
    Scratch<float> dP(n), dE(n), dD(3 * (size_t)n), dU(6 * (size_t)n), dO(11 * (size_t)n);
    HIP_TRY(hipMemcpy(dP.d, prev_ev, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dE.d, energies_ev, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dD.d, dir, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dU.d, u, 24 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_scatter, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, m.mech0 + mech, n, dP.d, dE.d,
                       dD.d, dU.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 44 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_barrier(monsel_ctx ctx, monsel_scene_id scene, int region_from, int region_to, int n,
                         const float* dir, const float* normal, const float* energies_ev, const float* u, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  if (region_from < 0 || region_from >= sd->dev.n_regions || region_to < 0 || region_to >= sd->dev.n_regions ||
      region_from == region_to || n < 0 || (n > 0 && (!dir || !normal || !energies_ev || !u || !out)))
    return set_err(MONSEL_E_INVALID, "probe_barrier arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    DRegion rf;
    HIP_TRY(hipMemcpy(&rf, sd->dev.regions + region_from, sizeof(DRegion), hipMemcpyDeviceToHost));
    DMSM m;
    HIP_TRY(hipMemcpy(&m, sd->dev.msms + rf.msm, sizeof(DMSM), hipMemcpyDeviceToHost));
    if (m.is_null || m.barrier_kind == MONSEL_BARRIER_NONE)
      throw MonselError(MONSEL_E_INVALID, "region has no barrier mechanism");
    Scratch<float> dD(3 * (size_t)n), dN(3 * (size_t)n), dE(n), dU(n), dO(6 * (size_t)n);
    HIP_TRY(hipMemcpy(dD.d, dir, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dN.d, normal, 12 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dE.d, energies_ev, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dU.d, u, 4 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_barrier, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, region_from, region_to, n, dD.d,
                       dN.d, dE.d, dU.d, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 24 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_philox(monsel_ctx ctx, int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out,
                        float* out_u) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  if (n < 0 || (n > 0 && (!ctr || !key || !out))) return set_err(MONSEL_E_INVALID, "probe_philox arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    HIP_TRY(hipSetDevice(ctx->device));
    Scratch<uint32_t> dC(4 * (size_t)n), dK(2 * (size_t)n), dO(4 * (size_t)n);
    Scratch<float> dU(out_u ? 4 * (size_t)n : 1);
    HIP_TRY(hipMemcpy(dC.d, ctr, 16 * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dK.d, key, 8 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_philox, dim3(grid_for(n)), dim3(256), 0, 0, n, dC.d, dK.d, dO.d,
                       out_u ? dU.d : nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 16 * (size_t)n, hipMemcpyDeviceToHost));
    if (out_u) HIP_TRY(hipMemcpy(out_u, dU.d, 16 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_hist_bin(monsel_ctx ctx, int n, const double* val, double min, double max, int nbins, int compat,
                          int* out) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  if (n < 0 || nbins < 1 || !(max > min) || (n > 0 && (!val || !out)))
    return set_err(MONSEL_E_INVALID, "probe_hist_bin arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    HIP_TRY(hipSetDevice(ctx->device));
    Scratch<double> dV(n);
    Scratch<int> dO(n);
    HIP_TRY(hipMemcpy(dV.d, val, 8 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_hist_bin, dim3(grid_for(n)), dim3(256), 0, 0, n, dV.d, min, max, nbins, compat,
                       dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

int monsel_probe_primaries(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, int n, float* out) {
  std::shared_ptr<SceneDev> sd;
  int st = probe_prologue(ctx, scene, sd);
  if (st) return st;
  st = validate_job(job);
  if (st) return st;
  if (n < 0 || (n > 0 && !out) || (uint64_t)n > (uint64_t)job->n_pixels * (uint64_t)job->primaries_per_pixel)
    return set_err(MONSEL_E_INVALID, "probe_primaries arguments");
  if (n == 0) return MONSEL_OK;
  return guarded([&] {
    Scratch<double> dCen(3 * (size_t)job->n_pixels);
    Scratch<JobRec> dJ(1);
    Scratch<float> dO(4 * (size_t)n);
    HIP_TRY(hipMemcpy(dCen.d, job->pixel_center, 24 * (size_t)job->n_pixels, hipMemcpyHostToDevice));
    JobRec R;
    memset(&R, 0, sizeof(R));
    R.lin0 = job_lineage(job->seed);
    R.E0 = (float)(job->beam_energy / kE);
    R.width = (float)(job->beam_width * 1e9);
    R.centers = dCen.d;
    R.n_pixels = job->n_pixels;
    R.ppp = job->primaries_per_pixel;
    R.prim_base = (uint32_t)job->primary_offset;
    R.root = 0;
    HIP_TRY(hipMemcpy(dJ.d, &R, sizeof(R), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe_primaries, dim3(grid_for(n)), dim3(256), 0, 0, sd->dev, (const JobRec*)dJ.d, n, dO.d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dO.d, 16 * (size_t)n, hipMemcpyDeviceToHost));
    return MONSEL_OK;
  });
}

#ifdef MONSEL_PHASE_TIMING
/* Phase-timing builds only: copies out and clears the transport kernel's
 * phase counters on `device` (out[3k..3k+2] = cycles, entries, active lanes
 * of phase k; n >= 3 * kPhases).  Waits for the device first. */
int monsel_phase_timing_read(int device, uint64_t* out, int n) {
  if (!out || n < 3 * kPhases) return set_err(MONSEL_E_INVALID, "phase_timing_read arguments");
  return guarded([&] {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase), sizeof(unsigned long long) * 3 * kPhases));
    std::vector<unsigned long long> zero(3 * kPhases, 0ull);
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_phase), zero.data(), sizeof(unsigned long long) * 3 * kPhases));
    return MONSEL_OK;
  });
}

/* Phase-timing builds only: per-launch step-count histogram of ending
 * electrons (hist[kStepBuckets]) and the wave spans recorded since the last
 * call (spans[2 * max_spans], start/end in 100 MHz ticks); returns the
 * number of spans (also clears both). */
int monsel_phase_timing_spans(int device, uint64_t* hist, uint64_t* spans, int max_spans, int* n_spans) {
  if (!hist || !spans || !n_spans || max_spans < 0) return set_err(MONSEL_E_INVALID, "phase_timing_spans arguments");
  return guarded([&] {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceSynchronize());
    unsigned int n = 0;
    HIP_TRY(hipMemcpyFromSymbol(&n, HIP_SYMBOL(g_span_count), sizeof(n)));
    n = std::min<unsigned int>(n, (unsigned int)std::min(max_spans, kMaxSpans));
    HIP_TRY(hipMemcpyFromSymbol(hist, HIP_SYMBOL(g_step_hist), sizeof(unsigned long long) * kStepBuckets));
    if (n) HIP_TRY(hipMemcpyFromSymbol(spans, HIP_SYMBOL(g_spans), sizeof(unsigned long long) * 2 * n));
    *n_spans = (int)n;
    std::vector<unsigned long long> zero(kStepBuckets, 0ull);
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_step_hist), zero.data(), sizeof(unsigned long long) * kStepBuckets));
    const unsigned int z = 0;
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_span_count), &z, sizeof(z)));
    return MONSEL_OK;
  });
}
#endif

}  /* extern "C" */
