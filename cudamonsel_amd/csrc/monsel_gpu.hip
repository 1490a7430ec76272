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
#include <map>
#include <memory>
#include <mutex>
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

struct MonselError {
  int status;
  std::string msg;
  MonselError(int s, std::string m) : status(s), msg(std::move(m)) {}
};

/* ===================================================================== */
/* device kernels                                                         */
/* ===================================================================== */

constexpr int kQF = 7; /* float fields per queued electron: p, d, E */
constexpr int kQU = 4; /* u32 fields: region, primary, lineage lo/hi */
constexpr int kEF = 7; /* float fields per escape record: tally origin, d, E */
constexpr int kEU = 2; /* u32 fields: primary, through-the-chamber-wall flag */
constexpr int kStatSteps = 0, kStatElectrons = 1, kStatEscapes = 2, kStatWatchdog = 3, kStatOverflow = 4;

struct JobDev {
  uint32_t k0, k1;
  float E0;         /* eV */
  float width;      /* nm */
  const double* centers; /* m */
  int n_pixels;
  int ppp;
  uint32_t prim_base;   /* RNG id of the chunk's first primary */
  uint32_t chunk_start; /* job-relative index of the chunk's first primary */
  uint32_t chunk_n;
  int nbins;
  double E0eV;
  int maxSEbin;
  int max_steps;
  uint32_t* prim_counts; /* [chunk_n]: low 16 bits SE, high 16 bits BSE */
  unsigned long long* fwd_px;
  double* back_energy_px;
  unsigned long long* hist_back;
  unsigned long long* hist_fwd;
  unsigned long long* hist_elev;
  unsigned long long* hist_azim;
  unsigned long long* stats;
};

struct GenDev {
  int mode; /* 0: create primaries; 1: read queue */
  const float* in_f;
  const uint32_t* in_u;
  uint32_t in_cap;
  uint32_t in_n; /* electrons in the input (mode 1) */
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
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long wave_ballot(bool p) { return __ballot(p); }

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

/* BackscatterEvent -> BackscatterStats::actionPerformed (BackscatterStats.cu:147-170)
 * at the chamber wall point reached from (px,py,pz) along (dx,dy,dz). */
__device__ __forceinline__ void tally_escape(const DevScene& S, const JobDev& J, const Elec& e, float px, float py, float pz,
                             float dx, float dy, float dz, bool fromRoot) {
  double x = px, y = py, z = pz;
  if (fromRoot) {
    double cx = x - (double)S.chamber_c[0], cy = y - (double)S.chamber_c[1], cz = z - (double)S.chamber_c[2];
    double ddx = dx, ddy = dy, ddz = dz;
    double dn = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    ddx /= dn;
    ddy /= dn;
    ddz /= dn;
    double R = (double)S.chamber_r;
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
  atomicAdd(&J.hist_elev[hist_bin(elevation, 0.0, PI, 180, compat) + 1], 1ull);
  atomicAdd(&J.hist_azim[hist_bin(azimuth, 0.0, 2.0 * PI, 360, compat) + 1], 1ull);
  const double kEeV = (double)e.E;
  int bin = hist_bin(kEeV, 0.0, J.E0eV, J.nbins, compat);
  uint32_t rel = e.primary - J.prim_base;
  uint32_t pix = (J.chunk_start + rel) / (uint32_t)J.ppp;
  if (elevation < (PI / 2.0)) {
    atomicAdd(&J.hist_fwd[bin + 1], 1ull);
    atomicAdd(&J.fwd_px[pix], 1ull);
  } else {
    atomicAdd(&J.hist_back[bin + 1], 1ull);
    atomicAdd(&J.back_energy_px[pix], kEeV);
    bool se = bin >= 0 && bin < J.maxSEbin;
    atomicAdd(&J.prim_counts[rel], se ? 1u : 0x10000u);
  }
}

enum StepResult { STEP_CONTINUE = 0, STEP_DEAD = 1, STEP_ESCAPED = 2, STEP_WATCHDOG = 3 };

/* ExpQMBarrierSM.cu:98-219: the crossing from the electron's region into
 * `next`, given the outward barrier normal (nbx, nby, nbz) resolved from the
 * struck shape.  Moves the electron a nudge along +nb (returns +1) or -nb
 * (returns -1); u1 decides a quantum-mechanical transmission. */
__device__ __forceinline__ int barrier_core(const DevScene& S, const DMSM& msm, const DMSM& nmsm, Elec& e, int next,
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
    rootPerpE = __builtin_sqrtf(perpE);
    rootDiff = __builtin_sqrtf(perpE - deltaU);
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

/* One MonteCarloSS::takeStep (MonteCarloSS.cu:107-163). */
__device__ __forceinline__ int take_step(const DevScene& S, const JobDev& J, Elec& e, Spawn& sp) {
  sp.has = false;
  int region = e.region;
  if (region < 0 || !shape_contains(S, S.regions[region].shape, e.px, e.py, e.pz)) {
    region = containing_sub_region(S, 0, e.px, e.py, e.pz);
    e.region = region;
    if (region < 0) return STEP_DEAD;
  }
  const DMSM& msm = S.msms[S.regions[region].msm];
  float u[kStepUniforms];
  step_uniforms(e, J.k0, J.k1, u);
  /* MONSEL_MaterialScatterModel::randomMeanPathLength (:64-91); each
   * mechanism's rate is kept in a per-lane LDS slot for the scatter choice */
  __shared__ float l_rate[kMaxMech][kLaneBlock];
  const int tid = threadIdx.x;
  float total = 0.f;
  float freePath;
  const float maxFreePath = 2.f * S.chamber_r;
  if (msm.is_null) {
    freePath = 1.0e9f; /* NullMaterialScatterModel.cu:32-35: 1 m */
  } else {
    for (int k = 0; k < msm.n_mech; ++k) {
      const float rk = mech_rate(S, S.mechs[msm.mech0 + k], e.E);
      l_rate[k][tid] = rk;
      total += rk;
    }
    if (total != 0.f) {
      float fp = -__logf(u[0]) / total;
      freePath = fp > maxFreePath ? maxFreePath : fp;
    } else {
      freePath = maxFreePath;
    }
  }
  const float px = e.px, py = e.py, pz = e.pz;
  StepEnd se = find_end_of_step(S, region, px, py, pz, e.dx, e.dy, e.dz, freePath);
  float len = se.t;
  e.px = px + len * e.dx;
  e.py = py + len * e.dy;
  e.pz = pz + len * e.dz;
  float prevE = e.E;
  if (!msm.is_null) e.E += csd_compute(msm.csd, len, e.E);
  e.steps++;
  if ((int)e.steps > J.max_steps || !(e.E == e.E)) return STEP_WATCHDOG;
  if (e.E < msm.minE) return STEP_DEAD;
  if (se.next == region) {
    /* MONSEL_MaterialScatterModel::scatter (:93-134), rates at prevE */
    if (msm.is_null || total == 0.f) return STEP_CONTINUE;
    const float r = u[1] * total;
    /* first mechanism whose cumulative rate reaches r (MONSEL_MaterialScatterModel.cu:129-131);
     * the running sum repeats the order of `total` */
    int idx = 0;
    float rate = l_rate[0][tid];
    float cum = rate;
    while (idx < msm.n_mech - 1 && cum < r) {
      ++idx;
      rate = l_rate[idx][tid];
      cum += rate;
    }
    bool trapped = false;
    sp = mech_scatter(S, S.mechs[msm.mech0 + idx], e, prevE, rate, trapped, u);
    if (sp.has) {
      /* trackSecondaryElectron (MonteCarloSS.cu:165-174) */
      if (!(sp.E > msm.minE)) sp.has = false;
    }
    if (trapped || e.E < msm.minE) return STEP_DEAD;
    return STEP_CONTINUE;
  }
  if (se.next >= 0) {
    /* ExpQMBarrierSM::barrierScatter (ExpQMBarrierSM.cu:50-219) */
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
            ShapeHit h = shape_first(S, S.regions[struck].shape, px, py, pz, e.dx, e.dy, e.dz, freePath);
            nbx = -h.nx; nby = -h.ny; nbz = -h.nz;
          }
        } else {
          nbx = se.ownNx; nby = se.ownNy; nbz = se.ownNz;
        }
      }
      barrier_core(S, msm, S.msms[S.regions[next].msm], e, next, nbx, nby, nbz, u[1]);
    }
    /* MonteCarloSS.cu:146-148: nudge along the new direction; the region is
     * re-located by the check that opens the next step */
    float dd = nudge_nm(e.px, e.py, e.pz);
    e.px += dd * e.dx;
    e.py += dd * e.dy;
    e.pz += dd * e.dz;
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

__device__ __forceinline__ void make_primary(const DevScene& S, const JobDev& J, uint32_t idx, Elec& e) {
  uint32_t job_idx = J.chunk_start + idx;
  uint32_t pix = job_idx / (uint32_t)J.ppp;
  e.primary = J.prim_base + idx;
  e.lin_lo = 0;
  e.lin_hi = 0;
  e.nchild = 0;
  e.steps = 0;
  /* GaussianBeam::createElectron (GaussianBeam.cu:52-65) */
  uint32_t a0, a1, a2, a3;
  philox4x32_10(kPrimaryBlock, e.lin_lo, e.lin_hi, e.primary, J.k0, J.k1, a0, a1, a2, a3);
  const float u1 = u01(a0), u2 = u01(a1);
  float r = __builtin_sqrtf(-2.f * __logf(u1)) * J.width;
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
  e.region = containing_sub_region(S, 0, e.px, e.py, e.pz);
}

/* waves per SIMD the register allocation of k_transport targets */
#ifndef MONSEL_TRANSPORT_WAVES_PER_EU
#define MONSEL_TRANSPORT_WAVES_PER_EU 4
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MONSEL_TRANSPORT_WAVES_PER_EU)))
k_transport(DevScene S, JobDev J, GenDev G) {
  Elec e;
  bool alive = false, done = false;
  uint32_t nsteps = 0, nelec = 0, nesc = 0, nwd = 0, nover = 0;
  const uint32_t in_count = G.mode == 0 ? J.chunk_n : G.in_n;
  const unsigned long long lanemask_lt = (1ull << lane_id()) - 1ull;
  for (;;) {
    /* claim work for idle lanes: one atomic per wave */
    bool need = !alive && !done;
    unsigned long long m = wave_ballot(need);
    if (m) {
      int leader = __ffsll((long long)m) - 1;
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(G.claim, (unsigned int)__popcll(m));
      base = __shfl(base, leader);
      if (need) {
        uint32_t idx = base + (uint32_t)__popcll(m & lanemask_lt);
        if (idx < in_count) {
          if (G.mode == 0) {
            make_primary(S, J, idx, e);
          } else {
            const uint32_t cap = G.in_cap;
            e.px = G.in_f[0 * cap + idx];
            e.py = G.in_f[1 * cap + idx];
            e.pz = G.in_f[2 * cap + idx];
            e.dx = G.in_f[3 * cap + idx];
            e.dy = G.in_f[4 * cap + idx];
            e.dz = G.in_f[5 * cap + idx];
            e.E = G.in_f[6 * cap + idx];
            e.region = (int)G.in_u[0 * cap + idx];
            e.primary = G.in_u[1 * cap + idx];
            e.lin_lo = G.in_u[2 * cap + idx];
            e.lin_hi = G.in_u[3 * cap + idx];
            e.nchild = 0;
            e.steps = 0;
          }
          alive = true;
          nelec++;
        } else {
          done = true;
        }
      }
    }
    if (wave_ballot(alive) == 0ull) break;
    Spawn sp;
    sp.has = false;
    int res = STEP_CONTINUE;
    if (alive) {
      res = take_step(S, J, e, sp);
      nsteps++;
    }
    /* append secondaries: one atomic per wave */
    unsigned long long sm = wave_ballot(sp.has);
    if (sm) {
      int leader = __ffsll((long long)sm) - 1;
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(G.out_count, (unsigned int)__popcll(sm));
      base = __shfl(base, leader);
      if (sp.has) {
        uint32_t slot = base + (uint32_t)__popcll(sm & lanemask_lt);
        uint64_t lin = child_lineage(((uint64_t)e.lin_hi << 32) | e.lin_lo, e.nchild++);
        if (slot < G.out_cap) {
          const uint32_t cap = G.out_cap;
          G.out_f[0 * cap + slot] = e.px;
          G.out_f[1 * cap + slot] = e.py;
          G.out_f[2 * cap + slot] = e.pz;
          G.out_f[3 * cap + slot] = sp.dx;
          G.out_f[4 * cap + slot] = sp.dy;
          G.out_f[5 * cap + slot] = sp.dz;
          G.out_f[6 * cap + slot] = sp.E;
          G.out_u[0 * cap + slot] = (uint32_t)e.region;
          G.out_u[1 * cap + slot] = e.primary;
          G.out_u[2 * cap + slot] = (uint32_t)lin;
          G.out_u[3 * cap + slot] = (uint32_t)(lin >> 32);
        } else {
          nover++;
        }
      }
    }
    /* append escape records: one atomic per wave */
    const bool esc = alive && res == STEP_ESCAPED;
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
          G.esc_u[1 * cap + slot] = (uint32_t)e.region;
        } else {
          nover++;
        }
      }
    }
    if (alive && res != STEP_CONTINUE) {
      alive = false;
      if (res == STEP_ESCAPED) nesc++;
      if (res == STEP_WATCHDOG) nwd++;
    }
  }
  /* flush counters: lane sums reduced across the wave, one atomic each */
  unsigned long long wsteps = nsteps, welec = nelec, wesc = nesc, wwd = nwd, wover = nover;
  for (int off = 32; off > 0; off >>= 1) {
    wsteps += __shfl_down(wsteps, off);
    welec += __shfl_down(welec, off);
    wesc += __shfl_down(wesc, off);
    wwd += __shfl_down(wwd, off);
    wover += __shfl_down(wover, off);
  }
  if (lane_id() == 0) {
    atomicAdd(&J.stats[kStatSteps], wsteps);
    atomicAdd(&J.stats[kStatElectrons], welec);
    if (wesc) atomicAdd(&J.stats[kStatEscapes], wesc);
    if (wwd) atomicAdd(&J.stats[kStatWatchdog], wwd);
    if (wover) atomicAdd(&J.stats[kStatOverflow], wover);
  }
}

/* Tallies of the escape records of one transport launch (grid-stride over
 * the record count the launch left in *count). */
__global__ void __launch_bounds__(256) k_tally_escapes(DevScene S, JobDev J, const float* ef, const uint32_t* eu,
                                                       uint32_t cap, const unsigned int* count) {
  const uint32_t n = min(*count, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    Elec e;
    e.E = ef[6 * cap + i];
    e.primary = eu[i];
    tally_escape(S, J, e, ef[i], ef[cap + i], ef[2 * cap + i], ef[3 * cap + i], ef[4 * cap + i], ef[5 * cap + i],
                 eu[cap + i] != 0u);
  }
}

/* Per-primary SE/BSE counts -> per-pixel sums of n and n^2 (exact u64).
 * Primaries of one pixel are consecutive, so a wave whose 64 primaries share
 * a pixel reduces in registers and issues one atomic per sum; a wave that
 * straddles pixels falls back to per-lane atomics. */
__global__ void k_reduce_primaries(JobDev J, unsigned long long* back_px, unsigned long long* se_px,
                                   unsigned long long* se_sq, unsigned long long* bse_sq) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < J.chunk_n;
  uint32_t c = valid ? J.prim_counts[i] : 0u;
  uint32_t iv = valid ? i : J.chunk_n - 1u;
  unsigned long long nse = c & 0xFFFFu, nbse = c >> 16;
  uint32_t pix = (J.chunk_start + iv) / (uint32_t)J.ppp;
  unsigned long long vb = nse + nbse, vs = nse, vss = nse * nse, vbb = nbse * nbse;
  const uint32_t pix0 = __shfl(pix, 0);
  if (wave_ballot(pix != pix0) == 0ull) {
    for (int off = 32; off > 0; off >>= 1) {
      vb += __shfl_xor(vb, off);
      vs += __shfl_xor(vs, off);
      vss += __shfl_xor(vss, off);
      vbb += __shfl_xor(vbb, off);
    }
    if (lane_id() == 0) {
      if (vb) atomicAdd(&back_px[pix0], vb);
      if (vs) atomicAdd(&se_px[pix0], vs);
      if (vss) atomicAdd(&se_sq[pix0], vss);
      if (vbb) atomicAdd(&bse_sq[pix0], vbb);
    }
  } else if (vb) {
    atomicAdd(&back_px[pix], vb);
    if (vs) atomicAdd(&se_px[pix], vs);
    if (vss) atomicAdd(&se_sq[pix], vss);
    if (vbb) atomicAdd(&bse_sq[pix], vbb);
  }
}

/* ---- probe kernels ---------------------------------------------------- */

__global__ void k_probe_rates(DevScene S, int msm, int n, const float* E, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DMSM& m = S.msms[msm];
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
  const DMech& m = S.mechs[mech];
  const DElem& el = S.elems[m.elem0 + element];
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
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DMech& m = S.mechs[mech];
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

__global__ void k_probe_locate(DevScene S, int n, const float* p, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = containing_sub_region(S, 0, p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__global__ void k_probe_nu(DevScene S, int table, int order, int n, const float* x, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int dim = S.nu[table].dim;
  out[i] = nu_interp(S, table, order, x + (size_t)i * dim, (S.compat & MONSEL_COMPAT_D3_ABSCISSAE) != 0);
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
  ~DBuf() {
    if (p) (void)hipFree(p);
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
  uint32_t compat = 0;
  float chamber_r = 1e8f;
  float chamber_c[3] = {0, 0, 0};
};

struct SceneDev {
  DBuf blob;
  DevScene dev{};
  size_t bytes = 0;
  monsel_scene_desc copyOfCounts{};
};

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
      for (int j = 0; j < t.n[k]; ++j) H.nudata.push_back((float)t.x[k][j]);
      tot *= (size_t)t.n[k];
    }
    dt.voff = (int)H.nudata.size();
    double r0 = INFINITY, r1 = -INFINITY;
    for (size_t j = 0; j < tot; ++j) {
      double v = t.values[j];
      H.nudata.push_back((float)v);
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
        if (sh.n_planes < 1 || sh.n_planes > kMaxPlanes || sh.first_plane < 0 || sh.first_plane + sh.n_planes > d->n_planes)
          throw MonselError(MONSEL_E_INVALID, "plane range");
        p.kind = PRIM_PLANES;
        p.first_plane = (int)H.planes.size();
        p.n_planes = sh.n_planes;
        for (int k = 0; k < sh.n_planes; ++k) {
          const monsel_plane_desc& pl = d->planes[sh.first_plane + k];
          double n[3] = {pl.normal[0], pl.normal[1], pl.normal[2]};
          double b = (n[0] * pl.point[0] + n[1] * pl.point[1] + n[2] * pl.point[2]) * 1e9;
          H.planes.push_back(make_float4((float)n[0], (float)n[1], (float)n[2], (float)b));
        }
      } else if (sh.kind == MONSEL_SHAPE_CYLINDER) {
        p.kind = PRIM_CYLINDER;
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
  return total;
}

void upload_scene(const HostScene& H, SceneDev& sd) {
  std::vector<char> blob;
  size_t o_reg = append(blob, H.regions), o_ch = append(blob, H.children), o_sh = append(blob, H.shapes);
  size_t o_pr = append(blob, H.prims), o_pl = append(blob, H.planes), o_msm = append(blob, H.msms);
  size_t o_me = append(blob, H.mechs), o_el = append(blob, H.elems), o_sp = append(blob, H.spwem);
  size_t o_x1 = append(blob, H.x1), o_nu = append(blob, H.nu), o_nd = append(blob, H.nudata);
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
  S.compat = H.compat;
  S.chamber_r = H.chamber_r;
  for (int k = 0; k < 3; ++k) S.chamber_c[k] = H.chamber_c[k];
  S.regions = (const DRegion*)(b + o_reg);
  S.children = (const int*)(b + o_ch);
  S.shapes = (const DShape*)(b + o_sh);
  S.prims = (const DPrim*)(b + o_pr);
  S.planes = (const float4*)(b + o_pl);
  S.msms = (const DMSM*)(b + o_msm);
  S.mechs = (const DMech*)(b + o_me);
  S.elems = (const DElem*)(b + o_el);
  S.mott_spwem = (const float*)(b + o_sp);
  S.mott_x1 = (const float*)(b + o_x1);
  S.nu = (const DNUTable*)(b + o_nu);
  S.nu_data = (const float*)(b + o_nd);
  sd.bytes = blob.size();
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
  std::chrono::steady_clock::time_point t_submit;
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
  /* job workspace (worker thread only) */
  DBuf q[2], esc, qcount, prim, pix, hist, stats, centers;
  unsigned int* host_count = nullptr; /* pinned */
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

void run_job(monsel_ctx_s* C, Job& job, SceneDev& sd) {
  const monsel_job_desc& jd = job.desc;
  const uint64_t total = (uint64_t)jd.n_pixels * (uint64_t)jd.primaries_per_pixel;
  const int np = jd.n_pixels;
  const int nbins = jd.n_energy_bins > 0 ? jd.n_energy_bins : 1;
  const double E0eV = jd.beam_energy / kE;
  /* pixel tallies + histograms */
  C->pix.ensure((size_t)np * 6 * 8);
  unsigned long long* back_px = C->pix.as<unsigned long long>();
  unsigned long long* fwd_px = back_px + np;
  unsigned long long* se_px = fwd_px + np;
  unsigned long long* se_sq = se_px + np;
  unsigned long long* bse_sq = se_sq + np;
  double* benergy = (double*)(bse_sq + np);
  size_t hn = (size_t)2 * (nbins + 2) + 182 + 362;
  C->hist.ensure(hn * 8);
  unsigned long long* hback = C->hist.as<unsigned long long>();
  unsigned long long* hfwd = hback + nbins + 2;
  unsigned long long* helev = hfwd + nbins + 2;
  unsigned long long* hazim = helev + 182;
  C->stats.ensure(16 * 8);
  C->qcount.ensure(4 * 256);
  C->centers.ensure(sizeof(double) * 3 * (size_t)np);
  HIP_TRY(hipMemcpyAsync(C->centers.p, job.centers.data(), sizeof(double) * 3 * (size_t)np, hipMemcpyHostToDevice,
                         C->stream));
  HIP_TRY(hipMemsetAsync(C->pix.p, 0, (size_t)np * 6 * 8, C->stream));
  HIP_TRY(hipMemsetAsync(C->hist.p, 0, hn * 8, C->stream));
  HIP_TRY(hipMemsetAsync(C->stats.p, 0, 16 * 8, C->stream));
  /* chunking: queue memory per chunk bounded */
  const uint32_t K = secondaries_per_primary(E0eV);
  const size_t perEntry = (size_t)(kQF + kQU) * 4;
  const size_t perEsc = (size_t)(kEF + kEU) * 4;
  const size_t budget = (size_t)6 << 30; /* both generation buffers + escape records */
  uint64_t chunkMax = budget / ((2 * perEntry + perEsc) * K);
  if (C->tuning.queue_capacity > 0) chunkMax = std::max<uint64_t>(1, (uint64_t)C->tuning.queue_capacity / K);
  if (chunkMax < 1024) chunkMax = 1024;
  int block = C->tuning.block_size > 0 ? C->tuning.block_size : 256;
  int grid = C->n_cu * (C->blocks_per_cu > 0 ? C->blocks_per_cu : 8);
  int maxGen = C->tuning.max_generations > 0 ? C->tuning.max_generations : 64;
  uint32_t k0 = (uint32_t)jd.seed, k1 = (uint32_t)(jd.seed >> 32);
  /* Event markers time the kernels; MONSEL_NO_EVENT_TIMING=1 leaves them out
   * (hardware-counter profiling sessions reject the marker packets). */
  const bool timing = !getenv("MONSEL_NO_EVENT_TIMING");
  if (timing) HIP_TRY(hipEventRecord(C->ev0, C->stream));
  int generations = 0;
  for (uint64_t start = 0; start < total; start += chunkMax) {
    uint32_t n = (uint32_t)std::min<uint64_t>(chunkMax, total - start);
    uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)n * K, 0xFFFFFFF0ull);
    C->q[0].ensure((size_t)cap * perEntry);
    C->q[1].ensure((size_t)cap * perEntry);
    C->esc.ensure((size_t)cap * perEsc);
    C->prim.ensure((size_t)n * 4);
    HIP_TRY(hipMemsetAsync(C->prim.p, 0, (size_t)n * 4, C->stream));
    JobDev J;
    memset(&J, 0, sizeof(J));
    J.k0 = k0;
    J.k1 = k1;
    J.E0 = (float)E0eV;
    J.width = (float)(jd.beam_width * 1e9);
    J.centers = nullptr;
    J.n_pixels = np;
    J.ppp = jd.primaries_per_pixel;
    J.prim_base = (uint32_t)(jd.primary_offset + start);
    J.chunk_start = (uint32_t)start;
    J.chunk_n = n;
    J.nbins = nbins;
    J.E0eV = E0eV;
    J.maxSEbin = (int)(jd.se_cutoff_ev / (E0eV / nbins));
    J.max_steps = jd.max_steps > 0 ? jd.max_steps : 1000000;
    J.prim_counts = C->prim.as<uint32_t>();
    J.fwd_px = fwd_px;
    J.back_energy_px = benergy;
    J.hist_back = hback;
    J.hist_fwd = hfwd;
    J.hist_elev = helev;
    J.hist_azim = hazim;
    J.stats = C->stats.as<unsigned long long>();
    J.centers = C->centers.as<double>();
    unsigned int* counts = C->qcount.as<unsigned int>();
    /* counts[0]: claim counter, counts[1]: spawned secondaries, counts[2]: escapes */
    uint32_t in_count = n;
    int cur = 0;
    for (int g = 0; g < maxGen && in_count > 0; ++g) {
      HIP_TRY(hipMemsetAsync(counts, 0, 3 * sizeof(unsigned int), C->stream));
      GenDev G;
      memset(&G, 0, sizeof(G));
      G.mode = g == 0 ? 0 : 1;
      G.in_f = C->q[cur].as<float>();
      G.in_u = (const uint32_t*)(C->q[cur].as<float>() + (size_t)kQF * cap);
      G.in_cap = cap;
      G.in_n = in_count;
      G.claim = counts;
      G.out_f = C->q[cur ^ 1].as<float>();
      G.out_u = (uint32_t*)(C->q[cur ^ 1].as<float>() + (size_t)kQF * cap);
      G.out_cap = cap;
      G.out_count = counts + 1;
      G.esc_f = C->esc.as<float>();
      G.esc_u = (uint32_t*)(C->esc.as<float>() + (size_t)kEF * cap);
      G.esc_cap = cap;
      G.esc_count = counts + 2;
      if (timing) HIP_TRY(hipEventRecord(C->evk0, C->stream));
      g_hip_call = "launch k_transport";
      hipLaunchKernelGGL(k_transport, dim3(grid), dim3(block), 0, C->stream, sd.dev, J, G);
      HIP_TRY(hipGetLastError());
      if (timing) HIP_TRY(hipEventRecord(C->evk1, C->stream));
      HIP_TRY(hipMemcpyAsync(C->host_count, counts + 1, sizeof(unsigned int), hipMemcpyDeviceToHost, C->stream));
      g_hip_call = "launch k_tally_escapes";
      hipLaunchKernelGGL(k_tally_escapes, dim3(C->n_cu * 4), dim3(256), 0, C->stream, sd.dev, J, G.esc_f, G.esc_u,
                         cap, (const unsigned int*)(counts + 2));
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(C->stream));
      uint32_t outc = *C->host_count;
      in_count = outc > cap ? cap : outc;
      float kms = 0.f;
      if (timing) HIP_TRY(hipEventElapsedTime(&kms, C->evk0, C->evk1));
      job.stats.transport_ms += kms;
      job.stats.transport_launches += 1;
      cur ^= 1;
      generations = std::max(generations, g + 1);
    }
    /* generation limit: any electrons left are dropped and counted */
    if (in_count > 0) job.stats.queue_overflow += in_count;
    int rb = 256;
    g_hip_call = "launch k_reduce_primaries";
    hipLaunchKernelGGL(k_reduce_primaries, dim3((n + rb - 1) / rb), dim3(rb), 0, C->stream, J, back_px, se_px, se_sq,
                       bse_sq);
    HIP_TRY(hipGetLastError());
  }
  if (timing) HIP_TRY(hipEventRecord(C->ev1, C->stream));
  /* results */
  job.back.resize(np);
  job.fwd.resize(np);
  job.se.resize(np);
  job.se_sq.resize(np);
  job.bse_sq.resize(np);
  job.back_energy.resize(np);
  job.hback.resize(nbins + 2);
  job.hfwd.resize(nbins + 2);
  job.helev.resize(182);
  job.hazim.resize(362);
  std::vector<unsigned long long> st(16), sq(2 * (size_t)np);
  HIP_TRY(hipMemcpyAsync(job.back.data(), back_px, 8 * np, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.fwd.data(), fwd_px, 8 * np, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.se.data(), se_px, 8 * np, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(sq.data(), se_sq, 16 * (size_t)np, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.back_energy.data(), benergy, 8 * np, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.hback.data(), hback, 8 * (nbins + 2), hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.hfwd.data(), hfwd, 8 * (nbins + 2), hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.helev.data(), helev, 8 * 182, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(job.hazim.data(), hazim, 8 * 362, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipMemcpyAsync(st.data(), C->stats.p, 8 * 16, hipMemcpyDeviceToHost, C->stream));
  HIP_TRY(hipStreamSynchronize(C->stream));
  for (size_t p = 0; p < (size_t)np; ++p) {
    job.se_sq[p] = (double)sq[p];
    job.bse_sq[p] = (double)sq[np + p];
  }
  float ms = 0.f;
  if (timing) HIP_TRY(hipEventElapsedTime(&ms, C->ev0, C->ev1));
  job.stats.primaries = total;
  job.stats.steps = st[kStatSteps];
  job.stats.electrons = st[kStatElectrons];
  job.stats.escapes = st[kStatEscapes];
  job.stats.watchdog_drops = st[kStatWatchdog];
  job.stats.queue_overflow += st[kStatOverflow];
  job.stats.generations = generations;
  job.stats.device_ms = ms;
}

void worker_main(monsel_ctx_s* C) {
  (void)hipSetDevice(C->device);
  for (;;) {
    std::shared_ptr<Job> job;
    std::shared_ptr<SceneDev> sd;
    {
      std::unique_lock<std::mutex> lk(C->mu);
      C->cv_work.wait(lk, [&] { return C->stop || !C->queue.empty(); });
      if (C->stop && C->queue.empty()) return;
      job = C->queue.front();
      C->queue.pop_front();
      auto it = C->scenes.find(job->scene);
      if (it != C->scenes.end()) sd = it->second;
    }
    int st = MONSEL_OK;
    std::string msg;
    if (!sd) {
      st = MONSEL_E_NOTFOUND;
      msg = "scene released before the job ran";
    } else {
      try {
        run_job(C, *job, *sd);
      } catch (const MonselError& e) {
        st = e.status;
        msg = e.msg;
      } catch (const std::exception& e) {
        st = MONSEL_E_DEVICE;
        msg = std::string(e.what()) + " (in " + g_hip_call + ")";
      }
    }
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
  } catch (const MonselError& e) {
    return set_err(e.status, e.msg);
  } catch (const std::exception& e) {
    return set_err(MONSEL_E_DEVICE, std::string(e.what()) + " (in " + g_hip_call + ")");
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
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_ordinal));
    C->n_cu = prop.multiProcessorCount;
    C->hbm = prop.totalGlobalMem;
    C->name = prop.gcnArchName;
    HIP_TRY(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void**)&C->host_count, 64, hipHostMallocDefault));
    HIP_TRY(hipEventCreate(&C->ev0));
    HIP_TRY(hipEventCreate(&C->ev1));
    HIP_TRY(hipEventCreate(&C->evk0));
    HIP_TRY(hipEventCreate(&C->evk1));
    int block = C->tuning.block_size > 0 ? C->tuning.block_size : 256;
    if (block % 64 != 0 || block > kLaneBlock)
      throw MonselError(MONSEL_E_INVALID, "block_size must be a multiple of 64, <= 256");
    if (C->tuning.blocks_per_cu > 0) C->blocks_per_cu = C->tuning.blocks_per_cu;
    else {
      int occ = 0;
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_transport, block, 0));
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
  if (ctx->worker.joinable()) ctx->worker.join();
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->scenes.clear();
  ctx->jobs.clear();
  if (ctx->host_count) (void)hipHostFree(ctx->host_count);
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
  uint64_t total = (uint64_t)job->n_pixels * (uint64_t)job->primaries_per_pixel;
  uint32_t K = secondaries_per_primary(job->beam_energy / kE);
  size_t perEntry = (size_t)(kQF + kQU) * 4;
  uint64_t chunkMax = ((size_t)6 << 30) / (2 * perEntry * K);
  uint64_t chunk = std::min<uint64_t>(total, chunkMax);
  *hbm_bytes = (size_t)(2 * chunk * K * perEntry + chunk * 4 + (size_t)job->n_pixels * 72 +
                        (size_t)(2 * (job->n_energy_bins + 2) + 544) * 8);
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
  return MONSEL_OK;
}

int monsel_job_submit(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, monsel_job_id* out) {
  if (!ctx || !out) return set_err(MONSEL_E_INVALID, "NULL argument");
  int v = validate_job(job);
  if (v) return v;
  if (!find_scene(ctx, scene)) return set_err(MONSEL_E_NOTFOUND, "unknown scene");
  auto j = std::make_shared<Job>();
  j->scene = scene;
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

int monsel_job_release(monsel_ctx ctx, monsel_job_id job) {
  if (!ctx) return set_err(MONSEL_E_INVALID, "NULL context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->jobs.find(job);
  if (it == ctx->jobs.end()) return set_err(MONSEL_E_NOTFOUND, "unknown job");
  if (!it->second->done) return set_err(MONSEL_E_NOTREADY, "job still running");
  ctx->jobs.erase(it);
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

}  /* extern "C" */
