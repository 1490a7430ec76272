/*
 * phase_timing.h -- optional per-phase wave timing of the transport kernel.
 *
 * Built with -DMONSEL_PHASE_TIMING (python -m cudamonsel_amd.build
 * --phase-timing), every wave of k_transport keeps, for each phase of the
 * step below, the core-clock cycles it spent there, how many times it entered
 * it and how many lanes were active on entry, in its own LDS slot; at exit the
 * slots are added to a device array that monsel_phase_timing_read() copies
 * out (and clears).  Lanes active / (entries * 64) is the phase's lane
 * utilisation; cycles include the time the wave waited behind other waves of
 * its SIMD, so they apportion the kernel's wall time rather than its issue
 * slots.  Without the macro every hook is empty and nothing is exported.
 */
#ifndef MONSEL_PHASE_TIMING_H
#define MONSEL_PHASE_TIMING_H

#include <hip/hip_runtime.h>

namespace monsel_dev {

enum Phase {
  PH_CLAIM = 0,    /* claiming work: primary creation or queue load */
  PH_CONTAIN,      /* containment check and relocation opening a step */
  PH_RATES,        /* step uniforms, mechanism rates, free path */
  PH_GEOM,         /* findEndOfStep */
  PH_CSD,          /* continuous slowing down */
  PH_SELECT,       /* mechanism choice and the scatter uniforms */
  PH_ELASTIC,      /* elastic scatter (Mott / Browning / Rutherford) */
  PH_FITTED,       /* fitted inelastic (secondary generation) */
  PH_POLARON,      /* polaron trap */
  PH_PHONON,       /* phonon scatter */
  PH_TABULATED,    /* tabulated inelastic */
  PH_BARRIER,      /* boundary crossing (barrier normal, transmission) */
  PH_STEP,         /* the whole step */
  PH_TAIL,         /* secondaries, regroup, escape records after the step */
  PH_RATES_UNI,    /* step starts whose lanes share one region (entries only) */
  PH_GEOM_UNI,     /* findEndOfStep passes whose lanes share one region (entries only) */
  PH_SAFETY,       /* the safety bound at a step's end point */
  PH_UNIFORMS,     /* the step's first four uniforms (within PH_RATES) */
  PH_RATE_LOOP,    /* the mechanism rates (within PH_RATES) */
  PH_GEOM_IV,      /* primitive entry/exit intervals (within PH_GEOM) */
  PH_GEOM_WALK,    /* crossing walks of the shapes (within PH_GEOM) */
  PH_GEOM_LOC,     /* locating the region beyond a crossing (within PH_GEOM) */
  kPhases
};

}  // namespace monsel_dev

#ifdef MONSEL_PHASE_TIMING

namespace monsel_dev {

constexpr int kPhaseWaves = 4;  /* waves per workgroup (blockDim.x <= 256) */
__device__ unsigned long long g_phase[kPhases * 3];
/* electrons ending after 2^k .. 2^(k+1)-1 steps in one launch (k < 24) */
constexpr int kStepBuckets = 24;
__device__ unsigned long long g_step_hist[kStepBuckets];
/* (start, end) of every wave of the launches since the last read, in the
 * constant-rate wall clock (wall_clock64, 100 MHz) */
constexpr int kMaxSpans = 1 << 18;
__device__ unsigned int g_span_count;
__device__ unsigned long long g_spans[2 * kMaxSpans];

__device__ __forceinline__ unsigned long long* pt_slot() {
  __shared__ unsigned long long acc[kPhaseWaves * kPhases * 3];
  return acc + (threadIdx.x >> 6) * (kPhases * 3);
}

__device__ __forceinline__ unsigned long long pt_now() { return clock64(); }

/* first active lane of the wave updates the wave's slot */
__device__ __forceinline__ void pt_add(int ph, unsigned long long t0) {
  const unsigned long long t1 = pt_now();
  const unsigned long long m = __ballot(1);
  if ((int)__lane_id() == __ffsll((long long)m) - 1) {
    unsigned long long* s = pt_slot() + ph * 3;
    s[0] += t1 - t0;
    s[1] += 1ull;
    s[2] += (unsigned long long)__popcll(m);
  }
}

/* counts an entry of `ph` when every active lane has the same `key` */
__device__ __forceinline__ void pt_uniform(int ph, int key) {
  const int k0 = __builtin_amdgcn_readfirstlane(key);
  const unsigned long long m = __ballot(1);
  if (__ballot(key == k0) == m && (int)__lane_id() == __ffsll((long long)m) - 1) {
    unsigned long long* s = pt_slot() + ph * 3;
    s[1] += 1ull;
    s[2] += (unsigned long long)__popcll(m);
  }
}

__device__ __forceinline__ void pt_end_electron(unsigned int steps) {
  const int k = steps ? min(31 - __clz(steps), kStepBuckets - 1) : 0;
  atomicAdd(&g_step_hist[k], 1ull);
}

__device__ __forceinline__ void pt_init() {
  if ((threadIdx.x & 63) == 0) {
    unsigned long long* s = pt_slot();
    for (int k = 0; k < kPhases * 3; ++k) s[k] = 0ull;
  }
}

__device__ __forceinline__ void pt_flush(unsigned long long start) {
  if ((threadIdx.x & 63) == 0) {
    const unsigned int k = atomicAdd(&g_span_count, 1u);
    if (k < (unsigned int)kMaxSpans) {
      g_spans[2 * k] = start;
      g_spans[2 * k + 1] = wall_clock64();
    }
    unsigned long long* s = pt_slot();
    for (int k = 0; k < kPhases * 3; ++k)
      if (s[k]) atomicAdd(&g_phase[k], s[k]);
  }
}

}  // namespace monsel_dev

#define PT_BEGIN(v) const unsigned long long v = ::monsel_dev::pt_now()
#define PT_END(ph, v) ::monsel_dev::pt_add((ph), (v))
#define PT_INIT() ::monsel_dev::pt_init(); const unsigned long long pt_start = wall_clock64()
#define PT_FLUSH() ::monsel_dev::pt_flush(pt_start)
#define PT_END_ELECTRON(steps) ::monsel_dev::pt_end_electron(steps)
#define PT_UNIFORM(ph, key) ::monsel_dev::pt_uniform((ph), (key))

#else

#define PT_BEGIN(v)
#define PT_END(ph, v)
#define PT_INIT()
#define PT_FLUSH()
#define PT_END_ELECTRON(steps)
#define PT_UNIFORM(ph, key)

#endif

#endif
