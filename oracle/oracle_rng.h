/*
 * TEST INFRASTRUCTURE -- part of the CPU oracle.  Only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may use anything under oracle/.
 *
 * Counter-based random streams for the oracle: Philox4x32-10 (Salmon, Moraes,
 * Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), written
 * from the published round function.  The reference draws from libc rand()
 * (gov/nist/microanalysis/Utility/Math2.cu:1328-1336); a keyed stream per
 * electron replaces that single global stream so that results do not depend
 * on scheduling.  Pinned by the Random123 known-answer vectors (tests).
 *
 * Stream layout (shared specification with the GPU engine):
 *   key     = (seed & 0xffffffff, seed >> 32)
 *   counter = (block, lineage & 0xffffffff, lineage >> 32, primary)
 *   block b yields uniforms 4b .. 4b+3;  u = (x >> 9) * 2^-23 + 2^-24,
 *   exactly representable in FP32 and in (0, 1).
 *   primary electron: lineage 0;  k-th secondary (k = 0, 1, ...) of an
 *   electron with lineage L: splitmix64(L ^ (0x9E3779B97F4A7C15 * (k + 1))).
 */
#ifndef ORACLE_RNG_H
#define ORACLE_RNG_H

#include <stdint.h>

namespace oracle {

struct Philox4 {
  uint32_t v[4];
};

inline void philox_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

inline Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, ctr.v[0], &hi0, &lo0);
    philox_mulhilo(M1, ctr.v[2], &hi1, &lo1);
    Philox4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
  }
  return ctr;
}

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t child_lineage(uint64_t parent, uint32_t k) {
  return splitmix64(parent ^ (0x9E3779B97F4A7C15ull * (uint64_t)(k + 1)));
}

inline double u01_from_bits(uint32_t x) {
  return (double)(x >> 9) * (1.0 / 8388608.0) + (1.0 / 16777216.0);
}

/* Stream mode, used only to compare the oracle with run logs the reference
 * produced with MSVC (Math2.cu:1328-1331, rand()/RAND_MAX, RAND_MAX = 32767):
 *   0: per-electron Philox (default);
 *   1: per-electron Philox quantised to rand()/RAND_MAX values in [0, 1];
 *   2: one MSVC rand() linear congruential stream per host thread,
 *      x <- 214013 x + 2531011, value ((x >> 16) & 0x7fff) / 32767. */
extern int g_rng15;
extern thread_local uint32_t g_lcg;

inline double msvc_rand01() {
  g_lcg = g_lcg * 214013u + 2531011u;
  return (double)((g_lcg >> 16) & 0x7fffu) / 32767.0;
}

/* One electron's stream. */
struct Stream {
  uint64_t seed = 0;
  uint64_t lineage = 0;
  uint32_t primary = 0;
  uint32_t draws = 0;
  Philox4 buf{};
  /* probes: when set, next() returns script[0], script[1], ... instead */
  const double* script = nullptr;

  void init(uint64_t s, uint64_t lin, uint32_t prim) {
    seed = s;
    lineage = lin;
    primary = prim;
    draws = 0;
    script = nullptr;
  }
  double next() {
    if (script) return script[draws++];
    if (g_rng15 == 2) return msvc_rand01();
    uint32_t slot = draws & 3u;
    if (slot == 0) {
      Philox4 c;
      c.v[0] = draws >> 2;
      c.v[1] = (uint32_t)lineage;
      c.v[2] = (uint32_t)(lineage >> 32);
      c.v[3] = primary;
      buf = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    }
    ++draws;
    if (g_rng15) return (double)(buf.v[slot] >> 17) / 32767.0;
    return u01_from_bits(buf.v[slot]);
  }
};

}  // namespace oracle

#endif
