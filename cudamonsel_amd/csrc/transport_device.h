/*
 * transport_device.h -- FP32 device implementation of one electron step.
 *
 * Restates, for one GPU lane per electron, the per-step work of
 * MonteCarloSS::takeStep (gov/nist/microanalysis/NISTMonte/MonteCarloSS.cu:107-163)
 * and everything it calls.  Units: nm, eV; directions are unit vectors
 * (the reference keeps theta/phi; Electron::updateDirection, Electron.cu:111-131,
 * is the same rotation written in the electron's local frame).
 *
 * Geometry: every region shape is compiled into convex primitives and a
 * truth table.  The first boundary of the composite along a ray is found by
 * walking the primitives' entry/exit crossings in order and testing where
 * the composite's inside state changes.  This finds the same boundary as the
 * depth-counter walks of NormalIntersectionShape.cu:19-239 and
 * NormalUnionShape.cu:30-271 without their recursive re-queries.
 */
#ifndef MONSEL_TRANSPORT_DEVICE_H
#define MONSEL_TRANSPORT_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/monsel_gpu.h"
#include "device_scene.h"
#include "phase_timing.h"

namespace monsel_dev {

#define MD __device__ __forceinline__

constexpr float kPI = 3.14159265358979323846f;
constexpr float kTwoPI = 6.28318530717958647692f;
constexpr float kInf = __builtin_huge_valf();
constexpr float kA0nm = 0.05291772083f;          /* Bohr radius (nm) */
constexpr float kA0nm2 = 0.05291772083f * 0.05291772083f;
constexpr float kMottLn50 = 3.9120230054281460f;    /* ln(50) */
constexpr float kMottInvParam = 60.0f / 5.9914645471079810f; /* 60 / ln(400) */
constexpr float kMaxNISTMott = 20000.0f;            /* eV */
constexpr float kEcharge = 1.60217653e-19f;         /* J per eV */

/* ----------------------------------------------------------- RNG ------- */

/* Square root, natural log and exponential straight from the hardware
 * instructions (v_sqrt_f32, v_log_f32, v_exp_f32: about 1 ulp; FP32
 * denormals are flushed in this code, so no rescaling is needed).  The
 * compiler's own expansions refine sqrt to correct rounding with a
 * reciprocal-root Newton step and scale log/exp around denormal ranges, some
 * ten instructions each, on the transport loop's critical path. */
MD float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
MD float flog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
MD float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.442695040888963407f); }

/* Philox4x32-10 (Salmon et al., SC'11).  Counter words: (block, lineage lo,
 * lineage hi, primary index) under one fixed key; a job's seed enters
 * through the lineage of its primaries (splitmix64 of the seed), a
 * secondary's lineage is derived from its parent's. */
MD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                      uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t& o3) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    /* each 32x32 -> 64-bit product is one v_mad_u64_u32 (a separate low
     * and high multiply would be two quarter-rate instructions) */
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  o0 = c0;
  o1 = c1;
  o2 = c2;
  o3 = c3;
}

MD uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

MD uint64_t child_lineage(uint64_t parent, uint32_t k) {
  return splitmix64(parent ^ (0x9E3779B97F4A7C15ull * (uint64_t)(k + 1)));
}

/* electron state held in registers */
struct Elec {
  float px, py, pz;
  float dx, dy, dz;
  float E;
  int region;
  uint32_t primary;
  uint32_t lin_lo, lin_hi;
  uint32_t job;
  uint32_t nchild;
  uint32_t steps;
  float safe;   /* lower bound on the distance to the nearest boundary findEndOfStep
                 * would test from here (own shape and children); 0 = unknown */
};

MD float u01(uint32_t x) { return (float)(x >> 9) * 1.1920928955078125e-07f + 5.9604644775390625e-08f; }

/* Uniforms per electron step: Philox blocks 2s and 2s+1 of the electron's
 * stream, s = its step count.  u[0] draws the free path, u[1] the mechanism
 * (or a barrier transmission); the scatter mechanism takes u[2], u[3], ...
 * in the order the reference's scatter() draws them (at most six). */
constexpr int kStepUniforms = 8;

/* u[0..3] from block 2s (drawn when the step starts) */
MD void step_uniforms_lo(const Elec& e, uint32_t k0, uint32_t k1, float* u) {
  uint32_t a0, a1, a2, a3;
  philox4x32_10(2u * e.steps, e.lin_lo, e.lin_hi, e.primary, k0, k1, a0, a1, a2, a3);
  u[0] = u01(a0);
  u[1] = u01(a1);
  u[2] = u01(a2);
  u[3] = u01(a3);
}

/* u[4..7] from block 2s+1, drawn only by a scatter event (s = the step
 * count before the step, `steps` has been advanced since) */
MD void step_uniforms_hi(const Elec& e, uint32_t k0, uint32_t k1, float* u) {
  uint32_t b0, b1, b2, b3;
  philox4x32_10(2u * (e.steps - 1u) + 1u, e.lin_lo, e.lin_hi, e.primary, k0, k1, b0, b1, b2, b3);
  u[4] = u01(b0);
  u[5] = u01(b1);
  u[6] = u01(b2);
  u[7] = u01(b3);
}

/* block reserved for the primary's initial position (never a step block) */
constexpr uint32_t kPrimaryBlock = 0xFFFFFFFFu;
constexpr uint32_t kPhiloxKey0 = 0x6D6F6E73u, kPhiloxKey1 = 0x656C3235u;

/* Lanes per block of every kernel that uses the per-lane LDS arrays below. */
constexpr int kLaneBlock = 256;

/* ----------------------------------------------------------- math ------ */

/* Electron::updateDirection (Electron.cu:111-131) on a unit vector: deflect
 * by polar angle (ca = cos, sa = sin) and azimuth (cb, sb). */
MD void rotate_dir(float& dx, float& dy, float& dz, float ca, float sa, float cb, float sb) {
  float st2 = dx * dx + dy * dy;
  float cp = 1.f, sp = 0.f, st = 0.f;
  if (st2 > 0.f) {
    st = fsqrt(st2);
    float inv = 1.f / st;
    cp = dx * inv;
    sp = dy * inv;
  }
  float ct = dz;
  float xx = cb * ct * sa + ca * st;
  float yy = sa * sb;
  float nx = cp * xx - sp * yy;
  float ny = cp * yy + sp * xx;
  float nz = ca * ct - cb * sa * st;
  float inv = __builtin_amdgcn_rsqf(nx * nx + ny * ny + nz * nz);
  dx = nx * inv;
  dy = ny * inv;
  dz = nz * inv;
}

/* direction from polar/azimuth angles (Electron(parent, theta, phi, kE)) */
MD void dir_from_angles(float ct, float st, float cph, float sph, float& dx, float& dy, float& dz) {
  dx = st * cph;
  dy = st * sph;
  dz = ct;
}

/* Lagrange cubic through unit-spaced nodes 0..3 evaluated at t (the value
 * Neville's scheme in ULagrangeInterpolation.cu:5-33 produces). */
MD float lagrange4(float t, float f0, float f1, float f2, float f3) {
  float t1 = t - 1.f, t2 = t - 2.f, t3 = t - 3.f;
  float l0 = -t1 * t2 * t3 * (1.f / 6.f);
  float l1 = t * t2 * t3 * 0.5f;
  float l2 = -t * t1 * t3 * 0.5f;
  float l3 = t * t1 * t2 * (1.f / 6.f);
  return l0 * f0 + l1 * f1 + l2 * f2 + l3 * f3;
}

MD void lagrange4_w(float t, float& l0, float& l1, float& l2, float& l3) {
  float t1 = t - 1.f, t2 = t - 2.f, t3 = t - 3.f;
  l0 = -t1 * t2 * t3 * (1.f / 6.f);
  l1 = t * t2 * t3 * 0.5f;
  l2 = -t * t1 * t3 * 0.5f;
  l3 = t * t1 * t2 * (1.f / 6.f);
}

/* ULagrangeInterpolation::d1 index choice (ULagrangeInterpolation.cu:41-47). */
MD int u_index0(float reduced, int len) {
  int i0 = (int)reduced - 1;
  i0 = i0 < 0 ? 0 : i0;
  return i0 > len - 4 ? len - 4 : i0;
}

/* ------------------------------------------------- elastic (per element) */

/* BrowningEmpiricalCrossSection.cu:24-29 (nm^2) */
MD float browning_sigma(const MONSEL_CS DElem& el, float EeV) {
  float e = EeV * 1.0e-3f;
  float re = fsqrt(e);
  return 3.0e-4f * el.zp17 / (e + 0.005f * el.zp17 * re + 0.0007f * el.zp2 / re);
}

/* BrowningEmpiricalCrossSection.cu:31-45 -> cos(theta) */
MD float browning_cos(const MONSEL_CS DElem& el, float EeV, float r1, float r2) {
  float e = EeV * 1.0e-3f;
  float r = (300.0f * e / el.z) + (el.zp3 / (3.0e5f * e));
  if (r1 <= r / (r + 1.f)) {
    float alpha = 7.0e-3f / e;
    return 1.0f - ((2.0f * alpha * r2) / (alpha - r2 + 1.f));
  }
  return 1.f - 2.0f * r2;
}

/* ScreenedRutherfordScatteringAngle.cu:24-39 (sigma nm^2, cos theta) */
MD float rutherford_sigma(const MONSEL_CS DElem& el, float EeV) {
  const float K1 = 7.670843088080456e-38f / 1.60217653e-19f * 1.0e18f;
  const float K2 = 5.44967975966321e-19f / 1.60217653e-19f;
  float zp = el.rz13;
  return K1 * zp * (1.f + el.z) / (EeV + K2 * zp * zp);
}
MD float rutherford_cos(const MONSEL_CS DElem& el, float EeV, float r) {
  const float K = 5.44968e-19f / 1.60217653e-19f;
  float alpha = K * el.rz23 / EeV;
  /* alpha - r + 1 summed as (1 - r) + alpha: 1 - r is exact for r >= 1/2,
   * so the denominator keeps its relative precision when alpha and 1 - r are
   * both small (E >= 20 keV, r near 1) */
  return 1.f - 2.0f * alpha * r / ((1.f - r) + alpha);
}

/* The hot tables: the cross-section nodes (Spwem, 61 floats) of the first
 * kLdsMott NIST Mott elements of a scene, which every step above 50 eV reads
 * for every element of its material, live in LDS; each kernel that evaluates
 * cross-sections copies them there at its start (stage_mott_sigma).  The
 * angle tables (61 x 201 per element, read once per elastic event) stay in
 * device memory. */
#ifndef MONSEL_LDS_MOTT_TABLES
#define MONSEL_LDS_MOTT_TABLES 16
#endif
constexpr int kLdsMott = MONSEL_LDS_MOTT_TABLES;
MD float* lds_spwem() {
  __shared__ float l_spwem[(kLdsMott > 0 ? kLdsMott : 1) * kMottNE];
  return l_spwem;
}
/* every thread of the block calls it before any cross-section */
MD void stage_mott_sigma(const DevScene& S) {
  const int n = min(S.n_mott, kLdsMott) * kMottNE;
  float* l = lds_spwem();
  for (int i = threadIdx.x; i < n; i += blockDim.x) l[i] = S.mott_spwem[i];
  __syncthreads();
}

/* NISTMottRS::totalCrossSection (NISTMottRS.cu:103-119), nm^2 */
MD float mott_table_sigma(const DevScene& S, int mott, float EeV) {
  float red = (flog(EeV) - kMottLn50) * kMottInvParam;
  int i0 = u_index0(red, kMottNE);
  float f0, f1, f2, f3;
  if (mott < kLdsMott) {
    const float* f = lds_spwem() + mott * kMottNE + i0;
    f0 = f[0];
    f1 = f[1];
    f2 = f[2];
    f3 = f[3];
  } else {
    const MONSEL_CS float* f = S.mott_spwem + mott * kMottNE + i0;
    f0 = f[0];
    f1 = f[1];
    f2 = f[2];
    f3 = f[3];
  }
  return kA0nm2 * lagrange4(red - (float)i0, f0, f1, f2, f3);
}

MD float elem_sigma(const DevScene& S, const MONSEL_CS DElem& el, float EeV) {
  if (el.kind == MONSEL_MECH_BROWNING_ELASTIC) return browning_sigma(el, EeV);
  if (el.kind == MONSEL_MECH_RUTHERFORD_ELASTIC) return rutherford_sigma(el, EeV);
  if (EeV < el.extrap) {
    if (el.method == 3) return el.mottAtMin * EeV / el.extrap;
    return el.sfBrowning * browning_sigma(el, EeV);
  }
  if (EeV < kMaxNISTMott) return mott_table_sigma(S, el.mott, EeV);
  return rutherford_sigma(el, EeV);
}

/* NISTMottRS::randomScatteringAngle table branch: bicubic Lagrange on X1
 * over (ln E, U) (NISTMottRS.cu:134-150, ULagrangeInterpolation.cu:55-78);
 * returns cos(theta) clamped as the reference does. */
MD float mott_cos(const DevScene& S, int mott, float EeV, float u) {
  float rx = (flog(EeV) - kMottLn50) * kMottInvParam;
  int i0 = u_index0(rx, kMottNE);
  float ry = u * 200.0f;
  int j0 = u_index0(ry, kMottNQ);
  float wx0, wx1, wx2, wx3, wy0, wy1, wy2, wy3;
  lagrange4_w(rx - (float)i0, wx0, wx1, wx2, wx3);
  lagrange4_w(ry - (float)j0, wy0, wy1, wy2, wy3);
  const float* base = S.mott_x1 + (size_t)mott * (kMottNE * kMottNQ) + i0 * kMottNQ + j0;
  float q = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float* row = base + i * kMottNQ;
    float v = wy0 * row[0] + wy1 * row[1] + wy2 * row[2] + wy3 * row[3];
    float wx = i == 0 ? wx0 : (i == 1 ? wx1 : (i == 2 ? wx2 : wx3));
    q += wx * v;
  }
  float com = 1.0f - 2.0f * q * q;
  return com > -1.0f ? (com < 1.0f ? com : 1.0f) : -1.0f;
}

/* ------------------------------------------------------------- CSD ----- */

/* (1 + eps)^(-2/3) - 1 for eps >= 0: the Taylor series below 1e-3 (error
 * < 5e-13 relative), the hardware log2/exp2 above (relative error of the
 * result < 1e-4 there, where it is at least 6.7e-4 in size). */
MD float nieminen_factor(float eps) {
  if (eps < 1e-3f) return eps * (-2.f / 3.f + eps * (5.f / 9.f - eps * (40.f / 81.f)));
  return __builtin_amdgcn_exp2f(-(2.f / 3.f) * __builtin_amdgcn_logf(1.f + eps)) - 1.f;
}

/* JoyLuoNieminenCSD::compute (JoyLuoNieminenCSD.cu:91-121) with the
 * recursive halving of :154-159 as an explicit depth-first walk: bit d of
 * `pending` marks a second half at depth d.  Below breakE the Nieminen form
 * kE (1 + 1.5 gamma len kE^1.5)^(-2/3) - kE; above it the Joy-Luo loss,
 * halved while it exceeds 10% of the energy. */
MD float csd_compute_split(const MONSEL_CS DCSD& c, float len, float kE);

/* The common case in straight-line code: one segment, no halving (the
 * first pass of csd_compute_split, with the same arithmetic). */
MD float csd_compute(const MONSEL_CS DCSD& c, float len, float kE) {
  if (!c.active || c.nce == 0) return 0.f;
  if (kE < c.minE || kE <= 0.f) return 0.f;
  if (kE <= c.breakE) {
    const float eps = 1.5f * c.gamma * len * kE * fsqrt(kE);
    return kE * nieminen_factor(eps);
  }
  float s = 0.f;
  for (int i = 0; i < c.nce; ++i) s += c.coef[i] * flog(c.recipJ[i] * kE + c.beta[i]);
  const float loss = s * len / kE;
  if (!(loss > 0.1f * kE)) return -loss;
  return csd_compute_split(c, len, kE);
}

MD float csd_compute_split(const MONSEL_CS DCSD& c, float len, float kE) {
  float E = kE, total = 0.f;
  uint32_t pending = 0;
  int d = 0;
  for (int it = 0; it < 128; ++it) {
    const float seg = ldexpf(len, -d);
    float dE = 0.f;
    if (!(E < c.minE || E <= 0.f)) {
      if (E <= c.breakE) {
        const float eps = 1.5f * c.gamma * seg * E * fsqrt(E);
        dE = E * nieminen_factor(eps);
      } else {
        float s = 0.f;
        for (int i = 0; i < c.nce; ++i) s += c.coef[i] * flog(c.recipJ[i] * E + c.beta[i]);
        const float loss = s * seg / E;
        if (loss > 0.1f * E && d < 30) {
          pending |= 1u << (d + 1);
          d += 1;
          continue;
        }
        dE = -loss;
      }
    }
    E += dE;
    total += dE;
    if (pending == 0) break;
    d = 31 - __clz(pending);
    pending &= ~(1u << d);
  }
  return total;
}

/* ------------------------------------------------------- NU tables ----- */

/* Non-uniform table interpolation, generic over the arithmetic type T:
 * float for the inelastic mean free path (a well-conditioned 1-D lookup),
 * double for the tabulated scatter, whose 3-D angle lookup in reference
 * mode extrapolates a cubic far outside its abscissae
 * (NULagrangeInterpolation.cu:236) with weights up to ~1e7, so that FP32
 * rounding of the table values would change the result. */
template <typename T>
MD const T* nu_base(const DevScene& S);
template <>
MD const float* nu_base<float>(const DevScene& S) { return S.nu_data; }
template <>
MD const double* nu_base<double>(const DevScene& S) { return S.nu_data64; }

/* NULagrangeInterpolation::locate (NULagrangeInterpolation.cu:5-68) on an
 * ascending or descending grid: first index of the order+1 window. */
template <typename T>
MD int nu_locate(const T* xs, int n, int order, T x) {
  int lowlim = -1, uplim = n, maxindex = n - 1;
  bool asc = xs[maxindex] > xs[0];
  while (uplim - lowlim > 1) {
    int mid = (uplim + lowlim) >> 1;
    if ((x > xs[mid]) == asc) lowlim = mid;
    else uplim = mid;
  }
  if (lowlim < 0) return 0;
  if (uplim > maxindex) return maxindex - order;
  T xl = xs[lowlim], xu = xs[uplim];
  int first = lowlim, last = uplim;
  while ((last - first) < order && first > 0 && last < maxindex) {
    if ((((xl - xs[first - 1]) + xu) - xs[last + 1] >= (T)0) == asc) last += 1;
    else first -= 1;
  }
  if (last >= maxindex) return maxindex - order;
  if (first <= 0) return 0;
  return first;
}

/* Lagrange basis weights of the ORDER+1 abscissae xs[0..ORDER] at x: the
 * interpolant sum_i w_i f_i is the polynomial Neville's scheme
 * (NULagrangeInterpolation.cu:85-112) evaluates.  Weights are formed once
 * per axis and shared by every row of a multi-dimensional window. */
template <typename T, int ORDER>
MD void nu_weights(const T* xs, T x, T* w) {
  /* differences are taken in units of the window's span, so that products
   * of ORDER of them stay far from underflow whatever the grid's units
   * (energies in J are ~1e-17) */
  T xv[ORDER + 1];
#pragma unroll
  for (int i = 0; i <= ORDER; ++i) xv[i] = xs[i];
  const T inv = (T)1 / (xv[ORDER] - xv[0]);
#pragma unroll
  for (int i = 0; i <= ORDER; ++i) {
    T num = 1, den = 1;
#pragma unroll
    for (int j = 0; j <= ORDER; ++j)
      if (j != i) {
        num *= (x - xv[j]) * inv;
        den *= (xv[i] - xv[j]) * inv;
      }
    w[i] = num / den;
  }
}

template <typename T, int ORDER>
MD T nu_dot(const T* w, const T* f) {
  T y = 0;
#pragma unroll
  for (int i = 0; i <= ORDER; ++i) y += w[i] * f[i];
  return y;
}

template <typename T, int ORDER>
MD T nu_poly(const T* xs, const T* f, T x) {
  T w[ORDER + 1];
  nu_weights<T, ORDER>(xs, x, w);
  return nu_dot<T, ORDER>(w, f);
}

/* NULagrangeInterpolation::d1 (:144-158). */
template <typename T, int ORDER>
MD T nu_d1(const DevScene& S, const MONSEL_CS DNUTable& t, T x0) {
  const T* d = nu_base<T>(S);
  const T* xa = d + t.xoff[0];
  int i0 = nu_locate<T>(xa, t.n[0], ORDER, x0);
  return nu_poly<T, ORDER>(xa + i0, d + t.voff + i0, x0);
}

/* NULagrangeInterpolation::d2 (:169-192); values are an [na][nb] block. */
template <typename T, int ORDER>
MD T nu_d2(const DevScene& S, const MONSEL_CS DNUTable& t, T x0, T x1) {
  const T* d = nu_base<T>(S);
  const T* xa = d + t.xoff[0];
  const T* xb = d + t.xoff[1];
  const T* f = d + t.voff;
  const int nb = t.n[1];
  int i0 = nu_locate<T>(xa, t.n[0], ORDER, x0);
  int j0 = nu_locate<T>(xb, nb, ORDER, x1);
  T wb[ORDER + 1], y[ORDER + 1];
  nu_weights<T, ORDER>(xb + j0, x1, wb);
#pragma unroll
  for (int i = 0; i <= ORDER; ++i) y[i] = nu_dot<T, ORDER>(wb, f + (size_t)(i0 + i) * nb + j0);
  return nu_poly<T, ORDER>(xa + i0, y, x0);
}

/* NULagrangeInterpolation::d3 (:215-240); compat_d3 reproduces line 236
 * (outer abscissae taken from the start of the grid).  voff overrides the
 * value offset (the 4-D lookup walks 3-D slices). */
template <typename T, int ORDER>
MD T nu_d3(const DevScene& S, const MONSEL_CS DNUTable& t, int ax, int voff, T x0, T x1, T x2, bool compat_d3) {
  const T* d = nu_base<T>(S);
  const T* xa = d + t.xoff[ax];
  const T* xb = d + t.xoff[ax + 1];
  const T* xc = d + t.xoff[ax + 2];
  const T* v = d + voff;
  const int nb = t.n[ax + 1], nc = t.n[ax + 2];
  int i0 = nu_locate<T>(xa, t.n[ax], ORDER, x0);
  int j0 = nu_locate<T>(xb, nb, ORDER, x1);
  int k0 = nu_locate<T>(xc, nc, ORDER, x2);
  T wb[ORDER + 1], wc[ORDER + 1], y[ORDER + 1];
  nu_weights<T, ORDER>(xb + j0, x1, wb);
  nu_weights<T, ORDER>(xc + k0, x2, wc);
#pragma unroll 1
  for (int i = 0; i <= ORDER; ++i) {
    T z[ORDER + 1];
    const T* blk = v + (size_t)(i0 + i) * nb * nc;
#pragma unroll 1
    for (int j = 0; j <= ORDER; ++j) {
      const T zj = nu_dot<T, ORDER>(wc, blk + (size_t)(j0 + j) * nc + k0);
#pragma unroll
      for (int q = 0; q <= ORDER; ++q)
        if (q == j) z[q] = zj;
    }
    const T yi = nu_dot<T, ORDER>(wb, z);
#pragma unroll
    for (int q = 0; q <= ORDER; ++q)
      if (q == i) y[q] = yi;
  }
  return nu_poly<T, ORDER>(compat_d3 ? xa : xa + i0, y, x0);
}

/* NULagrangeInterpolation::d4 (:259-306). */
template <typename T, int ORDER>
MD T nu_d4(const DevScene& S, const MONSEL_CS DNUTable& t, const T* x, bool compat_d3) {
  const T* xa = nu_base<T>(S) + t.xoff[0];
  int i0 = nu_locate<T>(xa, t.n[0], ORDER, x[0]);
  T y[ORDER + 1];
#pragma unroll
  for (int i = 0; i <= ORDER; ++i)
    y[i] = nu_d3<T, ORDER>(S, t, 1, t.voff + (i0 + i) * t.n[1] * t.n[2] * t.n[3], x[1], x[2], x[3], compat_d3);
  return nu_poly<T, ORDER>(xa + i0, y, x[0]);
}

template <typename T, int ORDER>
MD T nu_interp_o(const DevScene& S, int table, const T* x, bool compat_d3) {
  const MONSEL_CS DNUTable& t = S.nu[table];
  switch (t.dim) {
    case 1: return nu_d1<T, ORDER>(S, t, x[0]);
    case 2: return nu_d2<T, ORDER>(S, t, x[0], x[1]);
    case 3: return nu_d3<T, ORDER>(S, t, 0, t.voff, x[0], x[1], x[2], compat_d3);
    case 4: return nu_d4<T, ORDER>(S, t, x, compat_d3);
  }
  return 0;
}

/* NUTableInterpolation::interpolate (NUTableInterpolation.cu:98-115) in
 * the arithmetic the transport uses for that table's dimension: FP32 for
 * 1-D (mean free path), FP64 otherwise. */
MD double nu_interp(const DevScene& S, int table, int order, const double* x, bool compat_d3) {
  if (S.nu[table].dim == 1) {
    const float xf = (float)x[0];
    if (order == 1) return nu_interp_o<float, 1>(S, table, &xf, compat_d3);
    if (order == 2) return nu_interp_o<float, 2>(S, table, &xf, compat_d3);
    return nu_interp_o<float, 3>(S, table, &xf, compat_d3);
  }
  if (order == 1) return nu_interp_o<double, 1>(S, table, x, compat_d3);
  if (order == 2) return nu_interp_o<double, 2>(S, table, x, compat_d3);
  return nu_interp_o<double, 3>(S, table, x, compat_d3);
}

/* ------------------------------------------------------ mechanisms ----- */

MD float elastic_rate(const DevScene& S, const MONSEL_CS DMech& m, float EeV) {
  float tot = 0.f;
  for (int i = 0; i < m.nelem; ++i) {
    const MONSEL_CS DElem& el = S.elems[m.elem0 + i];
    tot += elem_sigma(S, el, EeV) * el.natoms;
  }
  return tot;
}

/* ScatterMechanism::scatterRate (1/nm) */
MD float mech_rate(const DevScene& S, const MONSEL_CS DMech& m, float EeV) {
  switch (m.kind) {
    case MONSEL_MECH_NISTMOTT_ELASTIC:
    case MONSEL_MECH_BROWNING_ELASTIC:
    case MONSEL_MECH_RUTHERFORD_ELASTIC:
      return elastic_rate(S, m, EeV);
    case MONSEL_MECH_FITTED_INELASTIC: /* FittedInelSM.cu:20-25 */
      if (EeV <= m.p[0]) return 0.f;
      return -csd_compute(m.csd, 0.1f, EeV) * 10.f * m.p[2];
    case MONSEL_MECH_POLARON_TRAP: /* GanachaudMokraniPolaronTrapSM.cu:21-27 */
      if (EeV > m.p[2]) return 0.f;
      return m.p[0] * fexp(-m.p[1] * EeV);
    case MONSEL_MECH_PHONON: { /* GanachaudMokraniPhononInelasticSM.cu:46-72 */
      if (EeV < m.p[0]) return 0.f;
      float x = m.p[0] / EeV;
      if (x < 0.1f) return m.p[1] * x * flog(4.f / x);
      if (x >= 1.f) return 0.f;
      float temp = fsqrt(1.f - x);
      return m.p[1] * x * flog((1.f + temp) / (1.f - temp));
    }
    case MONSEL_MECH_TABULATED_INELASTIC: { /* TabulatedInelasticSM.cu:436-459 */
      float kE = (EeV + m.p[TP_OFFSET]) * kEcharge;
      if (kE < m.p[TP_IIMFPDOM0]) return 0.f;
      return m.p[TP_RATEMULT] * nu_d1<float, 1>(S, S.nu[m.table[0]], kE) * 1.0e-9f;
    }
  }
  return 0.f;
}

/* secondary produced by a scatter event (new Electron(*pe, theta, phi, E)
 * followed by an optional relative rotation) */
struct Spawn {
  bool has;
  float dx, dy, dz;
  float E;
};

/* TabulatedInelasticSM.cu:103-151 (energies in J) */
MD float e0_from_dispersion(float Eq, float deltaE) {
  if (Eq == 0.f) return deltaE;
  float x = Eq / deltaE;
  float y = 3.77300614251479e17f * Eq;
  float x13 = cbrtf(x);
  float x2 = x * x;
  float c1 = x2 * x2 * (27.f + (18.f * y) + (2.f * y * y));
  float c2 = x2 * (27.f + (4.f * y));
  float c3 = c2 - 27.f;
  float c4 = x2 * (3.f + y);
  float c6 = 1.f - x2;
  if (c3 > 0.f) {
    float tanPart = atan2f(3.f * c6 * sqrtf(3.f * c3 * c6), (18.f * c4) - c1 - 27.f);
    float trigPart = cosf(tanPart / 3.f);
    float prefactor = 2.f * x * sqrtf(y * (y - (c6 * (y + 6.f))));
    trigPart *= prefactor;
    return deltaE * sqrtf(((3.f - c4) + trigPart) / 3.f);
  }
  float c5 = cbrtf(-c1 + (18.f * c4) + (3.f * (-9.f + (c6 * sqrtf(-(3.f * c3 * c6))))));
  float term1 = -2.f * c4;
  float x23 = x13 * x13;
  float x43 = x * x13;
  float y13 = cbrtf(y);
  float y23 = y13 * y13;
  float x43y23 = x43 * y23;
  float two13 = 1.2599210498948732f;
  float term2 = -12.f * two13 * x43y23 * c6;
  float term3 = 2.f * two13 * x43y23 * x2 * y;
  float term23 = (term2 + term3) / c5;
  float term4 = two13 * two13 * x23 * y13 * c5;
  return deltaE * sqrtf((6.f + term1 + term23 + term4) / 6.f);
}

/* TabulatedInelasticSM::pickBE (:153-207), default branching ratios; J. */
MD double pick_be(const MONSEL_CS DMech& m, double Eq, double deltaE) {
  const int ncore = (int)m.q[TP_NCORE];
  if (ncore > 0 && deltaE <= m.q[TP_CORE0]) return 0.;
  const double eIon = m.q[TP_E0DISP] != 0. ? (double)e0_from_dispersion((float)Eq, (float)deltaE) : deltaE;
  int i = 0;
  while (i < ncore && m.q[TP_CORE0 + i] <= eIon) ++i;
  return i == 0 ? 0. : m.q[TP_CORE0 + i - 1];
}

MD void isotropic_dir(float u1, float u2, float& dx, float& dy, float& dz) {
  /* theta = acos(1 - 2 u1), phi = 2 pi u2 */
  float ct = 1.f - 2.f * u1;
  float st = fsqrt(fmaxf(0.f, 1.f - ct * ct));
  float sp, cp;
  __sincosf(kTwoPI * u2, &sp, &cp);
  dx = st * cp;
  dy = st * sp;
  dz = ct;
}

/* TabulatedInelasticSM::scatter (TabulatedInelasticSM.cu:209-434) in FP64,
 * in the reference's SI units (J); only the final rotations are FP32.  The
 * parameters come from DMech::q (J). */
MD Spawn tabulated_scatter(const DevScene& S, const MONSEL_CS DMech& m, Elec& e, const float* u) {
  Spawn se;
  se.has = false;
  const bool cd3 = (S.compat & MONSEL_COMPAT_D3_ABSCISSAE) != 0;
  const double J = 1.60217653e-19;
  const double kE0 = (double)e.E * J;
  const double kE = kE0 + m.q[TP_OFFSET];
  if (kE < m.q[TP_EIDOM0]) return se;
  const double rnd0 = u[2], rnd1 = u[3], rnd2 = u[4], rnd3 = u[5];
  double deltaE = kE * nu_d2<double, 3>(S, S.nu[m.table[1]], kE, rnd0);
  const double gap = m.q[TP_GAP], bandgap = m.q[TP_BANDGAP];
  if (deltaE < gap && deltaE > 0.95 * gap) deltaE = gap;
  const float d0x = e.dx, d0y = e.dy, d0z = e.dz;
  double theta = 0.;
  float sphi = 0.f, cphi = 1.f;
  const double offsetFermi = m.q[TP_OFFSETFERMI];
  if (deltaE >= bandgap) {
    double red = (deltaE - gap) / (kE - offsetFermi - 2. * gap);
    red = red > 1. ? 1. : (red < 0. ? 0. : red);
    theta = nu_d3<double, 3>(S, S.nu[m.table[2]], 0, S.nu[m.table[2]].voff, kE, red, rnd1, cd3);
    __sincosf(kTwoPI * (float)rnd2, &sphi, &cphi);
    float st, ct;
    sincosf((float)theta, &st, &ct);
    rotate_dir(e.dx, e.dy, e.dz, ct, st, cphi, sphi);
  }
  e.E = (float)((kE0 - deltaE) / J);
  if (deltaE < bandgap) return se;
  const double Eq = (2. * kE) - deltaE - (2. * sqrt(kE * (kE - deltaE)) * (double)cosf((float)theta));
  const double cb = m.q[TP_CB], minEgen = m.q[TP_MINEGEN], bEref = m.q[TP_BEREF];
  const double offset = m.q[TP_OFFSET];
  double energySE;
  const int method = m.method;
  /* SE direction relative to the PE's original direction: thetaSE = pi/2 - theta,
   * phiSE = phi + pi */
  auto rel_dir = [&](float& sx, float& sy, float& sz) {
    sx = d0x;
    sy = d0y;
    sz = d0z;
    float st, ct;
    sincosf((float)(0.5 * 3.14159265358979323846 - theta), &st, &ct);
    rotate_dir(sx, sy, sz, ct, st, -cphi, -sphi);
  };
  auto se0 = [&](double dE) {
    double e0 = nu_d2<double, 3>(S, S.nu[m.table[3]], dE, rnd3);
    if (e0 < m.q[TP_SE0RANGE0]) e0 = m.q[TP_SE0RANGE0];
    else if (e0 > m.q[TP_SE0RANGE1]) e0 = m.q[TP_SE0RANGE1];
    return e0;
  };
  if (method == 1) {
    energySE = (deltaE + bEref) - pick_be(m, Eq, deltaE);
    if (energySE + cb < minEgen) return se;
    rel_dir(se.dx, se.dy, se.dz);
  } else if (method == 2) {
    const double be = pick_be(m, Eq, deltaE);
    if (be > 0.) energySE = (deltaE + bEref) - be;
    else energySE = (deltaE + se0(deltaE)) - offset;
    if (energySE + cb < minEgen) return se;
    rel_dir(se.dx, se.dy, se.dz);
  } else {
    const double be = pick_be(m, Eq, deltaE);
    if (be > 0.) {
      energySE = (deltaE + bEref) - be;
      if (energySE + cb < minEgen) return se;
      isotropic_dir(u[6], u[7], se.dx, se.dy, se.dz);
    } else {
      const double root = 2. * sqrt(offsetFermi * (offsetFermi + deltaE));
      const double sum = (2. * offsetFermi) + deltaE;
      if ((sum - root) <= Eq && Eq <= (sum + root)) {
        /* simESEf (TabulatedInelasticSM.cu:87-101) */
        const double q = sqrt(Eq);
        const double kz = (deltaE - Eq) / 2. / q;
        const double kzf = kz + q;
        const double Ezq = kzf * kzf;
        double minE = (offsetFermi + bandgap) - Ezq;
        if (minE < 0.) minE = 0.;
        const double maxE = offsetFermi - (kz * kz);
        const double Exy = (minE * (1. - rnd3)) + (maxE * rnd3);
        const double ESEf = Exy + Ezq;
        const float thSE = acosf((float)(kzf / sqrt(ESEf)));
        energySE = ESEf - offset;
        if (energySE + cb < minEgen) return se;
        /* TabulatedInelasticSM.cu:388-395: the q-vector direction (pi/2 - theta,
         * phi + pi) is deflected by (thSE, 2 pi U) in the PE frame, and the SE
         * is the PE's original direction rotated by the resulting angles. */
        float st, ct;
        sincosf((float)(0.5 * 3.14159265358979323846 - theta), &st, &ct);
        float wx, wy, wz;
        dir_from_angles(ct, st, -cphi, -sphi, wx, wy, wz);
        float sp2, cp2, st2, ct2;
        __sincosf(kTwoPI * u[6], &sp2, &cp2);
        sincosf(thSE, &st2, &ct2);
        rotate_dir(wx, wy, wz, ct2, st2, cp2, sp2);
        float wst = fsqrt(fmaxf(0.f, wx * wx + wy * wy));
        float wcp = 1.f, wsp = 0.f;
        if (wst > 0.f) {
          wcp = wx / wst;
          wsp = wy / wst;
        }
        se.dx = d0x;
        se.dy = d0y;
        se.dz = d0z;
        rotate_dir(se.dx, se.dy, se.dz, wz, wst, wcp, wsp);
      } else {
        energySE = (deltaE + se0(deltaE)) - offset;
        if (energySE + cb < minEgen) return se;
        isotropic_dir(u[6], u[7], se.dx, se.dy, se.dz);
      }
    }
  }
  se.has = true;
  se.E = (float)(energySE / J);
  return se;
}

/* ScatterMechanism::scatter.  `rate` is this mechanism's scatterRate at
 * prevE, as summed when the step's free path was drawn. */
/* TAB = false leaves the tabulated mechanism out of the code (its FP64
 * scatter sets the register budget of every kernel that contains it); only
 * for scenes without one. */
template <bool TAB = true>
MD Spawn mech_scatter(const DevScene& S, const MONSEL_CS DMech& m, Elec& e, float prevE, float rate, bool& trapped,
                      const float* u, const float* esig = nullptr) {
  Spawn se;
  se.has = false;
  switch (m.kind) {
    case MONSEL_MECH_NISTMOTT_ELASTIC:
    case MONSEL_MECH_BROWNING_ELASTIC:
    case MONSEL_MECH_RUTHERFORD_ELASTIC: { /* SelectableElasticSM.cu:49-67 */
      PT_BEGIN(t_el);
      /* first element whose running sum of sigma * atoms reaches u * rate
       * (the last if none does); the sum is re-formed in the order
       * elastic_rate formed it, from the rate's own terms when the caller
       * kept them (esig: the first nelem - 1 terms, lane stride kLaneBlock) */
      const float r = u[2] * rate;
      int idx = 0;
      float cum = 0.f;
      for (; idx < m.nelem - 1; ++idx) {
        const MONSEL_CS DElem& el = S.elems[m.elem0 + idx];
        cum += esig ? esig[idx * kLaneBlock] : elem_sigma(S, el, prevE) * el.natoms;
        if (!(cum < r)) break;
      }
      const MONSEL_CS DElem& el = S.elems[m.elem0 + idx];
      /* the polar angle takes one uniform (two for Browning), the azimuth
       * the next one (SelectableElasticSM.cu:62-64) */
      float ca, ub;
      if (el.kind == MONSEL_MECH_BROWNING_ELASTIC ||
          (el.kind == MONSEL_MECH_NISTMOTT_ELASTIC && prevE < el.extrap)) {
        ca = browning_cos(el, prevE, u[3], u[4]);
        ub = u[5];
      } else if (el.kind == MONSEL_MECH_RUTHERFORD_ELASTIC || prevE >= kMaxNISTMott) {
        ca = rutherford_cos(el, prevE, u[3]);
        ub = u[4];
      } else {
        ca = mott_cos(S, el.mott, prevE, u[3]);
        ub = u[4];
      }
      ca = fminf(1.f, fmaxf(-1.f, ca));
      float sa = fsqrt(fmaxf(0.f, 1.f - ca * ca));
      float sb, cb;
      __sincosf(kTwoPI * ub, &sb, &cb);
      rotate_dir(e.dx, e.dy, e.dz, ca, sa, cb, sb);
      PT_END(PH_ELASTIC, t_el);
      return se;
    }
    case MONSEL_MECH_FITTED_INELASTIC: { /* FittedInelSM.cu:13-18 */
      PT_BEGIN(t_fi);
      float u_phi = u[2];
      float u_th = u[3];
      isotropic_dir(u_th, u_phi, se.dx, se.dy, se.dz);
      se.E = m.p[1];
      se.has = true;
      PT_END(PH_FITTED, t_fi);
      return se;
    }
    case MONSEL_MECH_POLARON_TRAP: { /* GanachaudMokraniPolaronTrapSM.cu:15-20 */
      PT_BEGIN(t_po);
      e.E = 0.f;
      trapped = true;
      PT_END(PH_POLARON, t_po);
      return se;
    }
    case MONSEL_MECH_PHONON: { /* GanachaudMokraniPhononInelasticSM.cu:19-44 */
      PT_BEGIN(t_ph);
      float kE0 = e.E;
      if (kE0 < m.p[0]) {
        PT_END(PH_PHONON, t_ph);
        return se;
      }
      float x = m.p[0] / kE0;
      const float r0 = u[2], r1 = u[3];
      float ct;
      /* x^a = 2^(a log2 x) for x > 0, on the hardware exp2/log2 */
      if (x < 0.1f)
        ct = 1.f + ((x * x) - __builtin_amdgcn_exp2f(4.f * r0 + (2.f - 2.f * r0) * __builtin_amdgcn_logf(x))) / 8.f;
      else {
        float root = fsqrt(1.f - x);
        float temp = __builtin_amdgcn_exp2f(
            r0 * __builtin_amdgcn_logf(((-2.f * (1.f + root)) + x) / ((-2.f * (1.f - root)) + x)));
        ct = temp + (((x - 2.f) * (temp - 1.f)) / 2.f / root);
      }
      ct = fminf(1.f, fmaxf(-1.f, ct));
      float st = fsqrt(fmaxf(0.f, 1.f - ct * ct));
      float sb, cb;
      __sincosf(kTwoPI * r1, &sb, &cb);
      rotate_dir(e.dx, e.dy, e.dz, ct, st, cb, sb);
      e.E = kE0 - m.p[0];
      PT_END(PH_PHONON, t_ph);
      return se;
    }
    case MONSEL_MECH_TABULATED_INELASTIC:
      if constexpr (TAB) {
        PT_BEGIN(t_tb);
        Spawn r = tabulated_scatter(S, m, e, u);
        PT_END(PH_TABULATED, t_tb);
        return r;
      }
      break;
  }
  return se;
}

/* --------------------------------------------------------- geometry ---- */

struct Interval {
  float tin, tout;
  int fin, fout; /* face codes: plane index, or -1 cap0 / -2 cap1 / -3 side */
};

MD bool plane_tie_break(float nx, float ny, float nz) {
  /* NormalMultiPlaneShape.cu:26-38 */
  if (nx < 0.f) return false;
  if (nx == 0.f) {
    if (ny < 0.f) return false;
    if (ny == 0.f && nz < 0.f) return false;
  }
  return true;
}

/* plane k of a multi-plane primitive */
MD float4 prim_plane(const DevScene& S, const MONSEL_CS DPrim& pr, int k) {
  return k < kInlinePlanes ? pr.pl[k] : S.planes[pr.first_plane + k];
}

/* Parameter interval of the line p + t d inside a convex primitive. */
MD Interval prim_interval(const DevScene& S, const MONSEL_CS DPrim& pr, float px, float py, float pz, float dx, float dy,
                          float dz) {
  Interval iv;
  iv.tin = -kInf;
  iv.tout = kInf;
  iv.fin = iv.fout = -4;
  if (pr.kind == PRIM_PLANES) {
    for (int k = 0; k < pr.n_planes; ++k) {
      float4 pl = prim_plane(S, pr, k);
      float num = pl.x * px + pl.y * py + pl.z * pz - pl.w;
      float den = pl.x * dx + pl.y * dy + pl.z * dz;
      if (den == 0.f) {
        if (num < 0.f || (num == 0.f && plane_tie_break(pl.x, pl.y, pl.z))) continue;
        iv.tin = kInf;
        iv.tout = -kInf;
        return iv;
      }
      float t = -num / den;
      if (den > 0.f) {
        if (t < iv.tout) {
          iv.tout = t;
          iv.fout = k;
        }
      } else if (t > iv.tin) {
        iv.tin = t;
        iv.fin = k;
      }
    }
    return iv;
  }
  if (pr.kind == PRIM_CYLINDER) {
    float wx = px - pr.e0[0], wy = py - pr.e0[1], wz = pz - pr.e0[2];
    float apar = wx * pr.ax[0] + wy * pr.ax[1] + wz * pr.ax[2];
    float dpar = dx * pr.ax[0] + dy * pr.ax[1] + dz * pr.ax[2];
    float sin_ = -kInf, sout = kInf;
    int fsin = -4, fsout = -4;
    if (dpar != 0.f) {
      float inv = 1.f / dpar;
      float t0 = -apar * inv, t1 = (pr.len - apar) * inv;
      if (dpar > 0.f) {
        sin_ = t0; fsin = -1; sout = t1; fsout = -2;
      } else {
        sin_ = t1; fsin = -2; sout = t0; fsout = -1;
      }
    } else if (!(apar >= 0.f && apar <= pr.len)) {
      iv.tin = kInf;
      iv.tout = -kInf;
      return iv;
    }
    /* radial part from the components perpendicular to the axis */
    float qx = wx - apar * pr.ax[0], qy = wy - apar * pr.ax[1], qz = wz - apar * pr.ax[2];
    float vx = dx - dpar * pr.ax[0], vy = dy - dpar * pr.ax[1], vz = dz - dpar * pr.ax[2];
    float A = vx * vx + vy * vy + vz * vz;
    float B = qx * vx + qy * vy + qz * vz;
    float C = qx * qx + qy * qy + qz * qz - pr.r2;
    float rin = -kInf, rout = kInf;
    if (A > 1e-12f) {
      float disc = B * B - A * C;
      if (disc < 0.f) {
        iv.tin = kInf;
        iv.tout = -kInf;
        return iv;
      }
      float sq = fsqrt(disc);
      float q = -(B + (B >= 0.f ? sq : -sq));
      float ra = q / A, rb = (q != 0.f) ? C * __builtin_amdgcn_rcpf(q) : ra;
      rin = fminf(ra, rb);
      rout = fmaxf(ra, rb);
    } else if (C > 0.f) {
      iv.tin = kInf;
      iv.tout = -kInf;
      return iv;
    }
    if (sin_ >= rin) { iv.tin = sin_; iv.fin = fsin; }
    else { iv.tin = rin; iv.fin = -3; }
    if (sout <= rout) { iv.tout = sout; iv.fout = fsout; }
    else { iv.tout = rout; iv.fout = -3; }
    return iv;
  }
  /* sphere */
  float cx = px - pr.e0[0], cy = py - pr.e0[1], cz = pz - pr.e0[2];
  float B = cx * dx + cy * dy + cz * dz;
  float C = cx * cx + cy * cy + cz * cz - pr.r2;
  float disc = B * B - C;
  if (disc < 0.f) {
    iv.tin = kInf;
    iv.tout = -kInf;
    return iv;
  }
  float sq = fsqrt(disc);
  iv.tin = -B - sq;
  iv.tout = -B + sq;
  iv.fin = iv.fout = -3;
  return iv;
}

/* Outward normal of primitive face `f` at point q. */
MD void prim_normal(const DevScene& S, const MONSEL_CS DPrim& pr, int f, float qx, float qy, float qz, float& nx, float& ny,
                    float& nz) {
  if (pr.kind == PRIM_PLANES) {
    float4 pl = prim_plane(S, pr, f < 0 ? 0 : f);
    nx = pl.x;
    ny = pl.y;
    nz = pl.z;
    return;
  }
  if (pr.kind == PRIM_CYLINDER) {
    if (f == -1) { nx = -pr.ax[0]; ny = -pr.ax[1]; nz = -pr.ax[2]; return; }
    if (f == -2) { nx = pr.ax[0]; ny = pr.ax[1]; nz = pr.ax[2]; return; }
    float wx = qx - pr.e0[0], wy = qy - pr.e0[1], wz = qz - pr.e0[2];
    float apar = wx * pr.ax[0] + wy * pr.ax[1] + wz * pr.ax[2];
    wx -= apar * pr.ax[0];
    wy -= apar * pr.ax[1];
    wz -= apar * pr.ax[2];
    float inv = __builtin_amdgcn_rsqf(wx * wx + wy * wy + wz * wz);
    nx = wx * inv;
    ny = wy * inv;
    nz = wz * inv;
    return;
  }
  float wx = qx - pr.e0[0], wy = qy - pr.e0[1], wz = qz - pr.e0[2];
  float inv = __builtin_amdgcn_rsqf(wx * wx + wy * wy + wz * wz);
  nx = wx * inv;
  ny = wy * inv;
  nz = wz * inv;
}

MD bool prim_contains(const DevScene& S, const MONSEL_CS DPrim& pr, float px, float py, float pz) {
  if (pr.kind == PRIM_PLANES) {
    /* MultiPlaneShape.cu:23-27 for every plane */
    for (int k = 0; k < pr.n_planes; ++k) {
      float4 pl = prim_plane(S, pr, k);
      if (!(pl.x * px + pl.y * py + pl.z * pz - pl.w <= 0.f)) return false;
    }
    return true;
  }
  float wx = px - pr.e0[0], wy = py - pr.e0[1], wz = pz - pr.e0[2];
  if (pr.kind == PRIM_CYLINDER) {
    /* CylindricalShape.cu:41-45 */
    float u = (pr.axv[0] * wx + pr.axv[1] * wy + pr.axv[2] * wz) / pr.len2;
    float rx = wx - u * pr.axv[0], ry = wy - u * pr.axv[1], rz = wz - u * pr.axv[2];
    return u >= 0.f && u <= 1.f && (rx * rx + ry * ry + rz * rz) <= pr.r2;
  }
  return wx * wx + wy * wy + wz * wz <= pr.r2;
}

/* max_k g_k over a convex primitive's constraints, g_k the signed distance
 * to constraint k (plane rows scaled by sdscale <= 1/|n|): <= 0 inside.
 * Its magnitude bounds the distance from p to the primitive's surface from
 * below -- inside, -max g_k is the distance to the nearest face; outside,
 * max g_k is at most the distance to the primitive. */
MD float prim_gmax(const DevScene& S, const MONSEL_CS DPrim& pr, float px, float py, float pz) {
  if (pr.kind == PRIM_PLANES) {
    float g = -kInf;
    for (int k = 0; k < pr.n_planes; ++k) {
      const float4 pl = prim_plane(S, pr, k);
      g = fmaxf(g, pl.x * px + pl.y * py + pl.z * pz - pl.w);
    }
    return g * pr.sdscale;
  }
  const float wx = px - pr.e0[0], wy = py - pr.e0[1], wz = pz - pr.e0[2];
  if (pr.kind == PRIM_CYLINDER) {
    const float apar = wx * pr.ax[0] + wy * pr.ax[1] + wz * pr.ax[2];
    const float qx = wx - apar * pr.ax[0], qy = wy - apar * pr.ax[1], qz = wz - apar * pr.ax[2];
    const float radial = fsqrt(qx * qx + qy * qy + qz * qz) - fsqrt(pr.r2);
    return fmaxf(radial, fmaxf(-apar, apar - pr.len));
  }
  return fsqrt(wx * wx + wy * wy + wz * wz) - fsqrt(pr.r2);
}

MD bool truth(const MONSEL_CS DShape& sh, uint32_t bits) { return (sh.truth[bits >> 5] >> (bits & 31u)) & 1u; }

/* prim_gmax, and in `in` the primitive's inside bit as prim_contains
 * computes it.  Planes: the same row values.  A sphere or cylinder: the sign
 * of prim_gmax, except within a band of 1e-5 of the magnitudes involved
 * around the surface -- far wider than the few ulps by which the two
 * formulas can round apart -- where prim_contains itself decides. */
MD float prim_gmax_in(const DevScene& S, const MONSEL_CS DPrim& pr, float px, float py, float pz, bool& in) {
  if (pr.kind == PRIM_PLANES) {
    float g = -kInf;
    bool all = true;
    for (int k = 0; k < pr.n_planes; ++k) {
      const float4 pl = prim_plane(S, pr, k);
      const float v = pl.x * px + pl.y * py + pl.z * pz - pl.w;
      g = fmaxf(g, v);
      all = all && v <= 0.f;
    }
    in = all;
    return g * pr.sdscale;
  }
  const float wx = px - pr.e0[0], wy = py - pr.e0[1], wz = pz - pr.e0[2];
  const float r = fsqrt(pr.r2);
  float g, scale;
  if (pr.kind == PRIM_CYLINDER) {
    const float apar = wx * pr.ax[0] + wy * pr.ax[1] + wz * pr.ax[2];
    const float qx = wx - apar * pr.ax[0], qy = wy - apar * pr.ax[1], qz = wz - apar * pr.ax[2];
    const float q = fsqrt(qx * qx + qy * qy + qz * qz);
    g = fmaxf(q - r, fmaxf(-apar, apar - pr.len));
    scale = q + fabsf(apar) + pr.len + r;
  } else {
    const float w = fsqrt(wx * wx + wy * wy + wz * wz);
    g = w - r;
    scale = w + r;
  }
  in = g <= 0.f;
  if (fabsf(g) <= 1e-5f * scale) in = prim_contains(S, pr, px, py, pz);
  return g;
}

MD bool shape_contains(const DevScene& S, const MONSEL_CS DShape& sh, float px, float py, float pz) {
  uint32_t bits = 0;
  for (int k = 0; k < sh.n_prims; ++k)
    if (prim_contains(S, S.prims[sh.first_prim + k], px, py, pz)) bits |= 1u << k;
  return truth(sh, bits);
}

struct ShapeHit {
  float t;      /* kInf if none within tmax */
  float nx, ny, nz;
};

/* First boundary of the composite shape along p + t d, 0 < t <= tmax.
 * Each primitive's entry/exit parameters go to per-lane LDS slots; the walk
 * then visits the crossings in increasing t, toggling the primitives'
 * inside bits, until the truth table changes value.  Requires
 * blockDim.x <= kLaneBlock. */
MD ShapeHit shape_first(const DevScene& S, const MONSEL_CS DShape& sh, float px, float py, float pz, float dx,
                        float dy, float dz, float tmax) {
  __shared__ float l_tin[kMaxPrims][kLaneBlock];
  __shared__ float l_tout[kMaxPrims][kLaneBlock];
  __shared__ int8_t l_fin[kMaxPrims][kLaneBlock];
  __shared__ int8_t l_fout[kMaxPrims][kLaneBlock];
  const int tid = threadIdx.x;
  ShapeHit h;
  h.t = kInf;
  h.nx = h.ny = h.nz = 0.f;
  uint32_t bits = 0;
  const int n = sh.n_prims;
  auto one = [&](int k) {
    Interval iv = prim_interval(S, S.prims[sh.first_prim + k], px, py, pz, dx, dy, dz);
    if (!(iv.tin < iv.tout)) {
      iv.tin = -kInf;
      iv.tout = -kInf;
    }
    l_tin[k][tid] = iv.tin;
    l_tout[k][tid] = iv.tout;
    l_fin[k][tid] = (int8_t)iv.fin;
    l_fout[k][tid] = (int8_t)iv.fout;
    if (iv.tin <= 0.f && iv.tout > 0.f) bits |= 1u << k;
  };
  /* the first four primitives unrolled, so that their records' loads and
   * interval arithmetic can overlap */
  PT_BEGIN(t_iv);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < n) one(k);
  for (int k = 4; k < n; ++k) one(k);
  PT_END(PH_GEOM_IV, t_iv);
  PT_BEGIN(t_wk);
  bool state = truth(sh, bits);
  float tcur = 0.f;
  for (int it = 0; it < 2 * n; ++it) {
    /* next crossing after tcur and the primitives crossing there */
    float tnext = kInf;
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < kMaxPrims; ++k) {
      if (k >= n) break;
      const float a = l_tin[k][tid], b = l_tout[k][tid];
      const float nk = a > tcur ? a : (b > tcur ? b : kInf);
      if (nk < tnext) {
        tnext = nk;
        mask = 1u << k;
      } else if (nk == tnext) {
        mask |= 1u << k;
      }
    }
    if (!(tnext <= tmax)) {
      PT_END(PH_GEOM_WALK, t_wk);
      return h;
    }
    bits ^= mask;
    const bool ns = truth(sh, bits);
    if (ns != state) {
      const float qx = px + tnext * dx, qy = py + tnext * dy, qz = pz + tnext * dz;
      const int first = __ffs(mask) - 1;
      const uint32_t rest = mask & (mask - 1u);
      const int f1 = l_tin[first][tid] == tnext ? l_fin[first][tid] : l_fout[first][tid];
      float nx, ny, nz;
      prim_normal(S, S.prims[sh.first_prim + first], f1, qx, qy, qz, nx, ny, nz);
      if (rest) {
        /* simultaneous crossings: mean of the two lowest-index normals
         * (NormalIntersectionShape.cu:229-236) */
        const int second = __ffs(rest) - 1;
        const int f2 = l_tin[second][tid] == tnext ? l_fin[second][tid] : l_fout[second][tid];
        float mx, my, mz;
        prim_normal(S, S.prims[sh.first_prim + second], f2, qx, qy, qz, mx, my, mz);
        nx = 0.5f * (nx + mx);
        ny = 0.5f * (ny + my);
        nz = 0.5f * (nz + mz);
      }
      h.t = tnext;
      h.nx = nx;
      h.ny = ny;
      h.nz = nz;
      PT_END(PH_GEOM_WALK, t_wk);
      return h;
    }
    state = ns;
    tcur = tnext;
  }
  PT_END(PH_GEOM_WALK, t_wk);
  return h;
}

/* RegionBase::containingSubRegion (RegionBase.cu:50-60) from region r. */
MD int containing_sub_region(const DevScene& S, int r, float px, float py, float pz) {
  if (!shape_contains(S, S.regions[r].sh, px, py, pz)) return -1;
  for (int guard = 0; guard < 64; ++guard) {
    const MONSEL_CS DRegion& rg = S.regions[r];
    int next = -1;
    for (int c = 0; c < rg.n_children; ++c) {
      int ch = S.children[rg.first_child + c];
      if (shape_contains(S, S.regions[ch].sh, px, py, pz)) {
        next = ch;
        break;
      }
    }
    if (next < 0) return r;
    r = next;
  }
  return r;
}

MD bool is_containing_region(const DevScene& S, int r, int target) {
  int a = S.regions[target].parent;
  while (a >= 0) {
    if (a == r) return true;
    a = S.regions[a].parent;
  }
  return false;
}

/* T'[S] = T[S ^ B] for the entries of an n-input truth table (n <= 5,
 * wave-uniform; B < 2^n) */
MD uint32_t xor_permute(uint32_t T, uint32_t B, int n) {
  if (B & 1u) T = ((T & 0x55555555u) << 1) | ((T >> 1) & 0x55555555u);
  if (n > 1 && (B & 2u)) T = ((T & 0x33333333u) << 2) | ((T >> 2) & 0x33333333u);
  if (n > 2 && (B & 4u)) T = ((T & 0x0F0F0F0Fu) << 4) | ((T >> 4) & 0x0F0F0F0Fu);
  if (n > 3 && (B & 8u)) T = ((T & 0x00FF00FFu) << 8) | ((T >> 8) & 0x00FF00FFu);
  if (n > 4 && (B & 16u)) T = (T << 16) | (T >> 16);
  return T;
}

/* Lower bound on the distance from p to the surface of shape sh.  Moving
 * less than the distance D_k to primitive k's surface cannot change k's
 * inside bit, so within distance D_k only the bits of the primitives with
 * D_j <= D_k can have flipped; the bound is the least D_k for which some
 * flip of those bits changes the shape's truth value.  Surfaces inside the
 * shape, or shared by primitives whose crossing does not matter, therefore
 * do not limit it.  With at most kMaxCstPrims primitives the test is on
 * bit masks: M[S] = truth(B ^ S) != truth(B) over the 32 flip sets S, and
 * the flip sets within a set F of primitives are the AND, over the
 * primitives b outside F, of the sets with bit b clear.  Larger shapes: the
 * nearest primitive surface. */
MD float shape_safety(const DevScene& S, const MONSEL_CS DShape& sh, float px, float py, float pz, bool& inside) {
  const int n = sh.n_prims;
  if (n > kMaxCstPrims) {
    float d = kInf;
    uint32_t bits = 0;
    for (int k = 0; k < n; ++k) {
      bool in;
      d = fminf(d, fabsf(prim_gmax_in(S, S.prims[sh.first_prim + k], px, py, pz, in)));
      if (in) bits |= 1u << k;
    }
    inside = truth(sh, bits);
    return d;
  }
  float d[kMaxCstPrims];
  uint32_t bits = 0;
#pragma unroll
  for (int k = 0; k < kMaxCstPrims; ++k) {
    d[k] = kInf;
    if (k < n) {
      bool in;
      d[k] = fabsf(prim_gmax_in(S, S.prims[sh.first_prim + k], px, py, pz, in));
      if (in) bits |= 1u << k;
    }
  }
  const uint32_t Tp = xor_permute(sh.truth[0], bits, n);
  inside = (Tp & 1u) != 0u;   /* bit 0 of the permuted table is truth(bits) */
  const uint32_t M = Tp ^ (0u - (Tp & 1u));
  constexpr uint32_t kClear[kMaxCstPrims] = {0x55555555u, 0x33333333u, 0x0F0F0F0Fu, 0x00FF00FFu, 0x0000FFFFu};
  /* flip sets of primitives beyond n never occur */
  uint32_t exist = 0xFFFFFFFFu;
#pragma unroll
  for (int b = 0; b < kMaxCstPrims; ++b)
    if (b >= n) exist &= kClear[b];
  float best = kInf;
  /* the loops run over the n primitives only (n is wave-uniform) */
#pragma unroll
  for (int k = 0; k < kMaxCstPrims; ++k) {
    if (k < n) {
      uint32_t within = exist;
#pragma unroll
      for (int b = 0; b < kMaxCstPrims; ++b)
        if (b < n && b != k && !(d[b] <= d[k])) within &= kClear[b];
      if (M & within) best = fminf(best, d[k]);
    }
  }
  return best;
}

/* Lower bound on the distance from p to the surfaces findEndOfStep tests
 * from region r: r's shape and its children's shapes.  A step from p shorter
 * than this cannot end at a boundary.  0 when p is outside r's own shape by
 * the containment test of the step's start (shape_contains): such an
 * electron is relocated before its next step, as the reference relocates it
 * (MonteCarloSS.cu:107-115), so a bound measured from outside the shape
 * never lets it skip that test. */
MD float region_safety(const DevScene& S, int r, float px, float py, float pz) {
  const MONSEL_CS DRegion& rg = S.regions[r];
  bool inside;
  float d = shape_safety(S, rg.sh, px, py, pz, inside);
  if (!inside) return 0.f;
  for (int c = 0; c < rg.n_children; ++c) {
    bool in_child;
    d = fminf(d, shape_safety(S, S.regions[S.children[rg.first_child + c]].sh, px, py, pz, in_child));
  }
  return d;
}

/* Result of RegionBase::findEndOfStep. */
struct StepEnd {
  float t;        /* distance moved (<= s) */
  int next;       /* next region; == current if no boundary; -1 left the chamber */
  bool boundary;
  float ownNx, ownNy, ownNz;   /* current shape's crossing normal (this step) */
  int winner;                  /* child region whose shape won, -1 = own shape */
  float winNx, winNy, winNz;
};

/* Displacement that carries a point across a boundary at coordinate scale |p|. */
MD float nudge_nm(float px, float py, float pz) {
  float m = fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz)));
  return 1.0e-3f + 4.0e-6f * m;
}

/* RegionBase::findEndOfStep (RegionBase.cu:74-119) on the ray p + t d,
 * 0 < t <= s.  A crossing of the chamber's own boundary leaves the chamber. */
MD StepEnd find_end_of_step(const DevScene& S, int r, float px, float py, float pz, float dx, float dy, float dz,
                            float s) {
  StepEnd se;
  const MONSEL_CS DRegion& rg = S.regions[r];
  float t = kInf;
  int base = r;
  se.winner = -1;
  /* c = -1: the region's own shape; c >= 0: its children, each limited to
   * the nearest crossing found so far */
  for (int c = -1; c < rg.n_children; ++c) {
    const int reg = c < 0 ? r : S.children[rg.first_child + c];
    ShapeHit h = shape_first(S, S.regions[reg].sh, px, py, pz, dx, dy, dz, c < 0 ? s : fminf(s, t));
    if (c < 0) {
      se.ownNx = se.winNx = h.nx;
      se.ownNy = se.winNy = h.ny;
      se.ownNz = se.winNz = h.nz;
      t = h.t;
      base = (t <= s && rg.parent >= 0) ? rg.parent : r;
    } else if (h.t <= s && h.t < t) {
      t = h.t;
      base = reg;
      se.winner = reg;
      se.winNx = h.nx;
      se.winNy = h.ny;
      se.winNz = h.nz;
    }
  }
  if (t <= s) {
    se.t = t;
    se.boundary = true;
    if (rg.parent < 0 && se.winner < 0) {
      se.next = -1; /* through the chamber wall */
      return se;
    }
    float qx = px + t * dx, qy = py + t * dy, qz = pz + t * dz;
    float dd = nudge_nm(qx, qy, qz);
    float ox = qx + dd * dx, oy = qy + dd * dy, oz = qz + dd * dz;
    int res = -1;
    PT_BEGIN(t_lo);
    for (int guard = 0; guard < 64 && base >= 0; ++guard) {
      res = containing_sub_region(S, base, ox, oy, oz);
      if (res >= 0) break;
      base = S.regions[base].parent;
    }
    PT_END(PH_GEOM_LOC, t_lo);
    se.next = res;
    return se;
  }
  se.t = s;
  se.boundary = false;
  se.next = r;
  return se;
}

}  // namespace monsel_dev

#endif
