// Shared host/device primitives of the hsmd library: the counter-based random
// stream (Philox4x32-10), the deterministic exponential used for every
// accept/reject decision, and the uniform-variate conversions.
//
// The reference draws uniforms from a global dSFMT-216091 state
// (Codes_Folder/subrouts.cpp:828-855) and visits sites in a time-seeded
// Fisher-Yates order (Codes_Folder/subrouts.cpp:2666-2699); neither stream is
// reproducible.  This library replaces both with a counter-based generator so
// that every GPU update can be replayed bit for bit on the CPU:
//   key     = 64-bit user seed
//   counter = (site label, sweep*16 + pass, replica/sample id, stream tag)
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HSMD_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define HSMD_HD static inline
#endif

namespace hsmd {

// Stream tags (fourth counter word).
enum : uint32_t {
  STREAM_MC_UPDATE = 0u,   // heat-bath / Metropolis site draws
  STREAM_PT = 1u,          // replica-exchange draws
  STREAM_INIT = 2u,        // initial-state draws
  STREAM_NOISE = 3u,       // SMD sample noise
};

struct u32x4 { uint32_t x, y, z, w; };

HSMD_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

// Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11).
HSMD_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32 x 32 -> 64-bit product per multiplier (a single v_mad_u64_u32
    // on the GPU instead of separate low and high multiplies)
    const uint64_t p0 = (uint64_t)M0 * (uint64_t)c.x, p1 = (uint64_t)M1 * (uint64_t)c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

HSMD_HD u32x4 draw4(uint64_t seed, uint32_t site, uint32_t step, uint32_t inst,
                    uint32_t stream) {
  u32x4 c = {site, step, inst, stream};
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 53-bit uniform in [0,1) from two 32-bit words.
HSMD_HD double u01_f64(uint32_t hi, uint32_t lo) {
  const uint64_t a = ((uint64_t)hi << 32) | (uint64_t)lo;
  return (double)(a >> 11) * 0x1.0p-53;
}

// 24-bit uniform in [0,1) from one 32-bit word.
HSMD_HD float u01_f32(uint32_t x) { return (float)(x >> 8) * 0x1.0p-24f; }

// Deterministic exp(x) for x <= 0, built only from IEEE-754 operations that
// round identically on every conforming machine (multiply, fma, rint, ldexp),
// so that GPU and CPU make identical accept/reject decisions.
// Method: Cody-Waite reduction x = k*ln2 + r with the two-part ln2 of fdlibm,
// then the degree-13 Taylor polynomial of e^r on |r| <= ln2/2 (truncation
// error below 2e-17), then exact scaling by 2^k.  Inputs below -700 return 0
// (every uniform draw is >= 0, so "u < 0" is never taken either way).
HSMD_HD double det_exp_neg(double x) {
#if defined(__HIPCC__)
#pragma clang fp contract(off)
#endif
  if (!(x <= 0.0)) return 1.0;
  if (x < -700.0) return 0.0;
  const double INV_LN2 = 1.44269504088896338700e+00;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double k = rint(x * INV_LN2);
  double r = fma(-k, LN2_HI, x);
  r = fma(-k, LN2_LO, r);
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// Metropolis / replica-exchange acceptance: u < min(1, exp(a)).
HSMD_HD bool accept_exp(double u, double a) { return u < det_exp_neg(a); }

}  // namespace hsmd
