// Elementary functions written from their mathematics for the FP64 path.
// Host-compilable (plain C++ with <cmath>) so that tests can check them on
// the CPU.
#pragma once

#include <cmath>

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace thx {

using std::fma;
using std::frexp;

// 1/x for x in [1, 4]: an estimate refined by two Newton steps
// r <- r + r (1 - x r); on the GPU the estimate is the hardware reciprocal,
// on the host (where the tests run) a single-precision division, so both
// start from a relative error far below 2^-20 and end within an ulp.
__host__ __device__ __forceinline__ double recip_nr(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
#else
  double r = (double)(1.0f / (float)x);
#endif
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Natural logarithm of a positive, finite, normal double, from the
// mathematics (no library kernel): x = 2^e m with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(z), z = (m-1)/(m+1), |z| <= 3 - 2 sqrt(2) < 0.1716, summed as
// the Taylor series 2 z sum_{k>=0} z^{2k}/(2k+1) up to k = 11 (remainder
// < 0.0295^12/23 < 2e-20 relative), and e ln 2 added with ln 2 split into a
// head whose product with |e| < 2^11 is exact and a tail.  Error about
// 1 ulp (checked against 40-digit arithmetic on the CPU, tests/test_fastmath.py).
__host__ __device__ __forceinline__ double ln_pos(double x) {
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double f = m - 1.0;
  const double z = f * recip_nr(m + 1.0);  // its rounding only enters through z f below
  const double z2 = z * z;
  double s = 1.0 / 23.0;
  s = fma(s, z2, 1.0 / 21.0);
  s = fma(s, z2, 1.0 / 19.0);
  s = fma(s, z2, 1.0 / 17.0);
  s = fma(s, z2, 1.0 / 15.0);
  s = fma(s, z2, 1.0 / 13.0);
  s = fma(s, z2, 1.0 / 11.0);
  s = fma(s, z2, 1.0 / 9.0);
  s = fma(s, z2, 1.0 / 7.0);
  s = fma(s, z2, 1.0 / 5.0);
  s = fma(s, z2, 1.0 / 3.0);
  // ln m = 2z + 2z^3 s with 2z = f - z f: f = m - 1 is exact, so the rounding
  // of z only enters through the small product z f.
  const double lnm = fma((2.0 * z) * z2, s, fma(-z, f, f));
  const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000: 21 trailing zero bits
  const double ln2_lo = 1.90821492927058770002e-10;  // ln 2 - ln2_hi
  const double de = (double)e;
  return fma(de, ln2_hi, fma(de, ln2_lo, lnm));
}

}  // namespace thx
