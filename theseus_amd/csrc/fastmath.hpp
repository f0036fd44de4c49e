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

// Reciprocal refined to about an ulp: FP64 as recip_nr; FP32 from the
// hardware (device) or IEEE (host) reciprocal with one Newton step.
template <typename R>
__host__ __device__ __forceinline__ R recip_refined(R x);
template <>
__host__ __device__ __forceinline__ double recip_refined<double>(double x) {
  return recip_nr(x);
}
template <>
__host__ __device__ __forceinline__ float recip_refined<float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  float r = __builtin_amdgcn_rcpf(x);
#else
  float r = 1.0f / x;
#endif
  return std::fma(r, std::fma(-x, r, 1.0f), r);
}

// Logarithmic mean LM(x, y) = (y - x) / ln(y/x) of two positive numbers and
// its reciprocal, without a logarithm when the ratio is moderate.  With
// s = x + y, t = (y - x)/s and u = t^2, ln(y/x) = 2 atanh(t) = 2 t P(u),
// P(u) = sum_k u^k/(2k+1), so 1/LM = 2 P(u)/s and LM = (s/2)/P(u):
//   u <  eps    : the reference's series (theseus_kernels.hpp:106-123), P to
//                 u^3, with LM = (s/2) Q(u), Q = 1 - u/3 - 4u^2/45 - 44u^3/945
//                 the series of 1/P (one reciprocal gives both);
//   u <  0.0294 : (|t| < 0.1716, y/x within [0.707, 1.414]) P summed to
//                 u^11/23 in FP64 (remainder below 2e-20 relative), to u^5/11
//                 in FP32 (below 1e-10), LM from its refined reciprocal;
//   otherwise   : LM = (y - x)/ln(y/x), as the reference.
// eps is the reference's 1e-4 in FP64 and 1e-2 in FP32 (the truncation
// u^4/9 then stays below FP32 rounding).  The middle branch is the
// reference's log branch evaluated without the logarithm: it agrees with it
// to a few ulp, and keeps a wave in series arithmetic when a smooth flow
// puts some of its lanes just above eps.
template <typename R>
struct LogMeanPair {
  R lm, inv;
};

template <typename R, bool NEED_LM, bool NEED_INV>
__host__ __device__ __forceinline__ LogMeanPair<R> log_mean_pair(R x, R y) {
  constexpr bool F64 = sizeof(R) == sizeof(double);
  LogMeanPair<R> r{R(0), R(0)};
  const R s = x + y;
  const R d = y - x;
  if (d * d < R(F64 ? 1e-4 : 1e-2) * (s * s)) {
    const R rs = recip_refined<R>(s);
    const R t = d * rs;
    const R u = t * t;
    if constexpr (NEED_LM) r.lm = (R(0.5) * s) * (R(1) - u * (R(1.0 / 3.0) + u * (R(4.0 / 45.0) + u * R(44.0 / 945.0))));
    if constexpr (NEED_INV) r.inv = (R(2) * rs) * (R(1) + u * (R(1.0 / 3.0) + u * (R(0.2) + u * R(1.0 / 7.0))));
  } else if (d * d < R(0.0294) * (s * s)) {
    const R rs = recip_refined<R>(s);
    const R t = d * rs;
    const R u = t * t;
    R p;
    if constexpr (F64) {
      p = R(1.0 / 23.0);
      p = std::fma(p, u, R(1.0 / 21.0));
      p = std::fma(p, u, R(1.0 / 19.0));
      p = std::fma(p, u, R(1.0 / 17.0));
      p = std::fma(p, u, R(1.0 / 15.0));
      p = std::fma(p, u, R(1.0 / 13.0));
      p = std::fma(p, u, R(1.0 / 11.0));
    } else {
      p = R(1.0 / 11.0);
    }
    p = std::fma(p, u, R(1.0 / 9.0));
    p = std::fma(p, u, R(1.0 / 7.0));
    p = std::fma(p, u, R(1.0 / 5.0));
    p = std::fma(p, u, R(1.0 / 3.0));
    p = std::fma(p, u, R(1));
    if constexpr (NEED_INV) r.inv = (R(2) * rs) * p;
    if constexpr (NEED_LM) r.lm = (R(0.5) * s) * recip_refined<R>(p);
  } else {
    R l;
    if constexpr (F64) {
      l = ln_pos(y / x);
    } else {
#if defined(__HIP_DEVICE_COMPILE__)
      l = logf(y / x);
#else
      l = std::log(y / x);
#endif
    }
    if constexpr (NEED_LM) r.lm = d / l;
    if constexpr (NEED_INV) r.inv = l / d;
  }
  return r;
}

}  // namespace thx
