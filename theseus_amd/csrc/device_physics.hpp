// Device-side gas model and numerical fluxes for the gfx950 DGSEM engine.
//
// Every function is templated on the arithmetic type R (double for the FP64
// path, float for the mixed path's flux arithmetic) and on the spatial
// dimension.  States are handed over as "point samples": the conservative
// vector plus the velocity and pressure already derived from it, computed
// once per node (in LDS) instead of once per pair as the reference does.
// The derived values use the reference's own operation order
// (GasState.hpp:169-172 velocity = m/rho, EOS.hpp:75-138 pressure =
// (gamma-1)(rhoE - 0.5*|m|^2/rho)), so a flux sees the same operands as the
// reference's per-pair recomputation.
#pragma once

#include <hip/hip_runtime.h>

#include "fastmath.hpp"

namespace thx {

template <typename R>
struct GasC {
  R gamma, gammaM1, gammaM1Inverse, cp, PrInverse, mu, mu_bulk, R_gas;
};

template <typename R, int DIM>
struct Pt {
  R q[DIM + 2];  // rho, m_0..m_{DIM-1}, rhoE
  R v[DIM];      // velocity m/rho
  R p;           // pressure
  R ir;          // 1/rho
  R pd;          // pressure minus a constant background, for the momentum flux's pressure term
};

// The momentum fluxes below take their pressure term from Pt::pd.  A
// constant background pressure contributes nothing to the divergence (the
// discretisation preserves free streams), so it may be left out of every
// momentum flux; in FP64 pd == p (no background), in the mixed mode pd is
// p - p_ref formed in FP64 before rounding to FP32, which keeps the O(1)
// pressure fluctuation of a low-Mach flow from drowning in the rounding of
// a pressure 1/(gamma Ma^2) times larger.

// Square root: IEEE sqrt in FP64; in the mixed mode's FP32 the hardware
// v_sqrt_f32 (1 ulp), one instruction for the ~10 of the correctly rounded
// form.  (A reciprocal-square-root + Newton form in FP64 cost the FP64
// kernel a scratch spill at its 4-wave register budget.)
template <typename R>
__device__ __forceinline__ R dsqrt(R x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R>
__device__ __forceinline__ R dlog(R x) { return ln_pos(x); }
template <>
__device__ __forceinline__ float dlog<float>(float x) { return logf(x); }
// Reciprocal of a normal, finite x: hardware estimate refined by two Newton
// steps r <- r + r (1 - x r) (quadratic convergence, result within ~1 ulp),
// 5 VALU instead of the 19 of an IEEE-rounded division.
template <typename R>
__device__ __forceinline__ R drcp(R x) {
  R r = __builtin_amdgcn_rcp(x);
  R e = fma(-x, r, R(1));
  r = fma(r, e, r);
  e = fma(-x, r, R(1));
  return fma(r, e, r);
}
template <>
__device__ __forceinline__ float drcp<float>(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}
template <typename R>
__device__ __forceinline__ R dabs(R x) { return fabs(x); }
template <>
__device__ __forceinline__ float dabs<float>(float x) { return fabsf(x); }
template <typename R>
__device__ __forceinline__ R dmax(R a, R b) { return a > b ? a : b; }
template <typename R>
__device__ __forceinline__ R dmin(R a, R b) { return a < b ? a : b; }

// Series threshold of the logarithmic mean: the reference's FP64 value
// (theseus_kernels.hpp:106-123) for double; for float the series is used up
// to u < 1e-2, where the truncation (~u^4/9 ~ 1e-9) stays below FP32 epsilon
// and the log branch's cancellation error (eps/sqrt(u) ~ 1e-6) is avoided.
template <typename R>
struct LogMeanEps {
  static constexpr double value = 1e-4;
};
template <>
struct LogMeanEps<float> {
  static constexpr double value = 1e-2;
};

// Logarithmic mean and its reciprocal for the volume fluxes: log_mean_pair
// (fastmath.hpp), the atanh series for moderate ratios and the logarithm
// beyond, equal to the reference's formula (theseus_kernels.hpp:106-123) to
// rounding.
template <typename R>
using LogMean = LogMeanPair<R>;

template <typename R, bool NEED_LM, bool NEED_INV>
__device__ __forceinline__ LogMean<R> log_mean2(R x, R y) {
  return log_mean_pair<R, NEED_LM, NEED_INV>(x, y);
}

// Logarithmic mean in the reference's exact formulation (kept for the
// face fluxes, where it costs little).
template <typename R>
__device__ __forceinline__ R log_mean(R x, R y) {
  const R xi = y / x;
  const R u = (xi * (xi - R(2)) + R(1)) / (xi * (xi + R(2)) + R(1));
  if (u < R(LogMeanEps<R>::value)) {
    const R denom = R(105) + u * (R(35) + u * (R(21) + R(15) * u));
    return (x + y) * R(52.5) / denom;
  }
  return (y - x) / dlog(xi);
}

// Derive velocity and pressure from the conservative vector, as the
// reference's accessors do (GasState.hpp:169-172, EOS.hpp:75-138), with the
// divisions by rho replaced by one reciprocal.
template <typename R, int DIM>
__device__ __forceinline__ void derive(const GasC<R> &g, Pt<R, DIM> &s) {
  s.ir = drcp(s.q[0]);
  R m2 = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R m = s.q[1 + d];
    m2 += m * m;
  }
  const R ke = R(0.5) * m2 * s.ir;
  s.p = g.gammaM1 * (s.q[DIM + 1] - ke);
  s.pd = s.p;
#pragma unroll
  for (int d = 0; d < DIM; ++d) s.v[d] = s.q[1 + d] * s.ir;
}

template <typename R, int DIM>
__device__ __forceinline__ R sound_speed(const GasC<R> &g, const Pt<R, DIM> &s) {
  return dsqrt(g.gamma * s.p * s.ir);
}

// ---------------------------------------------------------------------------
// Two-point volume fluxes along the (mean) metric row `met`.
// ---------------------------------------------------------------------------

// Pressure term of Chandrashekar's momentum flux, p_hat = (rho1+rho2)/(2(beta1+beta2))
// = (beta1 p1 + beta2 p2)/(beta1 + beta2) with beta = rho/(2p); the second
// form carries a background shift through unchanged, so the mixed mode
// evaluates it on pd.  FP64 keeps the reference's expression.
template <typename R, int DIM>
__device__ __forceinline__ R chandrashekar_p_mom(const Pt<R, DIM> &a, const Pt<R, DIM> &b, R beta1, R beta2, R p_hat) {
  if constexpr (sizeof(R) == sizeof(double)) {
    return p_hat;
  } else {
    return (beta1 * a.pd + beta2 * b.pd) / (beta1 + beta2);
  }
}

// ChandrashekarFlux.hpp:20-89
template <typename R, int DIM>
__device__ __forceinline__ void vol_chandrashekar(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                                  const R *met, R *F) {
  const R rho1 = a.q[0], rho2 = b.q[0];
  const LogMean<R> lr = log_mean2<R, true, true>(rho1, rho2);
  const R rho_ln = lr.lm;
  R mom_hat[DIM];
  R h_hat = R(0), vn = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R v1 = a.v[d], v2 = b.v[d];
    const R vbar = R(0.5) * (v1 + v2);
    vn += vbar * met[d];
    mom_hat[d] = rho_ln * vbar;
    h_hat += -R(0.25) * (v1 * v1 + v2 * v2) + vbar * vbar;
  }
  const R beta1 = R(0.5) * rho1 / a.p;
  const R beta2 = R(0.5) * rho2 / b.p;
  const R inv_beta_ln = log_mean2<R, false, true>(beta1, beta2).inv;
  const R p_hat = R(0.5) * (rho1 + rho2) / (beta1 + beta2);
  const R gm1_av_inv = g.gammaM1Inverse;  // 2/(gamma1 + gamma2 - 2) with one gamma
  h_hat += R(0.5) * inv_beta_ln * gm1_av_inv + p_hat * lr.inv;
  F[0] = rho_ln * vn;
  const R p_mom = chandrashekar_p_mom(a, b, beta1, beta2, p_hat);
#pragma unroll
  for (int d = 0; d < DIM; ++d) F[1 + d] = vn * mom_hat[d] + p_mom * met[d];
  F[DIM + 1] = rho_ln * vn * h_hat;
}

// RanochaFlux.hpp:20-79
template <typename R, int DIM>
__device__ __forceinline__ void vol_ranocha(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b, const R *met,
                                            R *F) {
  const R rho1 = a.q[0], rho2 = b.q[0];
  const R rho_ln = log_mean2<R, true, false>(rho1, rho2).lm;
  const R p1 = a.p, p2 = b.p;
  const R pdbar = R(0.5) * (a.pd + b.pd);
  R vbar[DIM], ppm[DIM];
  R ke_flux = R(0), vn = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R v1 = a.v[d], v2 = b.v[d];
    vbar[d] = R(0.5) * (v1 + v2);
    vn += vbar[d] * met[d];
    ke_flux += R(0.5) * (v1 * v2);
    ppm[d] = R(0.5) * (p1 * v2 + p2 * v1);
  }
  const R inv_beta_ln = (p1 * p2) * log_mean2<R, false, true>(rho1 * p2, rho2 * p1).inv;
  const R inv_gm1 = g.gammaM1Inverse;
  const R ie_flux = inv_beta_ln * inv_gm1;
  R fe = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    F[1 + d] = rho_ln * vn * vbar[d] + pdbar * met[d];
    fe += ppm[d] * met[d];
  }
  F[0] = rho_ln * vn;
  F[DIM + 1] = fe + rho_ln * (ie_flux + ke_flux) * vn;
}

// PirozzoliFlux.hpp:19-71
template <typename R, int DIM>
__device__ __forceinline__ void vol_pirozzoli(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                              const R *met, R *F) {
  const R rho1 = a.q[0], rho2 = b.q[0];
  const R rho_mean = R(0.5) * (rho1 + rho2);
  const R p1 = a.p, p2 = b.p;
  R vbar[DIM];
  R ke_mean = R(0), vn = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R v1 = a.v[d], v2 = b.v[d];
    vbar[d] = R(0.5) * (v1 + v2);
    vn += vbar[d] * met[d];
    ke_mean += R(0.25) * ((v1 * v1) + (v2 * v2));
  }
  // specific internal energy e = (rhoE - ke)/rho, ke recomputed as the reference does
  R m2a = R(0), m2b = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    m2a += a.q[1 + d] * a.q[1 + d];
    m2b += b.q[1 + d] * b.q[1 + d];
  }
  const R e1 = (a.q[DIM + 1] - R(0.5) * m2a * a.ir) * a.ir;
  const R e2 = (b.q[DIM + 1] - R(0.5) * m2b * b.ir) * b.ir;
  const R h_mean = R(0.5) * (e1 + p1 * a.ir + e2 + p2 * b.ir);
  F[0] = rho_mean * vn;
#pragma unroll
  for (int d = 0; d < DIM; ++d) F[1 + d] = rho_mean * vn * vbar[d] + R(0.5) * (a.pd + b.pd) * met[d];
  F[DIM + 1] = rho_mean * (h_mean + ke_mean) * vn;
}

// Central flux 0.5*(f(q1)+f(q2)).met (HLLFlux.hpp:19-51, RoeFlux.hpp:18-45,
// LaxFriedrichsFlux.hpp:20-52), physical flux as NavierStokesFlux.hpp:339-387.
template <typename R, int DIM>
__device__ __forceinline__ void vol_central(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b, const R *met,
                                            R *F) {
  const R sv1 = a.ir, sv2 = b.ir;
  const R H1 = (a.q[DIM + 1] + a.p) * sv1, H2 = (b.q[DIM + 1] + b.p) * sv2;
  R f[DIM + 2];
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) f[e] = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R m1d = a.q[1 + d], m2d = b.q[1 + d];
    f[0] += R(0.5) * (m1d + m2d) * met[d];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      R f1 = a.q[1 + i] * m1d * sv1, f2 = b.q[1 + i] * m2d * sv2;
      if (i == d) {
        f1 += a.pd;
        f2 += b.pd;
      }
      f[1 + i] += R(0.5) * (f1 + f2) * met[d];
    }
    f[DIM + 1] += R(0.5) * (m1d * H1 + m2d * H2) * met[d];
  }
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) F[e] = f[e];
}

// ---------------------------------------------------------------------------
// Face fluxes with a non-unit normal n (magnitude = face area scale).
// ---------------------------------------------------------------------------

// Physical normal flux of one state (NavierStokesFlux.hpp:274-332).  FP64
// follows the reference's sums term by term; the mixed mode's FP32 form
// contracts with v.n first (m_c (v.n) + p n_c, (E + p)(v.n)): the same flux
// in 14 operations instead of ~40, its rounding far below the mode's FP32
// tolerance.
template <typename R, int DIM>
__device__ __forceinline__ void normal_flux(const Pt<R, DIM> &s, const R *n, R *fn) {
  if constexpr (sizeof(R) == sizeof(float)) {
    R mn = R(0), vn = R(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      mn += s.q[1 + d] * n[d];
      vn += s.v[d] * n[d];
    }
    fn[0] = mn;
#pragma unroll
    for (int c = 0; c < DIM; ++c) fn[1 + c] = s.q[1 + c] * vn + s.pd * n[c];
    fn[DIM + 1] = (s.q[DIM + 1] + s.p) * vn;
    return;
  }
  const R inv_rho = s.ir;
  const R H = (s.q[DIM + 1] + s.p) * inv_rho;
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) fn[e] = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R m = s.q[1 + d];
    fn[0] += m * n[d];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      R pf = s.q[1 + c] * m * inv_rho;
      if (c == d) pf += s.pd;
      fn[1 + c] += pf * n[d];
    }
    fn[DIM + 1] += m * H * n[d];
  }
}

// DissipationOperator.hpp:174-232
template <typename R, int DIM>
__device__ __forceinline__ void scalar_dissipation(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                                   const R *nor, R *diss) {
  const R rho1 = a.q[0], rho2 = b.q[0];
  const R rho_mean = R(0.5) * (rho1 + rho2);
  const R drho = rho2 - rho1;
  R vn1 = R(0), vn2 = R(0), nor_mag = R(0), de = R(0);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    nor_mag += nor[d] * nor[d];
    const R v1 = a.v[d], v2 = b.v[d];
    const R dv = v2 - v1;
    const R vbar = R(0.5) * (v1 + v2);
    vn1 += v1 * nor[d];
    vn2 += v2 * nor[d];
    de += R(0.5) * drho * v1 * v2 + rho_mean * dv * vbar;
  }
  nor_mag = dsqrt(nor_mag);
  const R lambda_max =
      dmax(dabs(vn1) + sound_speed(g, a) * nor_mag, dabs(vn2) + sound_speed(g, b) * nor_mag);
  const R gm1_av_inv = R(1) / (g.gamma - R(1));
  const R beta1 = R(0.5) * rho1 / a.p;
  const R beta2 = R(0.5) * rho2 / b.p;
  const R beta_ln = log_mean(beta1, beta2);
  de += R(0.5) * drho * gm1_av_inv / beta_ln + R(0.5) * rho_mean * gm1_av_inv * (R(1) / beta2 - R(1) / beta1);
  diss[0] = R(0.5) * lambda_max * drho;
#pragma unroll
  for (int d = 0; d < DIM; d++) diss[1 + d] = R(0.5) * lambda_max * (b.q[1 + d] - a.q[1 + d]);
  diss[DIM + 1] = de * (R(0.5) * lambda_max);
}

// DissipationOperator.hpp:57-172 (Harten-Hyman entropy fix)
// The gas enters through gm() = gamma of the first state and cs(s, side) =
// the sound speed of state s (side 0 for a, 1 for b): the ideal gas's
// uniform gamma and sqrt(gamma p/rho), or the tabulated gas's cp/cv and
// table sound speed.  Callables keep the ideal gas's evaluation order.
template <typename R, int DIM, typename GM, typename CS>
__device__ __forceinline__ void roe_dissipation_core(const Pt<R, DIM> &a, const Pt<R, DIM> &b, GM gm, CS cs,
                                                     const R *nor, R *diss) {
  const R rho1 = a.q[0], rho2 = b.q[0];
  const R H1 = (a.q[DIM + 1] + a.p) * a.ir;
  const R H2 = (b.q[DIM + 1] + b.p) * b.ir;
  const R ratio = dsqrt(rho2 * a.ir);
  const R ca = R(1) / (ratio + R(1));
  const R cb = ca * ratio;
  const R roe_rho = ratio * rho1;
  const R roe_H = ca * H1 + cb * H2;
  R roe_v[DIM], del_v[DIM];
  R ke = R(0), vn1 = R(0), vn2 = R(0), vn_roe = R(0), nor_mag = R(0);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    nor_mag += nor[d] * nor[d];
    const R v1 = a.v[d], v2 = b.v[d];
    del_v[d] = v2 - v1;
    roe_v[d] = ca * v1 + cb * v2;
    ke += R(0.5) * roe_v[d] * roe_v[d];
    vn1 += v1 * nor[d];
    vn2 += v2 * nor[d];
    vn_roe += roe_v[d] * nor[d];
  }
  nor_mag = dsqrt(nor_mag);
  const R inv_nor_mag = R(1) / nor_mag;
  vn1 *= inv_nor_mag;
  vn2 *= inv_nor_mag;
  vn_roe *= inv_nor_mag;
  const R c_roe = dsqrt((gm() - R(1)) * (roe_H - ke));
  const R c1 = cs(a, 0), c2 = cs(b, 1);
  R eig0, eig2;
  {
    const R er = vn_roe - c_roe;
    const R eps = dmax(R(0), dmax(er - (vn1 - c1), (vn2 - c2) - er));
    eig0 = (dabs(er) < eps) ? R(0.5) * (eps + (er * er) / eps) : dabs(er);
  }
  {
    const R er = vn_roe + c_roe;
    const R eps = dmax(R(0), dmax(er - (vn1 + c1), (vn2 + c2) - er));
    eig2 = (dabs(er) < eps) ? R(0.5) * (eps + (er * er) / eps) : dabs(er);
  }
  const R eig1 = dabs(vn_roe);
  const R del_rho = rho2 - rho1;
  const R del_vn = vn2 - vn1;
  const R del_p = b.p - a.p;
  const R ov_csq = R(1) / (c_roe * c_roe);
  const R pov_csq = del_p * ov_csq;
  const R rhoc_del_vn = roe_rho * del_vn * c_roe * ov_csq;
  const R alpha_Vn = eig1 * (del_rho - pov_csq);
  const R alpha_Vn_rho = eig1 * roe_rho;
  const R alpha_left = R(0.5) * (eig0 * (pov_csq - rhoc_del_vn));
  const R alpha_right = R(0.5) * (eig2 * (pov_csq + rhoc_del_vn));
  diss[0] = R(0.5) * (alpha_Vn + alpha_left + alpha_right);
  R de = R(0);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    R dm = R(0.5) * (alpha_Vn * roe_v[d] + alpha_Vn_rho * (del_v[d] - nor[d] * inv_nor_mag * del_vn));
    dm += R(0.5) * (alpha_left * (roe_v[d] - nor[d] * inv_nor_mag * c_roe) +
                    alpha_right * (roe_v[d] + nor[d] * inv_nor_mag * c_roe));
    diss[1 + d] = dm;
    de += R(0.5) * alpha_Vn_rho * (roe_v[d] * del_v[d]);
  }
  de -= R(0.5) * alpha_Vn_rho * vn_roe * del_vn;
  de += R(0.5) * alpha_Vn * ke + R(0.5) * alpha_left * (roe_H - vn_roe * c_roe) +
        R(0.5) * alpha_right * (roe_H + vn_roe * c_roe);
  diss[DIM + 1] = de;
#pragma unroll
  for (int e = 0; e < DIM + 2; e++) diss[e] *= nor_mag;
}

template <typename R, int DIM>
__device__ __forceinline__ void roe_dissipation(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                                const R *nor, R *diss) {
  roe_dissipation_core(
      a, b, [&] { return g.gamma; }, [&](const Pt<R, DIM> &s, int) { return sound_speed(g, s); }, nor, diss);
}

// ChandrashekarFlux.hpp:91-155 (EC + scalar dissipation)
template <typename R, int DIM>
__device__ __forceinline__ void face_chandrashekar(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                                   const R *nor, R *F) {
  const R rho1 = a.q[0], rho2 = b.q[0];
  const R rho_ln = log_mean(rho1, rho2);
  R mom[DIM];
  R hhat = R(0), vn = R(0);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const R v1 = a.v[d], v2 = b.v[d];
    const R vbar = R(0.5) * (v1 + v2);
    vn += vbar * nor[d];
    mom[d] = rho_ln * vbar;
    hhat += -R(0.25) * (v1 * v1 + v2 * v2) + vbar * vbar;
  }
  const R beta1 = R(0.5) * rho1 / a.p;
  const R beta2 = R(0.5) * rho2 / b.p;
  const R beta_ln = log_mean(beta1, beta2);
  const R p_hat = R(0.5) * (rho1 + rho2) / (beta1 + beta2);
  const R gm1_av_inv = R(2) / (g.gamma + g.gamma - R(2));
  hhat += R(0.5) / beta_ln * gm1_av_inv + p_hat / rho_ln;
  R diss[DIM + 2];
  scalar_dissipation(g, a, b, nor, diss);
  F[0] = rho_ln * vn - diss[0];
  const R p_mom = chandrashekar_p_mom(a, b, beta1, beta2, p_hat);
#pragma unroll
  for (int d = 0; d < DIM; d++) F[1 + d] = vn * mom[d] + p_mom * nor[d] - diss[1 + d];
  F[DIM + 1] = rho_ln * vn * hhat - diss[DIM + 1];
}

// RanochaFlux.hpp:81-134 (EC + Roe dissipation)
template <typename R, int DIM>
__device__ __forceinline__ void face_ranocha(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                             const R *nor, R *F) {
  R ec[DIM + 2], diss[DIM + 2];
  vol_ranocha<R, DIM>(g, a, b, nor, ec);
  roe_dissipation(g, a, b, nor, diss);
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) F[e] = ec[e] - diss[e];
}

// PirozzoliFlux.hpp:73-120 (KEP + Roe dissipation)
template <typename R, int DIM, typename GM, typename CS>
__device__ __forceinline__ void face_pirozzoli_core(const Pt<R, DIM> &a, const Pt<R, DIM> &b, GM gm, CS cs,
                                                    const R *nor, R *F) {
  R ec[DIM + 2], diss[DIM + 2];
  vol_pirozzoli<R, DIM>(GasC<R>{}, a, b, nor, ec);
  roe_dissipation_core(a, b, gm, cs, nor, diss);
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) F[e] = ec[e] - diss[e];
}
template <typename R, int DIM>
__device__ __forceinline__ void face_pirozzoli(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                               const R *nor, R *F) {
  face_pirozzoli_core(
      a, b, [&] { return g.gamma; }, [&](const Pt<R, DIM> &s, int) { return sound_speed(g, s); }, nor, F);
}

// RoeFlux.hpp:283-390 (no entropy fix)
template <typename R, int DIM>
__device__ __forceinline__ void face_roe(const GasC<R> &g, const Pt<R, DIM> &L, const Pt<R, DIM> &Rs, const R *nor,
                                         R *F) {
  R nor_mag2 = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) nor_mag2 += nor[d] * nor[d];
  const R nor_mag = dsqrt(nor_mag2);
  const R inv_nor_mag = R(1) / nor_mag;
  const R rhoL = L.q[0], rhoR = Rs.q[0];
  const R pL = L.p, pR = Rs.p;
  const R rootL = dsqrt(rhoL), rootR = dsqrt(rhoR);
  const R root_sum_inv = R(1) / (rootL + rootR);
  const R rho_roe = rootL * rootR;
  const R HL = (L.q[DIM + 1] + pL) * L.ir;
  const R HR = (Rs.q[DIM + 1] + pR) * Rs.ir;
  const R H = (rootL * HL + rootR * HR) * root_sum_inv;
  R u[DIM], du_t[DIM], nhat[DIM], du[DIM];
  R u2 = R(0), un = R(0), dun = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R uL = L.v[d], uR = Rs.v[d];
    nhat[d] = nor[d] * inv_nor_mag;
    u[d] = (rootL * uL + rootR * uR) * root_sum_inv;
    du[d] = uR - uL;
    u2 += u[d] * u[d];
    un += u[d] * nhat[d];
    dun += du[d] * nhat[d];
  }
  const R a2 = (g.gamma - R(1)) * (H - R(0.5) * u2);
  const R a = dsqrt(a2);
  const R dp = pR - pL;
  const R drho = rhoR - rhoL;
  const R alpha_minus = R(0.5) * (dp - rho_roe * a * dun) / a2;
  const R alpha_plus = R(0.5) * (dp + rho_roe * a * dun) / a2;
  const R alpha_zero = drho - dp / a2;
  const R lambda_minus = dabs(un - a) * nor_mag;
  const R lambda_zero = dabs(un) * nor_mag;
  const R lambda_plus = dabs(un + a) * nor_mag;
#pragma unroll
  for (int d = 0; d < DIM; ++d) du_t[d] = du[d] - dun * nhat[d];
  R diss[DIM + 2];
  diss[0] = lambda_minus * alpha_minus + lambda_zero * alpha_zero + lambda_plus * alpha_plus;
#pragma unroll
  for (int d = 0; d < DIM; ++d)
    diss[1 + d] = lambda_minus * alpha_minus * (u[d] - a * nhat[d]) +
                  lambda_zero * (alpha_zero * u[d] + rho_roe * du_t[d]) +
                  lambda_plus * alpha_plus * (u[d] + a * nhat[d]);
  R u_dot_du_t = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) u_dot_du_t += u[d] * du_t[d];
  diss[DIM + 1] = lambda_minus * alpha_minus * (H - a * un) +
                  lambda_zero * (R(0.5) * alpha_zero * u2 + rho_roe * u_dot_du_t) +
                  lambda_plus * alpha_plus * (H + a * un);
  R fl[DIM + 2], fr[DIM + 2];
  normal_flux(L, nor, fl);
  normal_flux(Rs, nor, fr);
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) F[e] = R(0.5) * (fl[e] + fr[e]) - R(0.5) * diss[e];
}

// HLLFlux.hpp:53-116
// cs(s, side): the sound speed of state s (side 0 for a, 1 for b)
template <typename R, int DIM, typename CS>
__device__ __forceinline__ void face_hll_core(const Pt<R, DIM> &a, const Pt<R, DIM> &b, CS cs, const R *nor, R *F) {
  R un1 = R(0), un2 = R(0), nor_mag2 = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    un1 += a.v[d] * nor[d];
    un2 += b.v[d] * nor[d];
    nor_mag2 += nor[d] * nor[d];
  }
  const R nor_mag = dsqrt(nor_mag2);
  const R c1 = cs(a, 0) * nor_mag;
  const R c2 = cs(b, 1) * nor_mag;
  const R sL1 = un1 - c1, sL2 = un2 - c2, sR1 = un1 + c1, sR2 = un2 + c2;
  const R sL = (sL1 < sL2) ? sL1 : sL2;
  const R sR = (sR1 > sR2) ? sR1 : sR2;
  R f1[DIM + 2], f2[DIM + 2];
  normal_flux(a, nor, f1);
  normal_flux(b, nor, f2);
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) {
    if (sL >= R(0)) F[e] = f1[e];
    else if (sR <= R(0)) F[e] = f2[e];
    else F[e] = (sR * f1[e] - sL * f2[e] + sL * sR * (b.q[e] - a.q[e])) / (sR - sL);
  }
}
template <typename R, int DIM>
__device__ __forceinline__ void face_hll(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b, const R *nor,
                                         R *F) {
  face_hll_core(a, b, [&](const Pt<R, DIM> &s, int) { return sound_speed(g, s); }, nor, F);
}

// LaxFriedrichsFlux.hpp:188-211 with MaximumNormalWaveSpeed (NavierStokesFlux.hpp:246-270)
template <typename R, int DIM, typename CS>
__device__ __forceinline__ void face_llf_core(const Pt<R, DIM> &a, const Pt<R, DIM> &b, CS cs, const R *nor, R *F) {
  R vd1 = R(0), vd2 = R(0), nn = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    vd1 += a.v[d] * nor[d];
    vd2 += b.v[d] * nor[d];
    nn += nor[d] * nor[d];
  }
  const R nrm = dsqrt(nn);
  const R lambda_max = dmax(dabs(vd1) + cs(a, 0) * nrm, dabs(vd2) + cs(b, 1) * nrm);
  R f1[DIM + 2], f2[DIM + 2];
  normal_flux(a, nor, f1);
  normal_flux(b, nor, f2);
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) F[e] = R(0.5) * (f1[e] + f2[e]) - R(0.5) * lambda_max * (b.q[e] - a.q[e]);
}
template <typename R, int DIM>
__device__ __forceinline__ void face_llf(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b, const R *nor,
                                         R *F) {
  face_llf_core(a, b, [&](const Pt<R, DIM> &s, int) { return sound_speed(g, s); }, nor, F);
}

// face_llf(a, b) evaluated from (own, oth) without ordering the operands:
// F(a, b) = (f(a) + f(b))/2 - (lambda/2)(q_b - q_a) with lambda symmetric, so
// with sg = +1 when own is a (the low side) and -1 when it is b,
// F = (f(own) + f(oth))/2 - (sg lambda/2)(q_oth - q_own): the sum commutes
// and the sign changes are exact, so the two elements of a face get the
// same bits, as with ordered operands.  Explicit fused multiply-adds fix the
// rounding on both sides.
// nrm = |nor|, computed once per face by the caller.
template <typename R, int DIM>
__device__ __forceinline__ void face_llf_signed(const GasC<R> &g, const Pt<R, DIM> &own, const Pt<R, DIM> &oth,
                                                const R *nor, R nrm, R sg, R *F) {
  R vd1 = R(0), vd2 = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    vd1 += own.v[d] * nor[d];
    vd2 += oth.v[d] * nor[d];
  }
  const R lambda_max = dmax(dabs(vd1) + sound_speed(g, own) * nrm, dabs(vd2) + sound_speed(g, oth) * nrm);
  const R hl = sg * (R(0.5) * lambda_max);
  R f1[DIM + 2], f2[DIM + 2];
  normal_flux(own, nor, f1);
  normal_flux(oth, nor, f2);
#pragma unroll
  for (int e = 0; e < DIM + 2; ++e) {
    const R avg = R(0.5) * (f1[e] + f2[e]);
    F[e] = std::fma(-hl, oth.q[e] - own.q[e], avg);
  }
}

// Compile-time flux selection (ids of include/theseus_rhs.h)
template <int VOL, typename R, int DIM>
__device__ __forceinline__ void volume_flux(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b, const R *met,
                                            R *F) {
  if constexpr (VOL == 0) vol_central<R, DIM>(g, a, b, met, F);
  else if constexpr (VOL == 1) vol_chandrashekar<R, DIM>(g, a, b, met, F);
  else if constexpr (VOL == 2) vol_ranocha<R, DIM>(g, a, b, met, F);
  else vol_pirozzoli<R, DIM>(g, a, b, met, F);
}

template <int SURF, typename R, int DIM>
__device__ __forceinline__ void surface_flux(const GasC<R> &g, const Pt<R, DIM> &a, const Pt<R, DIM> &b,
                                             const R *nor, R *F) {
  if constexpr (SURF == 0) face_chandrashekar<R, DIM>(g, a, b, nor, F);
  else if constexpr (SURF == 1) face_ranocha<R, DIM>(g, a, b, nor, F);
  else if constexpr (SURF == 2) face_pirozzoli<R, DIM>(g, a, b, nor, F);
  else if constexpr (SURF == 3) face_roe<R, DIM>(g, a, b, nor, F);
  else if constexpr (SURF == 4) face_hll<R, DIM>(g, a, b, nor, F);
  else face_llf<R, DIM>(g, a, b, nor, F);
}

// ---------------------------------------------------------------------------
// BR1 helpers (EOS.hpp:159-263, NavierStokesFlux.hpp:389-519)
// ---------------------------------------------------------------------------

// Node quantities shared by the entropy variables, the gradient conversion
// and the viscous flux (EOS.hpp:75-263): 1/rho, velocity, kinetic and
// internal energy density, pressure.
template <typename R, int DIM>
struct NodeThermo {
  R rho, ir, u[DIM], ke, ie, p, beta;  // beta = rho / p
};

template <typename R, int DIM>
__device__ __forceinline__ NodeThermo<R, DIM> node_thermo(const GasC<R> &g, const R *q) {
  NodeThermo<R, DIM> t;
  t.rho = q[0];
  t.ir = drcp(q[0]);
  R m2 = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    m2 += q[1 + d] * q[1 + d];
    t.u[d] = q[1 + d] * t.ir;
  }
  t.ke = R(0.5) * m2 * t.ir;
  t.ie = q[DIM + 1] - t.ke;
  t.p = g.gammaM1 * t.ie;
  t.beta = t.rho * drcp(t.p);
  return t;
}

// Entropy variables (EOS.hpp:206-234):
// ((gamma - s)/(gamma-1) - beta |u|^2/2, beta u, -beta), s = ln p - gamma ln rho, beta = rho/p.
template <typename R, int DIM>
__device__ __forceinline__ void entropy_state(const GasC<R> &g, const R *q, R *E) {
  const NodeThermo<R, DIM> t = node_thermo<R, DIM>(g, q);
  const R s = dlog(t.p) - g.gamma * dlog(t.rho);
  const R beta = t.beta;
  const R v2o2 = t.ke * t.ir;
  E[0] = (g.gamma - s) * g.gammaM1Inverse - beta * v2o2;
#pragma unroll
  for (int d = 0; d < DIM; ++d) E[1 + d] = beta * t.u[d];
  E[DIM + 1] = -beta;
}

// EOS.hpp:236-263 : gradient of entropy variables -> gradient of (rho, u, p)
template <typename R, int DIM>
__device__ __forceinline__ void grad_entropy_to_grad_prim(const NodeThermo<R, DIM> &t, const R *q, const R *dE,
                                                          R *dP) {
  const R p_over_rho = t.p * t.ir;
  R drho = R(0);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    dP[1 + d] = p_over_rho * (dE[1 + d] + t.u[d] * dE[DIM + 1]);
    drho += q[1 + d] * dP[1 + d];
  }
  drho = t.rho * dE[0] - dE[DIM + 1] * (t.ke - t.ie) + t.beta * drho;
  dP[0] = drho;
  dP[DIM + 1] = p_over_rho * (dP[0] + t.p * dE[DIM + 1]);
}

// Number of stored viscous-flux components per node: the symmetric stress
// (DIM*(DIM+1)/2) and the energy flux (DIM).  Layout: tau_00, tau_11[, tau_22],
// tau_01[, tau_02, tau_12], e_0, e_1[, e_2].
template <int DIM>
struct VisComp {
  static constexpr int NTAU = DIM * (DIM + 1) / 2;
  static constexpr int N = NTAU + DIM;
  __host__ __device__ static constexpr int tau(int a, int b) {
    if (a == b) return a;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return DIM == 2 ? 2 : (lo == 0 ? (hi == 1 ? 3 : 4) : 5);
  }
  __host__ __device__ static constexpr int en(int d) { return NTAU + d; }
};

// Physical viscous flux, NavierStokesFlux.hpp:160-290 (Cartesian), from the
// transport coefficients, the velocity u, the temperature gradient gt and the
// primitive gradients dP[dir][]; writes the unique components in VisComp order.
template <typename R, int DIM>
__device__ __forceinline__ void viscous_flux_core(R mu, R kappa, R mu_bulk, const R *u, const R *gt,
                                                  const R (*dP)[DIM + 2], R *out) {
  R div_vel = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) div_vel += dP[d][1 + d];
  R tau[DIM][DIM];
#pragma unroll
  for (int dir = 0; dir < DIM; ++dir)
#pragma unroll
    for (int mom = 0; mom < DIM; ++mom) {
      if (mom == dir) tau[mom][dir] = mu * (R(2) * dP[dir][1 + mom] - mu_bulk * div_vel);
      else tau[mom][dir] = mu * (dP[dir][1 + mom] + dP[mom][1 + dir]);
    }
#pragma unroll
  for (int a = 0; a < DIM; ++a)
#pragma unroll
    for (int b = a; b < DIM; ++b) out[VisComp<DIM>::tau(a, b)] = tau[a][b];
#pragma unroll
  for (int dir = 0; dir < DIM; ++dir) {
    R ef = kappa * gt[dir];
#pragma unroll
    for (int mom = 0; mom < DIM; ++mom) ef += u[mom] * tau[mom][dir];
    out[VisComp<DIM>::en(dir)] = ef;
  }
}

// The ideal gas: constant viscosity, kappa = mu cp / Pr (Transport.hpp:19-54)
// and grad T from the pressure and density gradients (EOS.hpp:159-173,
// fac = (1/(gamma-1)) / (cv rho), cv = cp/gamma).  The same arithmetic as
// viscous_flux_core, written out: computing grad T and div u in one loop
// keeps the FP64 gradient kernel of the 3-D p = 3 case within its registers.
template <typename R, int DIM>
__device__ __forceinline__ void viscous_flux(const GasC<R> &g, const NodeThermo<R, DIM> &t, const R (*dP)[DIM + 2],
                                             R *out) {
  const R mu = g.mu;
  const R kappa = g.mu * g.cp * g.PrInverse;
  const R mu_bulk = g.mu_bulk;
  const R pressor = t.p * t.ir;
  const R fac = g.gammaM1Inverse * g.gamma / g.cp * t.ir;
  R gt[DIM];
  R div_vel = R(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    gt[d] = fac * (dP[d][DIM + 1] - pressor * dP[d][0]);
    div_vel += dP[d][1 + d];
  }
  R tau[DIM][DIM];
#pragma unroll
  for (int dir = 0; dir < DIM; ++dir)
#pragma unroll
    for (int mom = 0; mom < DIM; ++mom) {
      if (mom == dir) tau[mom][dir] = mu * (R(2) * dP[dir][1 + mom] - mu_bulk * div_vel);
      else tau[mom][dir] = mu * (dP[dir][1 + mom] + dP[mom][1 + dir]);
    }
#pragma unroll
  for (int a = 0; a < DIM; ++a)
#pragma unroll
    for (int b = a; b < DIM; ++b) out[VisComp<DIM>::tau(a, b)] = tau[a][b];
#pragma unroll
  for (int dir = 0; dir < DIM; ++dir) {
    R ef = kappa * gt[dir];
#pragma unroll
    for (int mom = 0; mom < DIM; ++mom) ef += t.u[mom] * tau[mom][dir];
    out[VisComp<DIM>::en(dir)] = ef;
  }
}

}  // namespace thx
