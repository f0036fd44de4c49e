// TEST INFRASTRUCTURE ONLY.  CPU FP64 restatement of the reference's
// (chess-uiuc/Theseus) gas model and numerical fluxes.  Only tests/, the
// smoke check and the bench's cpu_baseline leg may load this code, and only
// as the checker; the product path never links or calls it.
//
// Each function follows the reference file:line named in its comment, with
// the same operation order so that rounding matches the reference's own CPU
// build as closely as plain C++ allows.
//
// Parity pins (see tests/test_oracle_*.py): consistency F*(u,u,n)=f(u).n,
// normal reversal, zero-normal-velocity pressure flux, Tadmor's entropy
// condition, Roe stationary contact, Chandrashekar dissipation boost
// invariance and the independent tensor restatements of LLF/HLL/Roe
// (tests/test_riemann_fluxes.cpp:64-801).
#pragma once

#include <algorithm>
#include <cmath>

namespace oracle {

constexpr int MAXEQ = 5;
constexpr int MAXDIM = 3;

// theseus_kernels.hpp:106-123
inline double ComputeLogMean(double x, double y, double eps) {
  const double xi = y / x;
  const double u = (xi * (xi - 2.0) + 1.0) / (xi * (xi + 2.0) + 1.0);
  if (u < eps) {
    const double denom = 105.0 + u * (35.0 + u * (21.0 + 15.0 * u));
    return (x + y) * 52.5 / denom;
  } else {
    return (y - x) / std::log(xi);
  }
}

// PhysicsConstants (Physics.hpp:13-43) + IdealSingleGasEOS (EOS.hpp:19-324)
// + constant-viscosity Transport (Transport.hpp:19-54, SUTHERLAND off).
// State layout q = [rho, rho*u_0..rho*u_{dim-1}, rho*E] (GasState.hpp:36-202).
struct Gas {
  int dim;
  double gamma, gammaM1, gammaM1Inverse, gamma_gammaM1Inverse;
  double Pr, PrInverse, R_gas, cp, mu, mu_bulk;

  Gas() = default;
  Gas(int dim_, double gamma_, double Pr_, double R_, double mu_, double mu_bulk_)
      : dim(dim_), gamma(gamma_), Pr(Pr_), R_gas(R_), mu(mu_), mu_bulk(mu_bulk_) {
    gammaM1 = gamma - 1.0;
    gammaM1Inverse = 1.0 / gammaM1;
    gamma_gammaM1Inverse = gamma * gammaM1Inverse;
    PrInverse = 1.0 / Pr;
    cp = gamma_gammaM1Inverse * R_gas;
  }
  int neq() const { return dim + 2; }
  int eq_energy() const { return dim + 1; }

  double density(const double *q) const { return q[0]; }
  double momentum(const double *q, int d) const { return q[1 + d]; }
  double velocity(const double *q, int d) const { return q[1 + d] / q[0]; }  // GasState.hpp:169-172
  double energy(const double *q) const { return q[dim + 1]; }
  // EOS.hpp:75-97
  double momentum_sq(const double *q) const {
    double m2 = 0;
    for (int d = 0; d < dim; ++d) {
      const double m = q[1 + d];
      m2 += m * m;
    }
    return m2;
  }
  double kinetic_energy_density(const double *q) const { return 0.5 * momentum_sq(q) / q[0]; }
  double internal_energy_density(const double *q) const { return energy(q) - kinetic_energy_density(q); }
  double specific_internal_energy(const double *q) const { return internal_energy_density(q) / q[0]; }
  double pressure(const double *q) const { return gammaM1 * internal_energy_density(q); }  // EOS.hpp:130-138
  double temperature(const double *q) const { return pressure(q) / (q[0] * R_gas); }
  double sound_speed(const double *q) const {                                              // EOS.hpp:175-184
    const double rho = q[0];
    const double p = pressure(q);
    return std::sqrt(gamma * p / rho);
  }
  double gamma_at(const double *) const { return gamma; }  // uniform gamma
  double viscosity(const double *) const { return mu; }
  double bulk_viscosity(const double *) const { return mu_bulk; }
  double thermal_conductivity(const double *q) const { return viscosity(q) * cp * PrInverse; }  // Transport.hpp:92-98

  // EOS.hpp:159-173
  void grad_temperature(const double *q, const double *grad_rho, const double *grad_p, double *grad_t) const {
    const double rho = q[0];
    const double pressor = pressure(q) / rho;
    const double cv = cp / gamma;
    const double fac = gammaM1Inverse / (cv * rho);
    for (int i = 0; i < dim; i++) grad_t[i] = fac * (grad_p[i] - pressor * grad_rho[i]);
  }

  // EOS.hpp:206-234
  void entropy_state(const double *q, double *E) const {
    const double p = pressure(q);
    const double rho = q[0];
    const double s = std::log(p) - gamma * std::log(rho);
    const double beta = rho / p;
    const double v2o2 = kinetic_energy_density(q) / rho;
    const double s_rho = (gamma - s) / (gamma - 1) - beta * v2o2;
    E[0] = s_rho;
    for (int d = 0; d < dim; d++) E[1 + d] = beta * velocity(q, d);
    E[dim + 1] = -beta;
  }

  // EOS.hpp:236-263
  void grad_entropy_to_grad_prim(const double *q, const double *dE, double *dP) const {
    const double ke = kinetic_energy_density(q);
    const double p = pressure(q);
    const double rho = q[0];
    const double ie = internal_energy_density(q);
    double drho = 0.0;
    for (int d = 0; d < dim; d++) {
      dP[1 + d] = p / rho * (dE[1 + d] + velocity(q, d) * dE[dim + 1]);
      drho += momentum(q, d) * dP[1 + d];
    }
    drho = rho * dE[0] - dE[dim + 1] * (ke - ie) + rho * drho / p;
    dP[0] = drho;
    dP[dim + 1] = p / rho * (dP[0] + p * dE[dim + 1]);
  }
};

// ---------------------------------------------------------------------------
// Physical fluxes (NavierStokesFlux.hpp)
// ---------------------------------------------------------------------------

// NavierStokesFlux.hpp:274-291
struct NormalFluxState {
  const double *conserved;
  double inverse_density, pressure, total_enthalpy;
};
template <class G>
inline NormalFluxState MakeInviscidNormalFluxState(const G &g, const double *q) {
  const double inverse_density = 1.0 / g.density(q);
  const double pressure = g.pressure(q);
  return {q, inverse_density, pressure, (g.energy(q) + pressure) * inverse_density};
}

// NavierStokesFlux.hpp:295-332
template <class G>
inline double ComputeInviscidNormalFluxEquation(const G &g, const NormalFluxState &s, const double *normal,
                                                int equation) {
  double result = 0.0;
  for (int direction = 0; direction < g.dim; ++direction) {
    const double momentum = g.momentum(s.conserved, direction);
    double physical_flux = 0.0;
    if (equation == 0) {
      physical_flux = momentum;
    } else if (equation >= 1 && equation < 1 + g.dim) {
      const int component = equation - 1;
      physical_flux = g.momentum(s.conserved, component) * momentum * s.inverse_density;
      if (component == direction) physical_flux += s.pressure;
    } else if (equation == g.dim + 1) {
      physical_flux = momentum * s.total_enthalpy;
    }
    result += physical_flux * normal[direction];
  }
  return result;
}

// NavierStokesFlux.hpp:246-270
template <class G>
inline double MaximumNormalWaveSpeed(const G &g, const double *q1, const double *q2, const double *n) {
  double vd1 = 0.0, vd2 = 0.0, nn = 0.0;
  for (int d = 0; d < g.dim; ++d) {
    vd1 += g.velocity(q1, d) * n[d];
    vd2 += g.velocity(q2, d) * n[d];
    nn += n[d] * n[d];
  }
  const double nrm = std::sqrt(nn);
  return std::max(std::abs(vd1) + g.sound_speed(q1) * nrm, std::abs(vd2) + g.sound_speed(q2) * nrm);
}

// NavierStokesFlux.hpp:339-387
template <class G>
inline void ComputeInviscidFluxKernel(const G &g, const double *q, double F[MAXEQ][MAXDIM]) {
  const int dim = g.dim;
  const double density = g.density(q);
  const double spec_vol = 1.0 / density;
  double momentum[MAXDIM] = {0., 0., 0.};
  for (int d = 0; d < dim; d++) momentum[d] = g.momentum(q, d);
  const double energy = g.energy(q);
  const double pressure = g.pressure(q);
  const int eq_ener = dim + 1;
  const double H = (energy + pressure) * spec_vol;
  for (int d = 0; d < dim; d++) {
    F[0][d] = momentum[d];
    for (int i = 0; i < dim; i++) F[1 + i][d] = momentum[i] * momentum[d] * spec_vol;
    F[1 + d][d] += pressure;
    F[eq_ener][d] = momentum[d] * H;
  }
}

// NavierStokesFlux.hpp:160-290 (Cartesian: axisymmetric=false); G is the
// ideal gas or the tabulated one (oracle_lte.hpp TableGas)
template <class G>
inline void ComputeViscousFluxKernel(const G &g, const double *q, const double *dx, const double *dy,
                                     const double *dz, double vf[MAXEQ][MAXDIM]) {
  const int dim = g.dim;
  for (int qq = 0; qq < MAXEQ; qq++)
    for (int d = 0; d < dim; d++) vf[qq][d] = 0.0;
  const double mu = g.viscosity(q);
  const double kappa = g.thermal_conductivity(q);
  const double mu_bulk = g.bulk_viscosity(q);
  const int eq_ener = dim + 1;
  double grad_rho[MAXDIM] = {0, 0, 0}, grad_p[MAXDIM] = {0, 0, 0}, grad_t[MAXDIM] = {0, 0, 0};
  grad_rho[0] = dx[0];
  grad_p[0] = dx[eq_ener];
  if (dim > 1) {
    grad_rho[1] = dy[0];
    grad_p[1] = dy[eq_ener];
    if (dim > 2) {
      grad_rho[2] = dz[0];
      grad_p[2] = dz[eq_ener];
    }
  }
  g.grad_temperature(q, grad_rho, grad_p, grad_t);
  const double *rows[3] = {dx, dy, dz};
  double vel[MAXDIM] = {0, 0, 0};
  double div_vel = 0.0;
  for (int i = 0; i < dim; i++) {
    vel[i] = g.velocity(q, i);
    div_vel += rows[i][1 + i];
  }
  for (int dir = 0; dir < dim; ++dir) {
    const double *dg = rows[dir];
    for (int mom = 0; mom < dim; ++mom) {
      double tau;
      if (mom == dir) {
        tau = mu * (2.0 * dg[1 + mom] - mu_bulk * div_vel);
      } else {
        const double *cg = rows[mom];
        tau = mu * (dg[1 + mom] + cg[1 + dir]);
      }
      vf[1 + mom][dir] = tau;
    }
    double eflux = kappa * grad_t[dir];
    for (int mom = 0; mom < dim; mom++) eflux += vel[mom] * vf[1 + mom][dir];
    vf[eq_ener][dir] = eflux;
  }
}

// ---------------------------------------------------------------------------
// Two-point volume fluxes
// ---------------------------------------------------------------------------

// ChandrashekarFlux.hpp:20-89
inline void ChandrashekarVolume(const Gas &g, const double *q1, const double *q2, const double *met1,
                                const double *met2, double *F) {
  const int dim = g.dim;
  double met[3] = {0, 0, 0};
  for (int i = 0; i < dim; ++i) met[i] = 0.5 * (met1[i] + met2[i]);
  const double rho1 = g.density(q1), rho2 = g.density(q2);
  const double rho_ln = ComputeLogMean(rho1, rho2, 1e-4);
  double mom_hat[3] = {0, 0, 0};
  double h_hat = 0, vn = 0;
  for (int d = 0; d < dim; ++d) {
    const double v1 = g.velocity(q1, d);
    const double v2 = g.velocity(q2, d);
    const double vbar = 0.5 * (v1 + v2);
    vn += vbar * met[d];
    mom_hat[d] = rho_ln * vbar;
    h_hat += -0.25 * (v1 * v1 + v2 * v2) + vbar * vbar;
  }
  const double p1 = g.pressure(q1), p2 = g.pressure(q2);
  const double beta1 = 0.5 * rho1 / p1;
  const double beta2 = 0.5 * rho2 / p2;
  const double beta_ln = ComputeLogMean(beta1, beta2, 1e-4);
  const double p_hat = 0.5 * (rho1 + rho2) / (beta1 + beta2);
  const double gm11 = g.gamma, gm12 = g.gamma;
  const double gm1_av_inv = 2.0 / (gm11 + gm12 - 2.0);
  h_hat += 0.5 / beta_ln * gm1_av_inv + p_hat / rho_ln;
  F[0] = rho_ln * vn;
  for (int d = 0; d < dim; ++d) F[1 + d] = vn * mom_hat[d] + p_hat * met[d];
  F[dim + 1] = rho_ln * vn * h_hat;
}

// RanochaFlux.hpp:20-79
inline void RanochaVolume(const Gas &g, const double *q1, const double *q2, const double *met1, const double *met2,
                          double *F) {
  const int dim = g.dim;
  double met[3] = {0, 0, 0};
  for (int i = 0; i < dim; ++i) met[i] = 0.5 * (met1[i] + met2[i]);
  const double rho1 = g.density(q1), rho2 = g.density(q2);
  const double rho_ln = ComputeLogMean(rho1, rho2, 1e-4);
  double vbar[3] = {0, 0, 0}, p_prod_mean[3] = {0, 0, 0};
  double ke_flux = 0, vn = 0;
  const double p1 = g.pressure(q1), p2 = g.pressure(q2);
  const double pbar = 0.5 * (p1 + p2);
  for (int d = 0; d < dim; ++d) {
    const double v1 = g.velocity(q1, d);
    const double v2 = g.velocity(q2, d);
    vbar[d] = 0.5 * (v1 + v2);
    vn += vbar[d] * met[d];
    ke_flux += 0.5 * (v1 * v2);
    p_prod_mean[d] = 0.5 * (p1 * v2 + p2 * v1);
  }
  const double inv_beta_ln = (p1 * p2) / ComputeLogMean(rho1 * p2, rho2 * p1, 1e-4);
  const double inv_gm1 = 1.0 / (g.gamma - 1.0);
  const double ie_flux = inv_beta_ln * inv_gm1;
  F[0] = rho_ln * vn;
  F[dim + 1] = 0.0;
  for (int d = 0; d < dim; ++d) {
    F[1 + d] = rho_ln * vn * vbar[d] + pbar * met[d];
    F[dim + 1] += p_prod_mean[d] * met[d];
  }
  F[dim + 1] += rho_ln * (ie_flux + ke_flux) * vn;
}

// PirozzoliFlux.hpp:19-71
template <class G>
inline void PirozzoliVolume(const G &g, const double *q1, const double *q2, const double *met1, const double *met2,
                            double *F) {
  const int dim = g.dim;
  double met[3] = {0, 0, 0};
  for (int i = 0; i < dim; ++i) met[i] = 0.5 * (met1[i] + met2[i]);
  const double rho1 = g.density(q1), rho2 = g.density(q2);
  const double rho_mean = 0.5 * (rho1 + rho2);
  double vbar[3] = {0, 0, 0};
  double ke_mean = 0, vn = 0;
  const double p1 = g.pressure(q1), p2 = g.pressure(q2);
  const double pbar = 0.5 * (p1 + p2);
  for (int d = 0; d < dim; ++d) {
    const double v1 = g.velocity(q1, d);
    const double v2 = g.velocity(q2, d);
    vbar[d] = 0.5 * (v1 + v2);
    vn += vbar[d] * met[d];
    ke_mean += 0.25 * ((v1 * v1) + (v2 * v2));
  }
  const double h_mean =
      0.5 * (g.specific_internal_energy(q1) + p1 / rho1 + g.specific_internal_energy(q2) + p2 / rho2);
  F[0] = rho_mean * vn;
  for (int d = 0; d < dim; ++d) F[1 + d] = rho_mean * vn * vbar[d] + pbar * met[d];
  F[dim + 1] = rho_mean * (h_mean + ke_mean) * vn;
}

// HLLFlux.hpp:19-51 (identical to RoeFlux.hpp:18-45 and LaxFriedrichsFlux.hpp:20-52)
template <class G>
inline void CentralVolume(const G &g, const double *q1, const double *q2, const double *met1, const double *met2,
                          double *F) {
  const int dim = g.dim, neq = g.neq();
  double met[3] = {0, 0, 0};
  for (int i = 0; i < dim; ++i) met[i] = 0.5 * (met1[i] + met2[i]);
  double f1[MAXEQ][MAXDIM], f2[MAXEQ][MAXDIM];
  ComputeInviscidFluxKernel(g, q1, f1);
  ComputeInviscidFluxKernel(g, q2, f2);
  for (int e = 0; e < neq; e++) {
    double s = 0;
    for (int d = 0; d < dim; d++) s += 0.5 * (f1[e][d] + f2[e][d]) * met[d];
    F[e] = s;
  }
}

// ---------------------------------------------------------------------------
// Dissipation operators (DissipationOperator.hpp)
// ---------------------------------------------------------------------------

// DissipationOperator.hpp:57-172
template <class G>
inline void RoeDissipation(const G &g, const double *S1, const double *S2, const double *nor, double *diss) {
  const int dim = g.dim, neq = g.neq(), ener = dim + 1;
  double vn1 = 0.0, vn2 = 0.0, vn_roe = 0.0, nor_mag = 0.0;
  const double rho1 = g.density(S1), rho2 = g.density(S2);
  const double H1 = (g.energy(S1) + g.pressure(S1)) / rho1;
  const double H2 = (g.energy(S2) + g.pressure(S2)) / rho2;
  double roe[MAXEQ] = {0, 0, 0, 0, 0};
  const double ratio = std::sqrt(rho2 / rho1);
  const double a = 1.0 / (ratio + 1.0);
  const double b = a * ratio;
  roe[0] = ratio * rho1;
  roe[ener] = a * H1 + b * H2;
  double del_mom[3] = {0, 0, 0};
  double ke = 0.0;
  for (int d = 0; d < dim; d++) {
    nor_mag += nor[d] * nor[d];
    const double v1 = g.velocity(S1, d), v2 = g.velocity(S2, d);
    del_mom[d] = v2 - v1;
    roe[1 + d] = a * v1 + b * v2;
    ke += 0.5 * roe[1 + d] * roe[1 + d];
    vn1 += v1 * nor[d];
    vn2 += v2 * nor[d];
    vn_roe += roe[1 + d] * nor[d];
  }
  nor_mag = std::sqrt(nor_mag);
  const double inv_nor_mag = 1.0 / nor_mag;
  vn1 *= inv_nor_mag;
  vn2 *= inv_nor_mag;
  vn_roe *= inv_nor_mag;
  const double c_roe = std::sqrt((g.gamma_at(S1) - 1) * (roe[ener] - ke));
  double eig[3] = {0., std::abs(vn_roe), 0.};
  const double c1 = g.sound_speed(S1), c2 = g.sound_speed(S2);
  {
    const double eig_roe = vn_roe - c_roe;
    const double eps = std::max(0.0, std::max(eig_roe - (vn1 - c1), (vn2 - c2) - eig_roe));
    eig[0] = (std::abs(eig_roe) < eps) ? 0.5 * (eps + (eig_roe * eig_roe) / eps) : std::abs(eig_roe);
  }
  {
    const double eig_roe = vn_roe + c_roe;
    const double eps = std::max(0.0, std::max(eig_roe - (vn1 + c1), (vn2 + c2) - eig_roe));
    eig[2] = (std::abs(eig_roe) < eps) ? 0.5 * (eps + (eig_roe * eig_roe) / eps) : std::abs(eig_roe);
  }
  const double del_rho = rho2 - rho1;
  const double del_vn = vn2 - vn1;
  const double del_p = g.pressure(S2) - g.pressure(S1);
  const double ov_csq = 1.0 / (c_roe * c_roe);
  const double pov_csq = del_p * ov_csq;
  const double rhoc_del_vn = roe[0] * del_vn * c_roe * ov_csq;
  const double alpha_Vn = eig[1] * (del_rho - pov_csq);
  const double alpha_Vn_rho = eig[1] * roe[0];
  const double alpha_left = 0.5 * (eig[0] * (pov_csq - rhoc_del_vn));
  const double alpha_right = 0.5 * (eig[2] * (pov_csq + rhoc_del_vn));
  diss[0] = 0.5 * (alpha_Vn + alpha_left + alpha_right);
  diss[ener] = 0.0;
  for (int d = 0; d < dim; d++) {
    diss[1 + d] = 0.5 * (alpha_Vn * roe[1 + d] + alpha_Vn_rho * (del_mom[d] - nor[d] * inv_nor_mag * del_vn));
    diss[1 + d] += 0.5 * (alpha_left * (roe[1 + d] - nor[d] * inv_nor_mag * c_roe) +
                          alpha_right * (roe[1 + d] + nor[d] * inv_nor_mag * c_roe));
    diss[ener] += 0.5 * alpha_Vn_rho * (roe[1 + d] * del_mom[d]);
  }
  diss[ener] -= 0.5 * alpha_Vn_rho * vn_roe * del_vn;
  diss[ener] += 0.5 * alpha_Vn * ke + 0.5 * alpha_left * (roe[ener] - vn_roe * c_roe) +
                0.5 * alpha_right * (roe[ener] + vn_roe * c_roe);
  for (int e = 0; e < neq; e++) diss[e] *= nor_mag;
}

// DissipationOperator.hpp:174-232
inline void ScalarDissipation(const Gas &g, const double *S1, const double *S2, const double *nor, double *diss) {
  const int dim = g.dim, ener = dim + 1;
  const double rho1 = g.density(S1), rho2 = g.density(S2);
  const double rho_mean = 0.5 * (rho1 + rho2);
  const double drho = rho2 - rho1;
  const double p1 = g.pressure(S1), p2 = g.pressure(S2);
  double vn1 = 0.0, vn2 = 0.0, nor_mag = 0.0;
  double mom1[3] = {0, 0, 0}, mom2[3] = {0, 0, 0};
  diss[ener] = 0.0;
  for (int d = 0; d < dim; d++) {
    nor_mag += nor[d] * nor[d];
    mom1[d] = g.momentum(S1, d);
    mom2[d] = g.momentum(S2, d);
    const double v1 = g.velocity(S1, d), v2 = g.velocity(S2, d);
    const double dv = v2 - v1;
    const double vbar = 0.5 * (v1 + v2);
    vn1 += v1 * nor[d];
    vn2 += v2 * nor[d];
    diss[ener] += 0.5 * drho * v1 * v2 + rho_mean * dv * vbar;
  }
  nor_mag = std::sqrt(nor_mag);
  const double lambda_max =
      std::max(std::abs(vn1) + g.sound_speed(S1) * nor_mag, std::abs(vn2) + g.sound_speed(S2) * nor_mag);
  const double gm1_av_inv = 1.0 / (g.gamma - 1.0);
  const double beta1 = 0.5 * rho1 / p1;
  const double beta2 = 0.5 * rho2 / p2;
  const double beta_ln = ComputeLogMean(beta1, beta2, 1e-4);
  diss[ener] += 0.5 * drho * gm1_av_inv / beta_ln + 0.5 * rho_mean * gm1_av_inv * (1.0 / beta2 - 1.0 / beta1);
  diss[0] = 0.5 * lambda_max * drho;
  for (int d = 0; d < dim; d++) diss[1 + d] = 0.5 * lambda_max * (mom2[d] - mom1[d]);
  diss[ener] *= 0.5 * lambda_max;
}

// ---------------------------------------------------------------------------
// Face fluxes
// ---------------------------------------------------------------------------

// ChandrashekarFlux.hpp:91-155
inline void ChandrashekarFace(const Gas &g, const double *s1, const double *s2, const double *nor, double *flux) {
  const int dim = g.dim;
  const double rho1 = g.density(s1), rho2 = g.density(s2);
  const double rho_ln = ComputeLogMean(rho1, rho2, 1e-4);
  double mom[3] = {0, 0, 0};
  double hhat = 0.0, vn = 0.0;
  double diss[MAXEQ] = {0, 0, 0, 0, 0};
  for (int d = 0; d < dim; d++) {
    const double v1 = g.velocity(s1, d), v2 = g.velocity(s2, d);
    const double vbar = 0.5 * (v1 + v2);
    vn += vbar * nor[d];
    mom[d] = rho_ln * vbar;
    hhat += -0.25 * (v1 * v1 + v2 * v2) + vbar * vbar;
  }
  const double p1 = g.pressure(s1), p2 = g.pressure(s2);
  const double beta1 = 0.5 * rho1 / p1;
  const double beta2 = 0.5 * rho2 / p2;
  const double beta_ln = ComputeLogMean(beta1, beta2, 1e-4);
  const double p_hat = 0.5 * (rho1 + rho2) / (beta1 + beta2);
  const double gm1_av_inv = 2.0 / (g.gamma + g.gamma - 2.0);
  hhat += 0.5 / beta_ln * gm1_av_inv + p_hat / rho_ln;
  ScalarDissipation(g, s1, s2, nor, diss);
  flux[0] = rho_ln * vn - diss[0];
  for (int d = 0; d < dim; d++) flux[1 + d] = vn * mom[d] + p_hat * nor[d] - diss[1 + d];
  flux[dim + 1] = rho_ln * vn * hhat - diss[dim + 1];
}

// RanochaFlux.hpp:81-134
inline void RanochaFace(const Gas &g, const double *s1, const double *s2, const double *nor, double *flux) {
  const int dim = g.dim;
  const double rho1 = g.density(s1), rho2 = g.density(s2);
  const double rho_ln = ComputeLogMean(rho1, rho2, 1e-4);
  double vbar[3] = {0, 0, 0}, ppm[3] = {0, 0, 0}, diss[MAXEQ] = {0, 0, 0, 0, 0};
  double ke_flux = 0.0, vn = 0.0;
  const double p1 = g.pressure(s1), p2 = g.pressure(s2);
  const double pbar = 0.5 * (p1 + p2);
  for (int d = 0; d < dim; d++) {
    const double v1 = g.velocity(s1, d), v2 = g.velocity(s2, d);
    vbar[d] = 0.5 * (v1 + v2);
    vn += vbar[d] * nor[d];
    ke_flux += 0.5 * (v1 * v2);
    ppm[d] = 0.5 * (p1 * v2 + p2 * v1);
  }
  const double inv_beta_ln = (p1 * p2) / ComputeLogMean(rho1 * p2, rho2 * p1, 1e-4);
  const double inv_gm1 = 1.0 / (g.gamma - 1.0);
  const double ie_flux = inv_beta_ln * inv_gm1;
  RoeDissipation(g, s1, s2, nor, diss);
  flux[0] = rho_ln * vn - diss[0];
  flux[dim + 1] = 0.0;
  for (int d = 0; d < dim; d++) {
    flux[1 + d] = rho_ln * vn * vbar[d] + pbar * nor[d] - diss[1 + d];
    flux[dim + 1] += ppm[d] * nor[d];
  }
  flux[dim + 1] += rho_ln * (ie_flux + ke_flux) * vn - diss[dim + 1];
}

// PirozzoliFlux.hpp:73-120
template <class G>
inline void PirozzoliFace(const G &g, const double *s1, const double *s2, const double *nor, double *flux) {
  const int dim = g.dim;
  const double rho1 = g.density(s1), rho2 = g.density(s2);
  const double rho_mean = 0.5 * (rho1 + rho2);
  double vbar[3] = {0, 0, 0}, diss[MAXEQ] = {0, 0, 0, 0, 0};
  double ke_mean = 0.0, vn = 0.0;
  const double p1 = g.pressure(s1), p2 = g.pressure(s2);
  const double pbar = 0.5 * (p1 + p2);
  for (int d = 0; d < dim; d++) {
    const double v1 = g.velocity(s1, d), v2 = g.velocity(s2, d);
    vbar[d] = 0.5 * (v1 + v2);
    vn += vbar[d] * nor[d];
    ke_mean += 0.25 * ((v1 * v1) + (v2 * v2));
  }
  const double h_mean =
      0.5 * (g.specific_internal_energy(s1) + p1 / rho1 + g.specific_internal_energy(s2) + p2 / rho2);
  RoeDissipation(g, s1, s2, nor, diss);
  flux[0] = rho_mean * vn - diss[0];
  flux[dim + 1] = 0.0;
  for (int d = 0; d < dim; d++) flux[1 + d] = rho_mean * vn * vbar[d] + pbar * nor[d] - diss[1 + d];
  flux[dim + 1] += rho_mean * (h_mean + ke_mean) * vn - diss[dim + 1];
}

// RoeFlux.hpp:283-390 (no entropy fix; calorically perfect gas, no scalars)
inline void RoeFace(const Gas &g, const double *qL, const double *qR, const double *nor, double *flux) {
  const int dim = g.dim, neq = g.neq(), ener = dim + 1;
  const NormalFluxState left = MakeInviscidNormalFluxState(g, qL);
  const NormalFluxState right = MakeInviscidNormalFluxState(g, qR);
  double nor_mag2 = 0.0;
  for (int d = 0; d < dim; ++d) nor_mag2 += nor[d] * nor[d];
  const double nor_mag = std::sqrt(nor_mag2);
  const double inv_nor_mag = 1.0 / nor_mag;
  const double rhoL = g.density(qL), rhoR = g.density(qR);
  const double pL = left.pressure, pR = right.pressure;
  const double rootL = std::sqrt(rhoL), rootR = std::sqrt(rhoR);
  const double root_sum_inv = 1.0 / (rootL + rootR);
  const double rho_roe = rootL * rootR;
  const double HL = (g.energy(qL) + pL) / rhoL;
  const double HR = (g.energy(qR) + pR) / rhoR;
  const double H = (rootL * HL + rootR * HR) * root_sum_inv;
  double u[3] = {0, 0, 0}, du[3] = {0, 0, 0}, du_t[3] = {0, 0, 0}, nhat[3] = {0, 0, 0};
  double u2 = 0.0, un = 0.0, dun = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double uL = g.velocity(qL, d), uR = g.velocity(qR, d);
    nhat[d] = nor[d] * inv_nor_mag;
    u[d] = (rootL * uL + rootR * uR) * root_sum_inv;
    du[d] = uR - uL;
    u2 += u[d] * u[d];
    un += u[d] * nhat[d];
    dun += du[d] * nhat[d];
  }
  const double gamma = g.gamma;
  const double a2 = (gamma - 1.0) * (H - 0.5 * u2);
  const double a = std::sqrt(a2);
  const double dp = pR - pL;
  const double drho = rhoR - rhoL;
  const double alpha_minus = 0.5 * (dp - rho_roe * a * dun) / a2;
  const double alpha_plus = 0.5 * (dp + rho_roe * a * dun) / a2;
  const double alpha_zero = drho - dp / a2;
  const double lambda_minus = std::abs(un - a) * nor_mag;
  const double lambda_zero = std::abs(un) * nor_mag;
  const double lambda_plus = std::abs(un + a) * nor_mag;
  for (int d = 0; d < dim; ++d) du_t[d] = du[d] - dun * nhat[d];
  const double dmass = lambda_minus * alpha_minus + lambda_zero * alpha_zero + lambda_plus * alpha_plus;
  double diss[MAXEQ] = {0.0};
  diss[0] = dmass;
  for (int d = 0; d < dim; ++d)
    diss[1 + d] = lambda_minus * alpha_minus * (u[d] - a * nhat[d]) +
                  lambda_zero * (alpha_zero * u[d] + rho_roe * du_t[d]) +
                  lambda_plus * alpha_plus * (u[d] + a * nhat[d]);
  double u_dot_du_t = 0.0;
  for (int d = 0; d < dim; ++d) u_dot_du_t += u[d] * du_t[d];
  diss[ener] = lambda_minus * alpha_minus * (H - a * un) + lambda_zero * (0.5 * alpha_zero * u2 + rho_roe * u_dot_du_t) +
               lambda_plus * alpha_plus * (H + a * un);
  for (int e = 0; e < neq; ++e) {
    const double central = 0.5 * (ComputeInviscidNormalFluxEquation(g, left, nor, e) +
                                  ComputeInviscidNormalFluxEquation(g, right, nor, e));
    flux[e] = central - 0.5 * diss[e];
  }
}

// HLLFlux.hpp:53-116
template <class G>
inline void HLLFace(const G &g, const double *s1, const double *s2, const double *nor, double *flux) {
  const int dim = g.dim, neq = g.neq();
  const NormalFluxState left = MakeInviscidNormalFluxState(g, s1);
  const NormalFluxState right = MakeInviscidNormalFluxState(g, s2);
  double un1 = 0.0, un2 = 0.0, nor_mag2 = 0.0;
  for (int d = 0; d < dim; ++d) {
    un1 += g.velocity(s1, d) * nor[d];
    un2 += g.velocity(s2, d) * nor[d];
    nor_mag2 += nor[d] * nor[d];
  }
  const double nor_mag = std::sqrt(nor_mag2);
  const double c1 = g.sound_speed(s1) * nor_mag;
  const double c2 = g.sound_speed(s2) * nor_mag;
  const double sL1 = un1 - c1, sL2 = un2 - c2, sR1 = un1 + c1, sR2 = un2 + c2;
  const double sL = (sL1 < sL2) ? sL1 : sL2;
  const double sR = (sR1 > sR2) ? sR1 : sR2;
  for (int e = 0; e < neq; ++e) {
    const double fn1 = ComputeInviscidNormalFluxEquation(g, left, nor, e);
    const double fn2 = ComputeInviscidNormalFluxEquation(g, right, nor, e);
    if (sL >= 0.0)
      flux[e] = fn1;
    else if (sR <= 0.0)
      flux[e] = fn2;
    else
      flux[e] = (sR * fn1 - sL * fn2 + sL * sR * (s2[e] - s1[e])) / (sR - sL);
  }
}

// LaxFriedrichsFlux.hpp:188-211
template <class G>
inline void LLFFace(const G &g, const double *s1, const double *s2, const double *nor, double *flux) {
  const int neq = g.neq();
  const NormalFluxState left = MakeInviscidNormalFluxState(g, s1);
  const NormalFluxState right = MakeInviscidNormalFluxState(g, s2);
  const double lambda_max = MaximumNormalWaveSpeed(g, s1, s2, nor);
  for (int e = 0; e < neq; ++e) {
    const double nl = ComputeInviscidNormalFluxEquation(g, left, nor, e);
    const double nr = ComputeInviscidNormalFluxEquation(g, right, nor, e);
    flux[e] = 0.5 * (nl + nr) - 0.5 * lambda_max * (s2[e] - s1[e]);
  }
}

// Dispatch by the C-ABI ids (include/theseus_rhs.h).
inline void VolumeFlux(int id, const Gas &g, const double *q1, const double *q2, const double *m1, const double *m2,
                       double *F) {
  switch (id) {
    case 0: CentralVolume(g, q1, q2, m1, m2, F); return;
    case 1: ChandrashekarVolume(g, q1, q2, m1, m2, F); return;
    case 2: RanochaVolume(g, q1, q2, m1, m2, F); return;
    case 3: PirozzoliVolume(g, q1, q2, m1, m2, F); return;
  }
}
inline void FaceFlux(int id, const Gas &g, const double *q1, const double *q2, const double *n, double *F) {
  switch (id) {
    case 0: ChandrashekarFace(g, q1, q2, n, F); return;
    case 1: RanochaFace(g, q1, q2, n, F); return;
    case 2: PirozzoliFace(g, q1, q2, n, F); return;
    case 3: RoeFace(g, q1, q2, n, F); return;
    case 4: HLLFace(g, q1, q2, n, F); return;
    case 5: LLFFace(g, q1, q2, n, F); return;
  }
}

}  // namespace oracle
