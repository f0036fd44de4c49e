// TEST INFRASTRUCTURE ONLY.  CPU FP64 restatement of the reference's
// (chess-uiuc/Theseus) boundary conditions: the inviscid boundary fluxes, the
// Navier-Stokes boundary fluxes and the BR1 gradient's boundary jumps.  Only
// tests/, the smoke check and the bench's cpu_baseline leg may load this
// code, and only as the checker.
//
// Parity pins (tests/test_oracle_bc.py): slip-wall p* against the exact
// solution of the 1-D wall Riemann problem's two branches at their closed
// forms (no normal velocity: p* = p; weak shock / rarefaction expansions),
// rotation invariance of p*, the symmetry mirror state's invariants,
// consistency of the prescribed-state flux with the interior flux when the
// exterior state equals the interior one, the no-slip wall fluxes against
// the viscous normal flux assembled from the full tensor, and the
// isothermal-wall gradient jump vanishing for a state at the wall
// temperature and velocity.
#pragma once

#include <cmath>

#include "oracle_physics.hpp"

namespace oracle {

// BCType (bc_cache_utilities.hpp:11-23)
enum BCKind : int {
  BC_SLIP_WALL = 0,
  BC_SUPERSONIC_INFLOW = 1,
  BC_SUPERSONIC_OUTFLOW = 2,
  BC_PRESCRIBED_STATE = 3,
  BC_SYMMETRY = 4,
  BC_AXIS = 5,
  BC_NO_SLIP_ADIABATIC = 6,
  BC_NO_SLIP_ISOTHERMAL = 7,
};
constexpr int BC_STRIDE = 8;

// slipwall_pstar (Flow.hpp:55-75) of the state seen in the wall frame
// (RotatedWallStateView, Flow.hpp:77-154): component 0 along the unit normal,
// then the tangents the reference picks; pressure and sound speed from the
// rotated momenta.  slipwall_pstar_normal (Flow.hpp:156-170) returns 0 for a
// zero normal.
inline double SlipWallPStarNormal(const Gas &g, const double *q, const double *n) {
  const int dim = g.dim;
  double m2 = 0.0;
  for (int d = 0; d < dim; ++d) m2 += n[d] * n[d];
  if (m2 == 0.0) return 0.0;
  const double inv = 1.0 / std::sqrt(m2);
  const double nx = n[0] * inv, ny = n[1] * inv, nz = dim > 2 ? n[2] * inv : 0.0;
  double basis[3][3] = {{nx, ny, nz}, {0, 0, 0}, {0, 0, 0}};
  double tx, ty, tz;
  if (dim == 2) {
    tx = -ny;
    ty = nx;
    tz = 0.0;
  } else if (std::abs(nx) <= std::abs(ny) && std::abs(nx) <= std::abs(nz)) {
    tx = 0.0;
    ty = nz;
    tz = -ny;
  } else if (std::abs(ny) <= std::abs(nx) && std::abs(ny) <= std::abs(nz)) {
    tx = -nz;
    ty = 0.0;
    tz = nx;
  } else {
    tx = ny;
    ty = -nx;
    tz = 0.0;
  }
  basis[1][0] = tx;
  basis[1][1] = ty;
  basis[1][2] = tz;
  basis[2][0] = ny * tz - nz * ty;
  basis[2][1] = nz * tx - nx * tz;
  basis[2][2] = nx * ty - ny * tx;
  double r[MAXEQ];
  r[0] = q[0];
  r[dim + 1] = q[dim + 1];
  for (int c = 0; c < dim; ++c) {
    double v = 0.0;
    v += q[1] * basis[c][0];
    v += q[2] * basis[c][1];
    if (dim == 3) v += q[3] * basis[c][2];
    r[1 + c] = v;
  }
  const double rho = g.density(r);
  const double c = g.sound_speed(r);
  const double p = g.pressure(r);
  const double v = g.velocity(r, 0);
  const double gamma = g.gamma;
  const double gammaP1 = gamma + 1.;
  const double gammaM1 = gamma - 1.;
  const double gammaP1Inverse = 1.0 / gammaP1;
  const double gammaM1Inverse = 1.0 / gammaM1;
  if (v > 0.0)
    return p + 0.25 * v * gammaP1 * rho *
                   (v + std::sqrt(v * v + 8.0 * gammaP1Inverse * p * (gammaM1 * gammaP1Inverse + 1.0) / rho));
  return p * std::pow(std::max(1.0 + 0.5 * gammaM1 * v / c, 0.0001), 2.0 * gamma * gammaM1Inverse);
}

// SlipWallInviscidFluxKernel (bc_kernels.hpp:138-152)
inline void SlipWallFlux(const Gas &g, const double *q, const double *n, double *F) {
  const double p = SlipWallPStarNormal(g, q, n);
  for (int k = 0; k < g.neq(); ++k) F[k] = 0.0;
  for (int d = 0; d < g.dim; ++d) F[1 + d] = p * n[d];
}

// Symmetry exterior state (bc_kernels.hpp:248-272): momentum mirrored about
// the unit normal, m - 2 (m.n^) n^ (Kernels::Normalize, theseus_kernels.hpp:24-33)
inline void SymmetryState(const Gas &g, const double *q, const double *n, double *out) {
  const int dim = g.dim;
  for (int k = 0; k < g.neq(); ++k) out[k] = q[k];
  double un[MAXDIM], mom[MAXDIM];
  double fac = 0.0;
  for (int d = 0; d < dim; ++d) {
    un[d] = n[d];
    mom[d] = q[1 + d];
    fac += un[d] * un[d];
  }
  fac = 1.0 / std::sqrt(fac);
  for (int d = 0; d < dim; ++d) un[d] *= fac;
  double nv = 0.0;
  for (int d = 0; d < dim; ++d) nv += mom[d] * un[d];
  for (int d = 0; d < dim; ++d) out[1 + d] = -2.0 * nv * un[d] + mom[d];
}

// ApplyBoundaryConditionInviscid (bc_kernels.hpp:219-287); `face` is the
// operator's surface flux F*(q_L, q_R, n).
template <typename FaceFn>
inline void BoundaryInviscidFlux(const Gas &g, FaceFn &&face, int type, const double *data, const double *q,
                                 const double *n, double *F) {
  switch (type) {
    case BC_SLIP_WALL:
      SlipWallFlux(g, q, n, F);
      return;
    case BC_SUPERSONIC_OUTFLOW:
      face(q, q, n, F);
      return;
    case BC_PRESCRIBED_STATE:
    case BC_SUPERSONIC_INFLOW:
      face(q, data, n, F);
      return;
    case BC_SYMMETRY: {
      double qb[MAXEQ];
      SymmetryState(g, q, n, qb);
      face(q, qb, n, F);
      return;
    }
    default:
      for (int k = 0; k < g.neq(); ++k) F[k] = 0.0;
      return;
  }
}

// ComputeViscousNormalFluxKernel (NavierStokesFlux.hpp:294-357), Cartesian:
// traction n.tau per momentum component, then the energy flux
// u.traction + n.(kappa grad T).
inline void ViscousNormalFlux(const Gas &g, const double *q, const double *const grads[3], const double *n,
                              double *F) {
  const int dim = g.dim;
  const double mu = g.viscosity(q), kappa = g.thermal_conductivity(q), mu_bulk = g.bulk_viscosity(q);
  double grad_rho[MAXDIM] = {0, 0, 0}, grad_thermal[MAXDIM] = {0, 0, 0}, grad_t[MAXDIM] = {0, 0, 0};
  double divergence = 0.0;
  for (int d = 0; d < dim; ++d) {
    grad_rho[d] = grads[d][0];
    grad_thermal[d] = grads[d][dim + 1];
    divergence += grads[d][1 + d];
  }
  g.grad_temperature(q, grad_rho, grad_thermal, grad_t);
  for (int k = 0; k < g.neq(); ++k) F[k] = 0.0;
  double energy_flux = 0.0;
  for (int comp = 0; comp < dim; ++comp) {
    double traction = 0.0;
    for (int dir = 0; dir < dim; ++dir) {
      const double stress = comp == dir ? mu * (2.0 * grads[dir][1 + comp] - mu_bulk * divergence)
                                        : mu * (grads[dir][1 + comp] + grads[comp][1 + dir]);
      traction += n[dir] * stress;
    }
    F[1 + comp] = traction;
    energy_flux += g.velocity(q, comp) * traction;
  }
  for (int dir = 0; dir < dim; ++dir) energy_flux += n[dir] * kappa * grad_t[dir];
  F[dim + 1] = energy_flux;
}

// ApplyViscousBoundaryCondition (bc_kernels.hpp:289-375) with the wall
// kernels NoSlipAdiabWallFluxKernel (:154-184) and NoSlipIsothWallFluxKernel
// (:186-216); grads[d] are the primitive gradients at the boundary point.
template <typename FaceFn>
inline void BoundaryViscousFlux(const Gas &g, FaceFn &&face, int type, const double *data, const double *q,
                                const double *const grads[3], const double *n, double *F) {
  const int dim = g.dim, neq = g.neq();
  switch (type) {
    case BC_SLIP_WALL:
    case BC_SUPERSONIC_INFLOW:
    case BC_SUPERSONIC_OUTFLOW:
    case BC_SYMMETRY:
      BoundaryInviscidFlux(g, face, type, data, q, n, F);
      return;
    case BC_PRESCRIBED_STATE: {
      BoundaryInviscidFlux(g, face, type, data, q, n, F);
      double vf[MAXEQ][MAXDIM];
      ComputeViscousFluxKernel(g, q, grads[0], grads[1], grads[2], vf);
      for (int k = 0; k < neq; ++k)
        for (int d = 0; d < dim; ++d) F[k] -= vf[k][d] * n[d];
      return;
    }
    case BC_NO_SLIP_ADIABATIC: {
      const double pressure = SlipWallPStarNormal(g, q, n);
      ViscousNormalFlux(g, q, grads, n, F);
      double m2 = 0.0;
      for (int d = 0; d < dim; ++d) m2 += n[d] * n[d];
      double energy_flux = data[dim] * std::sqrt(m2);
      for (int d = 0; d < dim; ++d) energy_flux += data[d] * F[1 + d];
      F[dim + 1] = energy_flux;
      for (int k = 0; k < neq; ++k) F[k] = -F[k];
      for (int d = 0; d < dim; ++d) F[1 + d] += pressure * n[d];
      return;
    }
    case BC_NO_SLIP_ISOTHERMAL: {
      const double pressure = SlipWallPStarNormal(g, q, n);
      ViscousNormalFlux(g, q, grads, n, F);
      for (int d = 0; d < dim; ++d) F[dim + 1] -= g.velocity(q, d) * F[1 + d];
      for (int d = 0; d < dim; ++d) F[dim + 1] += data[d] * F[1 + d];
      for (int k = 0; k < neq; ++k) F[k] = -F[k];
      for (int d = 0; d < dim; ++d) F[1 + d] += pressure * n[d];
      return;
    }
    default:
      for (int k = 0; k < neq; ++k) F[k] = 0.0;
      return;
  }
}

// ComputeBdrFaceGradFlux (bc_kernels.hpp:32-135): the jump S_boundary - S
// in entropy variables S (the gradient lift adds w * jump * n).  The
// isothermal wall's beta is 1 / (R T_wall) (EOS.hpp:42-46).
inline void BoundaryGradJump(const Gas &g, int type, const double *data, const double *S, double *jump) {
  const int dim = g.dim, neq = g.neq();
  for (int k = 0; k < neq; ++k) jump[k] = 0.0;
  switch (type) {
    case BC_PRESCRIBED_STATE: {
      double E[MAXEQ];
      g.entropy_state(data, E);
      for (int k = 0; k < neq; ++k) jump[k] = E[k] - S[k];
      return;
    }
    case BC_NO_SLIP_ADIABATIC: {
      const double v = -S[dim + 1];
      jump[0] = S[0];
      for (int d = 0; d < dim; ++d) jump[1 + d] = data[d] * v;
      jump[dim + 1] = -v;
      for (int k = 0; k < neq; ++k) jump[k] -= S[k];
      return;
    }
    case BC_NO_SLIP_ISOTHERMAL: {
      const double beta_like = 1.0 / (g.R_gas * data[dim]);
      jump[0] = S[0];
      for (int d = 0; d < dim; ++d) jump[1 + d] = data[d] * beta_like;
      jump[dim + 1] = -beta_like;
      for (int k = 0; k < neq; ++k) jump[k] -= S[k];
      return;
    }
    default:
      return;
  }
}

}  // namespace oracle
