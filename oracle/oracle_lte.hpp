// TEST INFRASTRUCTURE ONLY.  CPU FP64 restatement of the reference's
// tabulated LTE gas (chess-uiuc/Theseus): the six-plane Helmholtz table and
// its cell search, the bicubic Hermite interpolant of H(rho, T) and its
// exact divided differences, the thermodynamics derived from H, temperature
// recovery from (rho, e) by an inverse-table guess and Newton, the
// Itoh-Abe discrete gradients behind the entropy-conserving means, and the
// LTEEC two-point flux with Lax-Friedrichs dissipation.  Also the synthetic
// dissociating-gas table of the LTE Taylor-Green workload.  The product
// never links or calls this code.
//
// Restated files: include/LTEHelmholtzTable.hpp, LTEHelmholtzInterpolation.hpp,
// LTEHelmholtzEOS.hpp, LTEHelmholtzGasEOS.hpp, LTEECFlux.hpp,
// DissipationOperator.hpp:15-28, LTEHelmholtzTableBuilder.hpp:131-218 (the
// inverse table), LTETable.hpp:85-103 (grids).
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

namespace oracle {
namespace lte {

// LTEHelmholtzTable.hpp:13-31: planes H, H_rho, H_T, H_rhoT, mu, lambda;
// index property*nx*ny + T_idx*nx + rho_idx
struct Table {
  int nx = 0, ny = 0;
  std::vector<double> rho, T, e;   // grids (e: the inverse table's energy axis)
  std::vector<double> planes;      // 6 * nx * ny
  std::vector<double> inv;         // nx * ny: temperature guess at (rho_i, e_j)
  // viscosity by Sutherland's law instead of the mu plane (the reference's
  // SUTHERLAND build, LTETransport.hpp:22-43)
  bool sutherland = false;
  double mu0 = 0, T0 = 0, Ts = 0;
  int index(int prop, int i, int j) const { return prop * nx * ny + j * nx + i; }
};

// LTEHelmholtzTable.hpp:84-105 (bisection; -1 outside the grid)
inline int find_cell(const double *grid, int count, double value) {
  if (count < 2 || !std::isfinite(value) || value < grid[0] || value > grid[count - 1]) return -1;
  int lo = 0, hi = count - 1;
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (value < grid[mid]) hi = mid;
    else lo = mid;
  }
  return lo;
}

// LTEHelmholtzInterpolation.hpp:26-131: cubic Hermite bases on [0,1] and
// their exact divided differences
template <int D>
inline double position_basis(int side, double x) {
  const double x2 = x * x;
  if constexpr (D == 0) return side == 0 ? 2 * x2 * x - 3 * x2 + 1 : -2 * x2 * x + 3 * x2;
  else if constexpr (D == 1) return side == 0 ? 6 * x2 - 6 * x : -6 * x2 + 6 * x;
  else return side == 0 ? 12 * x - 6 : -12 * x + 6;
}
template <int D>
inline double slope_basis(int side, double x) {
  const double x2 = x * x;
  if constexpr (D == 0) return side == 0 ? x2 * x - 2 * x2 + x : x2 * x - x2;
  else if constexpr (D == 1) return side == 0 ? 3 * x2 - 4 * x + 1 : 3 * x2 - 2 * x;
  else return side == 0 ? 6 * x - 4 : 6 * x - 2;
}
template <int D>
inline double divided_position_basis(int side, double x, double y) {
  double v;
  if constexpr (D == 0) v = 2 * (x * x + x * y + y * y) - 3 * (x + y);
  else if constexpr (D == 1) v = 6 * (x + y) - 6;
  else v = 12;
  return side == 0 ? v : -v;
}
template <int D>
inline double divided_slope_basis(int side, double x, double y) {
  if constexpr (D == 0) return x * x + x * y + y * y - (side == 0 ? 2 : 1) * (x + y) + (side == 0 ? 1 : 0);
  else if constexpr (D == 1) return 3 * (x + y) - (side == 0 ? 4 : 2);
  else return 6;
}
template <int D>
inline double coordinate_scale(double w) {
  return D == 0 ? 1.0 : (D == 1 ? w : w * w);
}

// The four planes at the lower corner of one cell.
struct Cell {
  const double *H = nullptr, *Hr = nullptr, *HT = nullptr, *HrT = nullptr;
  int stride = 0;
  double wr = 0, wt = 0;  // cell widths in rho and T
  int i = -1, j = -1;
};

// LTEHelmholtzInterpolation.hpp:136-172
template <int DR, int DT>
inline double evaluate(const Cell &c, double xi, double eta) {
  double r = 0;
  for (int ds = 0; ds < 2; ++ds) {
    const double dv = position_basis<DR>(ds, xi), dsl = slope_basis<DR>(ds, xi);
    for (int ts = 0; ts < 2; ++ts) {
      const int o = ds + ts * c.stride;
      const double tv = position_basis<DT>(ts, eta), tsl = slope_basis<DT>(ts, eta);
      r += c.H[o] * dv * tv;
      r += (c.Hr[o] * c.wr) * dsl * tv;
      r += (c.HT[o] * c.wt) * dv * tsl;
      r += (c.HrT[o] * c.wr * c.wt) * dsl * tsl;
    }
  }
  return r / (coordinate_scale<DR>(c.wr) * coordinate_scale<DT>(c.wt));
}

// LTEHelmholtzInterpolation.hpp:177-211
template <int AXIS, int DR = 0, int DT = 0>
inline double divided_difference(const Cell &c, double xi1, double eta1, double xi2, double eta2) {
  double r = 0;
  for (int i = 0; i < 2; ++i) {
    const double rv = AXIS == 0 ? divided_position_basis<DR>(i, xi1, xi2) : position_basis<DR>(i, xi1);
    const double rs = AXIS == 0 ? divided_slope_basis<DR>(i, xi1, xi2) : slope_basis<DR>(i, xi1);
    for (int j = 0; j < 2; ++j) {
      const double tv = AXIS == 1 ? divided_position_basis<DT>(j, eta1, eta2) : position_basis<DT>(j, eta1);
      const double ts = AXIS == 1 ? divided_slope_basis<DT>(j, eta1, eta2) : slope_basis<DT>(j, eta1);
      const int o = i + j * c.stride;
      r += c.H[o] * rv * tv;
      r += (c.Hr[o] * c.wr) * rs * tv;
      r += (c.HT[o] * c.wt) * rv * ts;
      r += (c.HrT[o] * c.wr * c.wt) * rs * ts;
    }
  }
  const double scale = coordinate_scale<DR>(c.wr) * coordinate_scale<DT>(c.wt) * (AXIS == 0 ? c.wr : c.wt);
  return r / scale;
}

// LTEHelmholtzEOS.hpp:14-102
struct Eval {
  Cell cell;
  double xi = 0, eta = 0, rho = 0, T = 0;
  bool valid = false;
  template <int DR = 0, int DT = 0>
  double H() const { return evaluate<DR, DT>(cell, xi, eta); }
  double pressure() const { return rho * rho * H<1, 0>(); }
  double energy() const { return H() - T * H<0, 1>(); }
  double entropy() const { return -H<0, 1>(); }
  double gibbs() const { return H() + rho * H<1, 0>(); }
  double pressure_density_derivative() const { return 2 * rho * H<1, 0>() + rho * rho * H<2, 0>(); }
  double pressure_temperature_derivative() const { return rho * rho * H<1, 1>(); }
  double gibbs_density_derivative() const { return 2 * H<1, 0>() + rho * H<2, 0>(); }
  double gibbs_temperature_derivative() const { return H<0, 1>() + rho * H<1, 1>(); }
  double cv() const { return -T * H<0, 2>(); }
  // LTEHelmholtzEOS.hpp:76-80
  double cp() const {
    const double m = H<1, 1>();
    return cv() + T * rho * rho * m * m / pressure_density_derivative();
  }
  double sound_speed_squared() const {
    const double m = H<1, 1>();
    return pressure_density_derivative() - rho * rho * m * m / H<0, 2>();
  }
};

// LTEHelmholtzEOS.hpp:106-130
inline Eval at(const Table &t, double rho, double T) {
  Eval r;
  const int i = find_cell(t.rho.data(), t.nx, rho);
  const int j = find_cell(t.T.data(), t.ny, T);
  if (i < 0 || j < 0) return r;
  const int off = j * t.nx + i, size = t.nx * t.ny;
  const double dr = t.rho[i + 1] - t.rho[i], dt = t.T[j + 1] - t.T[j];
  r.cell.H = t.planes.data() + off;
  r.cell.Hr = t.planes.data() + size + off;
  r.cell.HT = t.planes.data() + 2 * size + off;
  r.cell.HrT = t.planes.data() + 3 * size + off;
  r.cell.stride = t.nx;
  r.cell.wr = dr;
  r.cell.wt = dt;
  r.cell.i = i;
  r.cell.j = j;
  r.xi = (rho - t.rho[i]) / dr;
  r.eta = (T - t.T[j]) / dt;
  r.rho = rho;
  r.T = T;
  r.valid = true;
  return r;
}

// LTEHelmholtzEOS.hpp:151-223: inverse-table guess (difference form), Newton on e(T)
// status: 0 converged, 1 outside table, 2 invalid thermodynamics, 3 iteration limit
struct Recovery {
  double T = 0;
  int iterations = 0;
  int status = 1;
};
inline Recovery temperature_from_energy(const Table &t, double rho, double e) {
  Recovery r;
  const int i = find_cell(t.rho.data(), t.nx, rho);
  const int j = find_cell(t.e.data(), t.ny, e);
  if (i < 0 || j < 0) return r;
  const double x = (rho - t.rho[i]) / (t.rho[i + 1] - t.rho[i]);
  const double y = (e - t.e[j]) / (t.e[j + 1] - t.e[j]);
  const double *p = t.inv.data() + j * t.nx + i;
  const double lower = p[0] + x * (p[1] - p[0]);
  const double upper = p[t.nx] + x * (p[t.nx + 1] - p[t.nx]);
  r.T = lower + y * (upper - lower);
  for (int it = 0; it < 100; ++it) {
    const Eval s = at(t, rho, r.T);
    if (!s.valid) return r;
    const double cv = s.cv(), guess = s.energy();
    if (!std::isfinite(cv) || cv <= 0 || !std::isfinite(guess)) {
      r.status = 2;
      return r;
    }
    const double step = (e - guess) / cv;
    const double next = r.T + step;
    ++r.iterations;
    const bool inside = std::isfinite(next) && next >= t.T[0] && next <= t.T[t.ny - 1];
    if (std::fabs(step) <= 1e-9 + 1e-12 * r.T) {
      if (next < t.T[0]) r.T = t.T[0];
      else if (next > t.T[t.ny - 1]) r.T = t.T[t.ny - 1];
      else r.T = next;
      r.status = 0;
      return r;
    }
    r.T = next;
    if (!inside) return r;
  }
  r.status = 3;
  return r;
}

// LTEHelmholtzEOS.hpp:225-397: Itoh-Abe discrete gradient of f/T in (rho, beta)
enum class Q { pressure, gibbs };
template <Q P>
inline double value(const Eval &s) {
  return P == Q::pressure ? s.pressure() : s.gibbs();
}
template <Q P>
inline double density_secant(const Eval &a, const Eval &b, double eta) {
  const Cell &c = a.cell;
  const double r1 = a.rho, r2 = b.rho;
  const double mean_hr = 0.5 * (evaluate<1, 0>(c, a.xi, eta) + evaluate<1, 0>(c, b.xi, eta));
  const double div_hr = divided_difference<0, 1, 0>(c, a.xi, eta, b.xi, eta);
  if constexpr (P == Q::pressure) {
    return 0.5 * (r1 * r1 + r2 * r2) * div_hr + (r1 + r2) * mean_hr;
  } else {
    const double div_h = divided_difference<0>(c, a.xi, eta, b.xi, eta);
    return div_h + 0.5 * (r1 + r2) * div_hr + mean_hr;
  }
}
template <Q P>
inline double beta_secant(const Eval &a, const Eval &b, double xi, double rho) {
  const Cell &c = a.cell;
  const double mean_t = 0.5 * (a.T + b.T);
  const double mean_hr = 0.5 * (evaluate<1, 0>(c, xi, a.eta) + evaluate<1, 0>(c, xi, b.eta));
  const double div_hr = divided_difference<1, 1, 0>(c, xi, a.eta, xi, b.eta);
  const double hr_beta = mean_hr - mean_t * div_hr;
  if constexpr (P == Q::pressure) {
    return rho * rho * hr_beta;
  } else {
    const double mean_h = 0.5 * (evaluate<0, 0>(c, xi, a.eta) + evaluate<0, 0>(c, xi, b.eta));
    const double div_h = divided_difference<1>(c, xi, a.eta, xi, b.eta);
    return mean_h - mean_t * div_h + rho * hr_beta;
  }
}
struct Gradient {
  double density = 0, beta = 0;
};
template <Q P>
inline Gradient discrete_gradient(const Table &t, const Eval &left, const Eval &right) {
  const bool reverse = left.rho > right.rho || (left.rho == right.rho && left.T > right.T);
  const Eval &a = reverse ? right : left;
  const Eval &b = reverse ? left : right;
  const double b1 = 1 / a.T, b2 = 1 / b.T;
  const double dr = b.rho - a.rho, db = b2 - b1;
  const double rmax = a.rho > b.rho ? a.rho : b.rho;
  const double bmax = b1 > b2 ? b1 : b2;
  const bool dd = std::fabs(dr) >= 2.22e-15 + 1.49e-8 * rmax;
  const bool bd = std::fabs(db) >= 2.22e-15 + 1.49e-8 * bmax;
  const double f11 = value<P>(a), f22 = value<P>(b);
  Gradient g;
  const bool same_cell = a.cell.H == b.cell.H;
  if (same_cell) {
    if (dd) g.density = 0.5 * (density_secant<P>(a, b, a.eta) * b1 + density_secant<P>(a, b, b.eta) * b2);
    if (bd) g.beta = 0.5 * (beta_secant<P>(a, b, a.xi, a.rho) + beta_secant<P>(a, b, b.xi, b.rho));
  } else if (dd || bd) {
    const double f12 = value<P>(at(t, a.rho, b.T));
    const double f21 = value<P>(at(t, b.rho, a.T));
    if (dd) g.density = 0.5 * ((f21 * b1 - f11 * b1) + (f22 * b2 - f12 * b2)) / dr;
    if (bd) g.beta = 0.5 * ((f12 * b2 - f11 * b1) + (f22 * b2 - f21 * b1)) / db;
  }
  if (!dd) {
    if constexpr (P == Q::pressure)
      g.density = 0.5 * (a.pressure_density_derivative() * b1 + b.pressure_density_derivative() * b2);
    else
      g.density = 0.5 * (a.gibbs_density_derivative() * b1 + b.gibbs_density_derivative() * b2);
  }
  if (!bd) {
    double d1, d2;
    if constexpr (P == Q::pressure) {
      d1 = a.pressure_temperature_derivative();
      d2 = b.pressure_temperature_derivative();
    } else {
      d1 = a.gibbs_temperature_derivative();
      d2 = b.gibbs_temperature_derivative();
    }
    g.beta = 0.5 * ((f11 - a.T * d1) + (f22 - b.T * d2));
  }
  return g;
}
struct Means {
  double density = 0, energy = 0;
};
inline Means thermodynamic_means(const Table &t, const Eval &l, const Eval &r) {
  const Gradient p = discrete_gradient<Q::pressure>(t, l, r);
  const Gradient g = discrete_gradient<Q::gibbs>(t, l, r);
  const double rho = p.density / g.density;
  return {rho, g.beta - p.beta / rho};
}

// Per-point state of the gas: temperature recovered once (LTEHelmholtzGasEOS.hpp:60-77)
inline Eval evaluate_state(const Table &t, const double *q, int dim) {
  double m2 = 0;
  for (int d = 0; d < dim; ++d) m2 += q[1 + d] * q[1 + d];
  const double rho = q[0];
  const double e = (q[dim + 1] - 0.5 * m2 / rho) / rho;
  const Recovery rec = temperature_from_energy(t, rho, e);
  if (rec.status != 0) throw std::runtime_error("Helmholtz temperature recovery failed");
  const Eval s = at(t, rho, rec.T);
  if (!s.valid) throw std::runtime_error("Helmholtz state outside the table");
  return s;
}

// LTEHelmholtzEOS.hpp:99-110: viscosity (plane 4) or thermal conductivity
// (plane 5), bilinear in the evaluated cell's coordinates
inline double transport(const Table &t, const Eval &s, int prop) {
  const double *p = t.planes.data() + t.index(prop, s.cell.i, s.cell.j);
  return p[0] * (1 - s.xi) * (1 - s.eta) + p[1] * s.xi * (1 - s.eta) + p[t.nx] * (1 - s.xi) * s.eta +
         p[t.nx + 1] * s.xi * s.eta;
}

// LTETransport.hpp:22-43: the mu plane, or with the SUTHERLAND build
// viscosity_from_temperature at the recovered temperature
inline double viscosity(const Table &t, const Eval &s) {
  if (!t.sutherland) return transport(t, s, 4);
  const double relative_temperature = s.T / t.T0;
  return t.mu0 * (t.T0 + t.Ts) * relative_temperature * std::sqrt(relative_temperature) / (s.T + t.Ts);
}

// The tabulated gas on the viscous path, LTEGasModel<LTEHelmholtzGasEOS,
// LTETransport> (LTEGasModel.hpp:339): viscosity and conductivity from the
// table (LTETransport.hpp:32-65, the build without SUTHERLAND), the constant
// bulk-viscosity ratio, the entropy variables ((g - |u|^2/2)/T, u/T, -1/T)
// and their gradient conversion (LTEHelmholtzGasEOS.hpp:133-171).  The same
// members as the ideal gas, so the Navier-Stokes assembly takes either.
struct TableGas {
  const Table *t = nullptr;
  int dim = 3;
  double mu_bulk = 2.0 / 3.0;

  Eval state(const double *q) const { return evaluate_state(*t, q, dim); }
  // The accessors the Pirozzoli, HLL and LLF fluxes take from the gas
  // (LTEGasModel.hpp:160-267, LTEHelmholtzGasEOS.hpp:40-110): the state's
  // own density, momentum and energy; pressure, sound speed and
  // gamma = cp/cv of the interpolant at the recovered temperature.
  int neq() const { return dim + 2; }
  double density(const double *q) const { return q[0]; }
  double momentum(const double *q, int d) const { return q[1 + d]; }
  double energy(const double *q) const { return q[dim + 1]; }
  double kinetic_energy_density(const double *q) const {
    double m2 = 0;
    for (int d = 0; d < dim; ++d) m2 += q[1 + d] * q[1 + d];
    return 0.5 * m2 / q[0];
  }
  double specific_internal_energy(const double *q) const { return (energy(q) - kinetic_energy_density(q)) / q[0]; }
  double pressure(const double *q) const { return state(q).pressure(); }
  double sound_speed(const double *q) const { return std::sqrt(state(q).sound_speed_squared()); }
  double gamma_at(const double *q) const {
    const Eval s = state(q);
    return s.cp() / s.cv();
  }
  double velocity(const double *q, int d) const { return q[1 + d] / q[0]; }
  double viscosity(const double *q) const { return lte::viscosity(*t, state(q)); }
  double thermal_conductivity(const double *q) const { return transport(*t, state(q), 5); }
  double bulk_viscosity(const double *) const { return mu_bulk; }
  // LTEHelmholtzGasEOS.hpp:133-142: the energy slot of the primitive gradient is grad T
  void grad_temperature(const double *, const double *, const double *grad_p, double *grad_t) const {
    for (int d = 0; d < dim; ++d) grad_t[d] = grad_p[d];
  }
  // LTEHelmholtzGasEOS.hpp:144-157 (kinetic energy density / rho, :40-50)
  void entropy_state(const double *q, double *E) const {
    const Eval s = state(q);
    const double beta = 1 / s.T;
    double m2 = 0;
    for (int d = 0; d < dim; ++d) m2 += q[1 + d] * q[1 + d];
    const double kinetic = 0.5 * m2 / q[0] / q[0];
    E[0] = beta * (s.gibbs() - kinetic);
    for (int d = 0; d < dim; ++d) E[1 + d] = beta * velocity(q, d);
    E[dim + 1] = -beta;
  }
  // LTEHelmholtzGasEOS.hpp:159-171
  void grad_entropy_to_grad_prim(const double *q, const double *dE, double *dP) const {
    const double T = state(q).T;
    dP[0] = 0.0;
    for (int d = 0; d < dim; ++d) dP[1 + d] = T * dE[1 + d] + T * velocity(q, d) * dE[dim + 1];
    dP[dim + 1] = T * T * dE[dim + 1];
  }
};

// LTEECFlux.hpp:11-85 (volume: mean metric row; face: the normal, with
// Lax-Friedrichs dissipation, DissipationOperator.hpp:15-28)
inline void lteec_flux(const Table &t, bool face, int dim, const double *q1, const double *q2, const double *dir1,
                       const double *dir2, double *F) {
  const Eval s1 = evaluate_state(t, q1, dim);
  const Eval s2 = evaluate_state(t, q2, dim);
  const Means m = thermodynamic_means(t, s1, s2);
  const double p1 = s1.pressure(), p2 = s2.pressure();
  double c1 = 0, c2 = 0;
  if (face) {
    c1 = std::sqrt(s1.sound_speed_squared());
    c2 = std::sqrt(s2.sound_speed_squared());
  }
  const double pm = 0.5 * (p1 + p2);
  double vn = 0, ck = 0, pw = 0, v1n = 0, v2n = 0, n2 = 0;
  for (int d = 0; d < dim; ++d) {
    const double dir = face ? dir1[d] : 0.5 * (dir1[d] + dir2[d]);
    const double u1 = q1[1 + d] / q1[0], u2 = q2[1 + d] / q2[0];
    vn += 0.5 * (u1 + u2) * dir;
    ck += 0.5 * u1 * u2;
    pw += 0.5 * (p1 * u2 + p2 * u1) * dir;
    if (face) {
      v1n += u1 * dir;
      v2n += u2 * dir;
      n2 += dir * dir;
    }
  }
  double diss = 0;
  if (face) {
    const double nm = std::sqrt(n2);
    const double sp1 = std::fabs(v1n) + c1 * nm, sp2 = std::fabs(v2n) + c2 * nm;
    diss = 0.5 * (sp1 > sp2 ? sp1 : sp2);
  }
  const double mass = m.density * vn;
  F[0] = mass - diss * (q2[0] - q1[0]);
  for (int d = 0; d < dim; ++d) {
    const double dir = face ? dir1[d] : 0.5 * (dir1[d] + dir2[d]);
    const double vel = 0.5 * (q1[1 + d] / q1[0] + q2[1 + d] / q2[0]);
    F[1 + d] = mass * vel + pm * dir - diss * (q2[1 + d] - q1[1 + d]);
  }
  F[dim + 1] = mass * (m.energy + ck) + pw - diss * (q2[dim + 1] - q1[dim + 1]);
}

// ---------------------------------------------------------------------------
// Synthetic dissociating-gas table (the LTE Taylor-Green workload):
// gamma_eff(T) = 1.4 - 0.2 tanh((T - Tc)/w), cv = R/(gamma_eff - 1),
// e(T) = int_{T0}^T cv, s(rho, T) = int_{T0}^T cv/T - R ln rho,
// H = e - T s, H_rho = R T/rho, H_T = -s, H_rhoT = R/rho.
// ---------------------------------------------------------------------------
struct SyntheticGas {
  double R = 287.05, Tc = 4000.0, w = 1000.0, g0 = 1.4, g1 = 0.2;
  double cv(double T) const { return R / (g0 - 1 - g1 * std::tanh((T - Tc) / w)); }
  // int cv dT in closed form: cv = R/(a - b tanh x), x = (T - Tc)/w,
  // int dx/(a - b tanh x) = (a x + b ln(a cosh x - b sinh x))/(a^2 - b^2)
  double energy(double T, double T0) const {
    const double a = g0 - 1, b = g1;
    auto F = [&](double TT) {
      const double x = (TT - Tc) / w;
      // a cosh x - b sinh x = exp(|x|)/2 ((a - b sgn) + (a + b sgn) exp(-2|x|)), kept in logs
      const double ax = std::fabs(x), sg = x >= 0 ? 1.0 : -1.0;
      const double lg = ax + std::log(0.5 * ((a - b * sg) + (a + b * sg) * std::exp(-2 * ax)));
      return R * w * (a * x + b * lg) / (a * a - b * b);
    };
    return F(T) - F(T0);
  }
};

// LTETable.hpp:94-102 (log10-spaced)
inline std::vector<double> log_grid(int n, double lo, double hi) {
  std::vector<double> g(n);
  const double step = std::log10(hi / lo) / (n - 1);
  for (int i = 0; i < n; ++i) g[i] = lo * std::pow(10.0, i * step);
  return g;
}

// Entropy integral int_{T0}^{T_j} cv/T dT at every grid temperature:
// 16-point Gauss-Legendre on each grid interval, cumulative.
inline std::vector<double> entropy_integral(const SyntheticGas &gas, const std::vector<double> &T) {
  static const double x16[8] = {0.0950125098376374, 0.2816035507792589, 0.4580167776572274, 0.6178762444026438,
                                0.7554044083550030, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499};
  static const double w16[8] = {0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                                0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541};
  std::vector<double> s(T.size(), 0.0);
  for (size_t j = 1; j < T.size(); ++j) {
    const double a = T[j - 1], b = T[j], c = 0.5 * (a + b), h = 0.5 * (b - a);
    double acc = 0;
    for (int k = 0; k < 8; ++k) {
      const double t1 = c - h * x16[k], t2 = c + h * x16[k];
      acc += w16[k] * (gas.cv(t1) / t1 + gas.cv(t2) / t2);
    }
    s[j] = s[j - 1] + h * acc;
  }
  return s;
}

// Forward planes from the analytic model and the inverse table as the
// reference builds it (LTEHelmholtzTableBuilder.hpp:148-216): energy axis
// uniform over the padded range of H - T H_T, bisection (60 halvings) on the
// interpolant's energy per density column, endpoint guesses outside it.
inline Table synthetic_table(int nx, int ny, double rho_lo, double rho_hi, double T_lo, double T_hi,
                             const SyntheticGas &gas = SyntheticGas()) {
  Table t;
  t.nx = nx;
  t.ny = ny;
  t.rho = log_grid(nx, rho_lo, rho_hi);
  t.T = log_grid(ny, T_lo, T_hi);
  t.planes.assign(6 * (size_t)nx * ny, 0.0);
  t.inv.assign((size_t)nx * ny, 0.0);
  t.e.assign(ny, 0.0);
  const std::vector<double> sT = entropy_integral(gas, t.T);
  double emin = 1e300, emax = -1e300, scale = 1;
  for (int j = 0; j < ny; ++j) {
    const double T = t.T[j];
    const double e = gas.energy(T, t.T[0]);
    for (int i = 0; i < nx; ++i) {
      const double r = t.rho[i];
      const double s = sT[j] - gas.R * std::log(r);
      const double H = e - T * s;
      t.planes[t.index(0, i, j)] = H;
      t.planes[t.index(1, i, j)] = gas.R * T / r;
      t.planes[t.index(2, i, j)] = -s;
      t.planes[t.index(3, i, j)] = gas.R / r;
      const double mu = 1.8e-5 * std::pow(T / 300.0, 0.7);
      t.planes[t.index(4, i, j)] = mu;
      t.planes[t.index(5, i, j)] = mu * (gas.cv(T) + gas.R) / 0.72;
      const double tht = T * t.planes[t.index(2, i, j)];
      const double en = H - tht;
      emin = std::min(emin, en);
      emax = std::max(emax, en);
      scale = std::max(scale, std::fabs(H) + std::fabs(tht));
    }
  }
  const double pad = 64 * 2.220446049250313e-16 * scale;
  emin -= pad;
  emax += pad;
  const double step = (emax - emin) / (ny - 1);
  for (int j = 0; j < ny; ++j) t.e[j] = emin + j * step;
  for (int i = 0; i < nx; ++i) {
    const double rho = t.rho[i];
    const double lowT = t.T[0], highT = t.T[ny - 1];
    const double lowE = at(t, rho, lowT).energy(), highE = at(t, rho, highT).energy();
    for (int j = 0; j < ny; ++j) {
      const double en = t.e[j];
      double lo = lowT, hi = highT;
      if (en <= lowE) hi = lo;
      else if (en >= highE) lo = hi;
      else
        for (int it = 0; it < 60; ++it) {
          const double mid = lo + (hi - lo) / 2;
          if (at(t, rho, mid).energy() < en) lo = mid;
          else hi = mid;
        }
      t.inv[(size_t)j * nx + i] = lo + (hi - lo) / 2;
    }
  }
  return t;
}

}  // namespace lte
}  // namespace oracle
