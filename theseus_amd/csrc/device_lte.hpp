// Device side of the tabulated LTE gas (FP64): the six-plane Helmholtz
// table, the bicubic Hermite interpolant of H(rho, T) and its exact divided
// differences, thermodynamics from H, temperature recovery (inverse-table
// guess + Newton), the Itoh-Abe discrete gradients behind the
// entropy-conserving means, and the LTEEC flux with Lax-Friedrichs
// dissipation.
//
// Restates include/LTEHelmholtzTable.hpp:84-105, LTEHelmholtzInterpolation.hpp:26-211,
// LTEHelmholtzEOS.hpp:14-397, LTEECFlux.hpp:11-85, DissipationOperator.hpp:15-28.
// A query outside the table (or a recovery that does not converge) yields
// NaN instead of trapping; the grid searches never read outside the grids.
#pragma once

#include <hip/hip_runtime.h>

// No multiply-add contraction in the table arithmetic: the Hermite sums and
// the thermodynamic combinations built on them (c^2, the divided differences)
// cancel strongly, and fusing changes their rounding enough to move the
// right-hand side of a smooth state by ~2e-12 relative (an uncontracted CPU
// build and a contracted one of the same code differ by that much).  Restored
// to the default at the end of this header.
#pragma clang fp contract(off)

namespace thx {
namespace lte {

struct View {
  int nx = 0, ny = 0;
  const double *rho = nullptr, *T = nullptr, *e = nullptr;  // grids
  const double *planes = nullptr;                           // 6 * nx * ny: H, H_rho, H_T, H_rhoT, mu, lambda
  const double *inv = nullptr;                              // nx * ny temperature guesses
  // per node of the state being evaluated (written by k_lte_nodes): the
  // recovered temperature and the table cell i | j << 16 (-1 outside)
  double *node_T = nullptr;
  int *node_cell = nullptr;
};

// bisection: largest i with grid[i] <= value (count-2 at the top end); -1 outside
__device__ __forceinline__ int find_cell(const double *grid, int count, double value) {
  if (!(value >= grid[0]) || !(value <= grid[count - 1])) return -1;
  int lo = 0, hi = count - 1;
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (value < grid[mid]) hi = mid;
    else lo = mid;
  }
  return lo;
}

template <int D>
__device__ __forceinline__ double pbasis(int side, double x) {
  const double x2 = x * x;
  if constexpr (D == 0) return side == 0 ? 2 * x2 * x - 3 * x2 + 1 : -2 * x2 * x + 3 * x2;
  else if constexpr (D == 1) return side == 0 ? 6 * x2 - 6 * x : -6 * x2 + 6 * x;
  else return side == 0 ? 12 * x - 6 : -12 * x + 6;
}
template <int D>
__device__ __forceinline__ double sbasis(int side, double x) {
  const double x2 = x * x;
  if constexpr (D == 0) return side == 0 ? x2 * x - 2 * x2 + x : x2 * x - x2;
  else if constexpr (D == 1) return side == 0 ? 3 * x2 - 4 * x + 1 : 3 * x2 - 2 * x;
  else return side == 0 ? 6 * x - 4 : 6 * x - 2;
}
template <int D>
__device__ __forceinline__ double dpbasis(int side, double x, double y) {
  double v;
  if constexpr (D == 0) v = 2 * (x * x + x * y + y * y) - 3 * (x + y);
  else if constexpr (D == 1) v = 6 * (x + y) - 6;
  else v = 12;
  return side == 0 ? v : -v;
}
template <int D>
__device__ __forceinline__ double dsbasis(int side, double x, double y) {
  if constexpr (D == 0) return x * x + x * y + y * y - (side == 0 ? 2 : 1) * (x + y) + (side == 0 ? 1 : 0);
  else if constexpr (D == 1) return 3 * (x + y) - (side == 0 ? 4 : 2);
  else return 6;
}
template <int D>
__device__ __forceinline__ double cscale(double w) {
  return D == 0 ? 1.0 : (D == 1 ? w : w * w);
}

// The four planes at the lower corner of one cell
struct Cell {
  const double *H = nullptr;  // plane 0 at the corner; plane k at H + k*plane
  int stride = 0, plane = 0;
  double wr = 0, wt = 0;
};

template <int DR, int DT>
__device__ __forceinline__ double evaluate(const Cell &c, double xi, double eta) {
  double r = 0;
#pragma unroll
  for (int ds = 0; ds < 2; ++ds) {
    const double dv = pbasis<DR>(ds, xi), dsl = sbasis<DR>(ds, xi);
#pragma unroll
    for (int ts = 0; ts < 2; ++ts) {
      const int o = ds + ts * c.stride;
      const double tv = pbasis<DT>(ts, eta), tsl = sbasis<DT>(ts, eta);
      r += c.H[o] * dv * tv;
      r += (c.H[c.plane + o] * c.wr) * dsl * tv;
      r += (c.H[2 * c.plane + o] * c.wt) * dv * tsl;
      r += (c.H[3 * c.plane + o] * c.wr * c.wt) * dsl * tsl;
    }
  }
  return r / (cscale<DR>(c.wr) * cscale<DT>(c.wt));
}

template <int AXIS, int DR = 0, int DT = 0>
__device__ __forceinline__ double divided(const Cell &c, double xi1, double eta1, double xi2, double eta2) {
  double r = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double rv = AXIS == 0 ? dpbasis<DR>(i, xi1, xi2) : pbasis<DR>(i, xi1);
    const double rs = AXIS == 0 ? dsbasis<DR>(i, xi1, xi2) : sbasis<DR>(i, xi1);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double tv = AXIS == 1 ? dpbasis<DT>(j, eta1, eta2) : pbasis<DT>(j, eta1);
      const double ts = AXIS == 1 ? dsbasis<DT>(j, eta1, eta2) : sbasis<DT>(j, eta1);
      const int o = i + j * c.stride;
      r += c.H[o] * rv * tv;
      r += (c.H[c.plane + o] * c.wr) * rs * tv;
      r += (c.H[2 * c.plane + o] * c.wt) * rv * ts;
      r += (c.H[3 * c.plane + o] * c.wr * c.wt) * rs * ts;
    }
  }
  return r / (cscale<DR>(c.wr) * cscale<DT>(c.wt) * (AXIS == 0 ? c.wr : c.wt));
}

struct Eval {
  Cell cell;
  double xi = 0, eta = 0, rho = 0, T = 0;
  int ci = 0, cj = 0;  // the cell's grid indices
  bool valid = false;
  template <int DR = 0, int DT = 0>
  __device__ double H() const { return evaluate<DR, DT>(cell, xi, eta); }
  __device__ double pressure() const { return rho * rho * H<1, 0>(); }
  __device__ double energy() const { return H() - T * H<0, 1>(); }
  __device__ double gibbs() const { return H() + rho * H<1, 0>(); }
  __device__ double p_rho() const { return 2 * rho * H<1, 0>() + rho * rho * H<2, 0>(); }
  __device__ double p_T() const { return rho * rho * H<1, 1>(); }
  __device__ double g_rho() const { return 2 * H<1, 0>() + rho * H<2, 0>(); }
  __device__ double g_T() const { return H<0, 1>() + rho * H<1, 1>(); }
  __device__ double cv() const { return -T * H<0, 2>(); }
  __device__ double cp() const {  // LTEHelmholtzEOS.hpp:76-80
    const double m = H<1, 1>();
    return cv() + T * rho * rho * m * m / p_rho();
  }
  // LTEHelmholtzEOS.hpp:99-110: viscosity (plane 4) or thermal conductivity
  // (plane 5), bilinear in the cell's coordinates
  template <int PROP>
  __device__ double transport() const {
    const double *p = cell.H + PROP * cell.plane;
    return p[0] * (1 - xi) * (1 - eta) + p[1] * xi * (1 - eta) + p[cell.stride] * (1 - xi) * eta +
           p[cell.stride + 1] * xi * eta;
  }
  __device__ double c2() const {
    const double m = H<1, 1>();
    return p_rho() - rho * rho * m * m / H<0, 2>();
  }
};

// Evaluation at (rho, T) in a known cell (i, j)
__device__ __forceinline__ Eval at_cell(const View &v, double rho, double T, int i, int j) {
  Eval r;
  const int off = j * v.nx + i;
  const double dr = v.rho[i + 1] - v.rho[i], dt = v.T[j + 1] - v.T[j];
  r.cell.H = v.planes + off;
  r.cell.stride = v.nx;
  r.cell.plane = v.nx * v.ny;
  r.cell.wr = dr;
  r.cell.wt = dt;
  r.xi = (rho - v.rho[i]) / dr;
  r.eta = (T - v.T[j]) / dt;
  r.rho = rho;
  r.T = T;
  r.ci = i;
  r.cj = j;
  r.valid = true;
  return r;
}

// Evaluation at (rho, T) whose cell is known (i | j << 16, from
// find_cell of the two grids), or a flagged NaN evaluation for -1, as at()
// gives for a query outside the table.
__device__ __forceinline__ Eval at_packed(const View &v, double rho, double T, int cell) {
  Eval r = at_cell(v, rho, T, cell < 0 ? 0 : (cell & 0xFFFF), cell < 0 ? 0 : (cell >> 16));
  if (cell < 0) {
    r.valid = false;
    r.xi = r.eta = __builtin_nan("");
  }
  return r;
}

// LTEHelmholtzEOS.hpp:106-130; an invalid query reads cell (0,0) and is flagged
__device__ __forceinline__ Eval at(const View &v, double rho, double T) {
  const int i = find_cell(v.rho, v.nx, rho);
  const int j = find_cell(v.T, v.ny, T);
  Eval r = at_cell(v, rho, T, i < 0 ? 0 : i, j < 0 ? 0 : j);
  r.valid = i >= 0 && j >= 0;
  if (!r.valid) r.xi = r.eta = __builtin_nan("");  // every derived quantity is NaN
  return r;
}

// LTEHelmholtzEOS.hpp:151-223; NaN unless converged
__device__ __forceinline__ double temperature_from_energy(const View &v, double rho, double e) {
  const int i = find_cell(v.rho, v.nx, rho);
  const int j = find_cell(v.e, v.ny, e);
  if (i < 0 || j < 0) return __builtin_nan("");
  const double x = (rho - v.rho[i]) / (v.rho[i + 1] - v.rho[i]);
  const double y = (e - v.e[j]) / (v.e[j + 1] - v.e[j]);
  const double *p = v.inv + j * v.nx + i;
  const double lower = p[0] + x * (p[1] - p[0]);
  const double upper = p[v.nx] + x * (p[v.nx + 1] - p[v.nx]);
  double T = lower + y * (upper - lower);
  const double Tlo = v.T[0], Thi = v.T[v.ny - 1];
  for (int it = 0; it < 100; ++it) {
    const Eval s = at(v, rho, T);
    if (!s.valid) return __builtin_nan("");
    const double cv = s.cv(), guess = s.energy();
    if (!(cv > 0) || !isfinite(guess)) return __builtin_nan("");
    const double step = (e - guess) / cv;
    const double next = T + step;
    const bool inside = isfinite(next) && next >= Tlo && next <= Thi;
    if (fabs(step) <= 1e-9 + 1e-12 * T) return next < Tlo ? Tlo : (next > Thi ? Thi : next);
    T = next;
    if (!inside) return __builtin_nan("");
  }
  return __builtin_nan("");
}

enum class Q { pressure, gibbs };
template <Q P>
__device__ __forceinline__ double value(const Eval &s) {
  if constexpr (P == Q::pressure) return s.pressure();
  else return s.gibbs();
}
template <Q P>
__device__ __forceinline__ double density_secant(const Eval &a, const Eval &b, double eta) {
  const Cell &c = a.cell;
  const double r1 = a.rho, r2 = b.rho;
  const double h1 = evaluate<1, 0>(c, a.xi, eta);
  const double h2 = evaluate<1, 0>(c, b.xi, eta);
  const double mean_hr = 0.5 * (h1 + h2);
  const double div_hr = divided<0, 1, 0>(c, a.xi, eta, b.xi, eta);
  if constexpr (P == Q::pressure) {
    return 0.5 * (r1 * r1 + r2 * r2) * div_hr + (r1 + r2) * mean_hr;
  } else {
    const double div_h = divided<0>(c, a.xi, eta, b.xi, eta);
    return div_h + 0.5 * (r1 + r2) * div_hr + mean_hr;
  }
}
template <Q P>
__device__ __forceinline__ double beta_secant(const Eval &a, const Eval &b, double xi, double rho) {
  const Cell &c = a.cell;
  const double mean_t = 0.5 * (a.T + b.T);
  const double h1 = evaluate<1, 0>(c, xi, a.eta);
  const double h2 = evaluate<1, 0>(c, xi, b.eta);
  const double mean_hr = 0.5 * (h1 + h2);
  const double div_hr = divided<1, 1, 0>(c, xi, a.eta, xi, b.eta);
  const double hr_beta = mean_hr - mean_t * div_hr;
  if constexpr (P == Q::pressure) {
    return rho * rho * hr_beta;
  } else {
    const double g1 = evaluate<0, 0>(c, xi, a.eta);
    const double g2 = evaluate<0, 0>(c, xi, b.eta);
    const double mean_h = 0.5 * (g1 + g2);
    const double div_h = divided<1>(c, xi, a.eta, xi, b.eta);
    return mean_h - mean_t * div_h + rho * hr_beta;
  }
}
struct Gradient {
  double density = 0, beta = 0;
};
template <Q P>
__device__ __forceinline__ Gradient discrete_gradient(const View &v, const Eval &left, const Eval &right) {
  const bool reverse = left.rho > right.rho || (left.rho == right.rho && left.T > right.T);
  const Eval &a = reverse ? right : left;
  const Eval &b = reverse ? left : right;
  const double b1 = 1 / a.T, b2 = 1 / b.T;
  const double dr = b.rho - a.rho, db = b2 - b1;
  const double rmax = a.rho > b.rho ? a.rho : b.rho;
  const double bmax = b1 > b2 ? b1 : b2;
  const bool dd = fabs(dr) >= 2.22e-15 + 1.49e-8 * rmax;
  const bool bd = fabs(db) >= 2.22e-15 + 1.49e-8 * bmax;
  const double f11 = value<P>(a), f22 = value<P>(b);
  Gradient g;
  const bool same_cell = a.cell.H == b.cell.H;
  if (same_cell) {
    if (dd) {
      const double s1 = density_secant<P>(a, b, a.eta);
      const double s2 = density_secant<P>(a, b, b.eta);
      g.density = 0.5 * (s1 * b1 + s2 * b2);
    }
    if (bd) {
      const double s1 = beta_secant<P>(a, b, a.xi, a.rho);
      const double s2 = beta_secant<P>(a, b, b.xi, b.rho);
      g.beta = 0.5 * (s1 + s2);
    }
  } else if (dd || bd) {
    // the cross corners (rho_a, T_b), (rho_b, T_a) lie in the cells (i_a, j_b)
    // and (i_b, j_a): the searches of at() would find exactly these
    const bool ok = a.valid && b.valid;
    const double f12 = value<P>(ok ? at_cell(v, a.rho, b.T, a.ci, b.cj) : at(v, a.rho, b.T));
    const double f21 = value<P>(ok ? at_cell(v, b.rho, a.T, b.ci, a.cj) : at(v, b.rho, a.T));
    if (dd) g.density = 0.5 * ((f21 * b1 - f11 * b1) + (f22 * b2 - f12 * b2)) / dr;
    if (bd) g.beta = 0.5 * ((f12 * b2 - f11 * b1) + (f22 * b2 - f21 * b1)) / db;
  }
  if (!dd) {
    if constexpr (P == Q::pressure) g.density = 0.5 * (a.p_rho() * b1 + b.p_rho() * b2);
    else g.density = 0.5 * (a.g_rho() * b1 + b.g_rho() * b2);
  }
  if (!bd) {
    double d1, d2;
    if constexpr (P == Q::pressure) {
      d1 = a.p_T();
      d2 = b.p_T();
    } else {
      d1 = a.g_T();
      d2 = b.g_T();
    }
    g.beta = 0.5 * ((f11 - a.T * d1) + (f22 - b.T * d2));
  }
  return g;
}

// LTEHelmholtzEOS.hpp:386-397
__device__ __forceinline__ void thermodynamic_means(const View &v, const Eval &l, const Eval &r, double &rho_mean,
                                                    double &e_mean) {
  const Gradient p = discrete_gradient<Q::pressure>(v, l, r);
  const Gradient g = discrete_gradient<Q::gibbs>(v, l, r);
  rho_mean = p.density / g.density;
  e_mean = g.beta - p.beta / rho_mean;
}

// LTEECFlux.hpp:11-85 on evaluated endpoints (p and, for faces, c per endpoint)
template <int DIM, bool FACE>
__device__ __forceinline__ void lteec_flux(const View &v, const double *q1, const double *q2, const Eval &s1,
                                           const Eval &s2, double p1, double p2, double c1, double c2,
                                           const double *dir, double *F) {
  double rho_m, e_m;
  thermodynamic_means(v, s1, s2, rho_m, e_m);
  const double pm = 0.5 * (p1 + p2);
  const double ir1 = 1.0 / q1[0], ir2 = 1.0 / q2[0];
  double vn = 0, ck = 0, pw = 0, v1n = 0, v2n = 0, n2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double u1 = q1[1 + d] * ir1, u2 = q2[1 + d] * ir2;
    vn += 0.5 * (u1 + u2) * dir[d];
    ck += 0.5 * u1 * u2;
    pw += 0.5 * (p1 * u2 + p2 * u1) * dir[d];
    if (FACE) {
      v1n += u1 * dir[d];
      v2n += u2 * dir[d];
      n2 += dir[d] * dir[d];
    }
  }
  double diss = 0;
  if (FACE) {
    const double nm = sqrt(n2);
    const double sp1 = fabs(v1n) + c1 * nm, sp2 = fabs(v2n) + c2 * nm;
    diss = 0.5 * (sp1 > sp2 ? sp1 : sp2);
  }
  const double mass = rho_m * vn;
  F[0] = mass - diss * (q2[0] - q1[0]);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double vel = 0.5 * (q1[1 + d] * ir1 + q2[1 + d] * ir2);
    F[1 + d] = mass * vel + pm * dir[d] - diss * (q2[1 + d] - q1[1 + d]);
  }
  F[DIM + 1] = mass * (e_m + ck) + pw - diss * (q2[DIM + 1] - q1[DIM + 1]);
}

// Temperature and table cell of every node of a state (one lane per node):
// the recovery of state_eval, done once per node and stage, for the RHS
// kernel's staged nodes and its neighbours' face points.
template <int DIM>
__global__ void __launch_bounds__(256) k_lte_nodes(const double *U, long long Ntot, View v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Ntot) return;
  double q[DIM + 2];
#pragma unroll
  for (int c = 0; c < DIM + 2; ++c) q[c] = U[c * Ntot + t];
  double m2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
  const double rho = q[0];
  const double e = (q[DIM + 1] - 0.5 * m2 / rho) / rho;
  const double T = temperature_from_energy(v, rho, e);
  const int i = find_cell(v.rho, v.nx, rho);
  const int j = find_cell(v.T, v.ny, T);
  v.node_T[t] = T;
  v.node_cell[t] = (i < 0 || j < 0) ? -1 : (i | (j << 16));
}

// ---- the viscous path (LTEGasModel<LTEHelmholtzGasEOS, LTETransport>,
//      LTEGasModel.hpp:339; LTETransport.hpp:32-65 without SUTHERLAND)

// A node's quantities for the viscous flux: velocity, recovered temperature,
// tabulated viscosity and conductivity
template <int DIM>
struct ViscNode {
  double u[DIM];
  double T, mu, kappa;
};

// Entropy variables ((g - |u|^2/2)/T, u/T, -1/T) of a node whose temperature
// and cell were recovered (LTEHelmholtzGasEOS.hpp:144-157; the kinetic term
// is the kinetic energy density over rho, :40-50)
template <int DIM>
__device__ __forceinline__ void entropy_state(const View &v, const double *q, double T, int cell, double *E) {
  const Eval s = at_packed(v, q[0], T, cell);
  const double beta = 1 / s.T;
  double m2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
  const double kinetic = 0.5 * m2 / q[0] / q[0];
  E[0] = beta * (s.gibbs() - kinetic);
#pragma unroll
  for (int d = 0; d < DIM; ++d) E[1 + d] = beta * (q[1 + d] / q[0]);
  E[DIM + 1] = -beta;
}

template <int DIM>
__device__ __forceinline__ ViscNode<DIM> visc_node(const View &v, const double *q, double T, int cell) {
  const Eval s = at_packed(v, q[0], T, cell);
  ViscNode<DIM> n;
#pragma unroll
  for (int d = 0; d < DIM; ++d) n.u[d] = q[1 + d] / q[0];
  n.T = T;
  n.mu = s.transport<4>();
  n.kappa = s.transport<5>();
  return n;
}

// A point's evaluated state from its conservative vector (LTEHelmholtzGasEOS.hpp:60-77)
template <int DIM>
__device__ __forceinline__ Eval state_eval(const View &v, const double *q) {
  double m2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
  const double rho = q[0];
  const double e = (q[DIM + 1] - 0.5 * m2 / rho) / rho;
  const double T = temperature_from_energy(v, rho, e);
  return at(v, rho, T);
}

}  // namespace lte
}  // namespace thx

#pragma clang fp contract(fast)
