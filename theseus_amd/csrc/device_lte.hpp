// Device side of the tabulated LTE gas: the six-plane Helmholtz table, the
// bicubic Hermite interpolant of H(rho, T) and its exact divided
// differences, thermodynamics from H, temperature recovery (inverse-table
// guess + Newton, always FP64), the Itoh-Abe discrete gradients behind the
// entropy-conserving means, and the LTEEC flux with Lax-Friedrichs
// dissipation.  The interpolant and the flux arithmetic take the type R:
// double, or float for the mixed-precision form (the planes then come from
// an FP32 copy of the table, and the discrete gradients switch to the
// derivative form below FP32's resolution instead of FP64's).
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
  const float *planes32 = nullptr;                          // the planes rounded to FP32 (mixed precision)
  const double *inv = nullptr;                              // nx * ny temperature guesses
  // per node of the state being evaluated (written by k_lte_nodes): the
  // recovered temperature and the table cell i | j << 16 (-1 outside)
  double *node_T = nullptr;
  int *node_cell = nullptr;
  // the Pirozzoli, HLL and LLF pairs: each node's pressure and sound speed
  // (and, for Pirozzoli, gamma = cp/cv), evaluated by k_lte_nodes; null for LTEEC
  double *node_p = nullptr, *node_c = nullptr, *node_g = nullptr;
  // viscosity of the Navier-Stokes form: the mu plane, or Sutherland's law
  // mu0 (T0 + Ts) (T/T0)^1.5 / (T + Ts) when sutherland is set
  int sutherland = 0;
  double mu0 = 0, T0 = 0, Ts = 0;
};

// LTETransport.hpp:22-29 (viscosity_from_temperature)
__device__ __forceinline__ double sutherland_viscosity(const View &v, double T) {
  const double rel = T / v.T0;
  return v.mu0 * (v.T0 + v.Ts) * rel * sqrt(rel) / (T + v.Ts);
}

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

template <int D, typename R>
__device__ __forceinline__ R pbasis(int side, R x) {
  const R x2 = x * x;
  if constexpr (D == 0) return side == 0 ? 2 * x2 * x - 3 * x2 + 1 : -2 * x2 * x + 3 * x2;
  else if constexpr (D == 1) return side == 0 ? 6 * x2 - 6 * x : -6 * x2 + 6 * x;
  else return side == 0 ? 12 * x - 6 : -12 * x + 6;
}
template <int D, typename R>
__device__ __forceinline__ R sbasis(int side, R x) {
  const R x2 = x * x;
  if constexpr (D == 0) return side == 0 ? x2 * x - 2 * x2 + x : x2 * x - x2;
  else if constexpr (D == 1) return side == 0 ? 3 * x2 - 4 * x + 1 : 3 * x2 - 2 * x;
  else return side == 0 ? 6 * x - 4 : 6 * x - 2;
}
template <int D, typename R>
__device__ __forceinline__ R dpbasis(int side, R x, R y) {
  R v;
  if constexpr (D == 0) v = 2 * (x * x + x * y + y * y) - 3 * (x + y);
  else if constexpr (D == 1) v = 6 * (x + y) - 6;
  else v = 12;
  return side == 0 ? v : -v;
}
template <int D, typename R>
__device__ __forceinline__ R dsbasis(int side, R x, R y) {
  if constexpr (D == 0) return x * x + x * y + y * y - (side == 0 ? 2 : 1) * (x + y) + (side == 0 ? 1 : 0);
  else if constexpr (D == 1) return 3 * (x + y) - (side == 0 ? 4 : 2);
  else return 6;
}
template <int D, typename R>
__device__ __forceinline__ R cscale(R w) {
  return D == 0 ? R(1) : (D == 1 ? w : w * w);
}

// The four planes at the lower corner of one cell
template <typename R = double>
struct Cell {
  const R *H = nullptr;  // plane 0 at the corner; plane k at H + k*plane
  int stride = 0, plane = 0;
  R wr = 0, wt = 0;
};

// the table's planes in the arithmetic type
template <typename R>
__device__ __forceinline__ const R *planes_of(const View &v) {
  if constexpr (sizeof(R) == sizeof(double)) return v.planes;
  else return v.planes32;
}

template <int DR, int DT, typename R>
__device__ __forceinline__ R evaluate(const Cell<R> &c, R xi, R eta) {
  R r = 0;
#pragma unroll
  for (int ds = 0; ds < 2; ++ds) {
    const R dv = pbasis<DR>(ds, xi), dsl = sbasis<DR>(ds, xi);
#pragma unroll
    for (int ts = 0; ts < 2; ++ts) {
      const int o = ds + ts * c.stride;
      const R tv = pbasis<DT>(ts, eta), tsl = sbasis<DT>(ts, eta);
      r += c.H[o] * dv * tv;
      r += (c.H[c.plane + o] * c.wr) * dsl * tv;
      r += (c.H[2 * c.plane + o] * c.wt) * dv * tsl;
      r += (c.H[3 * c.plane + o] * c.wr * c.wt) * dsl * tsl;
    }
  }
  return r / (cscale<DR>(c.wr) * cscale<DT>(c.wt));
}

template <int AXIS, int DR = 0, int DT = 0, typename R>
__device__ __forceinline__ R divided(const Cell<R> &c, R xi1, R eta1, R xi2, R eta2) {
  R r = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const R rv = AXIS == 0 ? dpbasis<DR>(i, xi1, xi2) : pbasis<DR>(i, xi1);
    const R rs = AXIS == 0 ? dsbasis<DR>(i, xi1, xi2) : sbasis<DR>(i, xi1);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const R tv = AXIS == 1 ? dpbasis<DT>(j, eta1, eta2) : pbasis<DT>(j, eta1);
      const R ts = AXIS == 1 ? dsbasis<DT>(j, eta1, eta2) : sbasis<DT>(j, eta1);
      const int o = i + j * c.stride;
      r += c.H[o] * rv * tv;
      r += (c.H[c.plane + o] * c.wr) * rs * tv;
      r += (c.H[2 * c.plane + o] * c.wt) * rv * ts;
      r += (c.H[3 * c.plane + o] * c.wr * c.wt) * rs * ts;
    }
  }
  return r / (cscale<DR>(c.wr) * cscale<DT>(c.wt) * (AXIS == 0 ? c.wr : c.wt));
}

template <typename R = double>
struct Eval {
  Cell<R> cell;
  R xi = 0, eta = 0, rho = 0, T = 0;
  int ci = 0, cj = 0;  // the cell's grid indices
  bool valid = false;
  template <int DR = 0, int DT = 0>
  __device__ R H() const { return evaluate<DR, DT>(cell, xi, eta); }
  __device__ R pressure() const { return rho * rho * H<1, 0>(); }
  __device__ R energy() const { return H() - T * H<0, 1>(); }
  __device__ R gibbs() const { return H() + rho * H<1, 0>(); }
  __device__ R p_rho() const { return 2 * rho * H<1, 0>() + rho * rho * H<2, 0>(); }
  __device__ R p_T() const { return rho * rho * H<1, 1>(); }
  __device__ R g_rho() const { return 2 * H<1, 0>() + rho * H<2, 0>(); }
  __device__ R g_T() const { return H<0, 1>() + rho * H<1, 1>(); }
  __device__ R cv() const { return -T * H<0, 2>(); }
  __device__ R cp() const {  // LTEHelmholtzEOS.hpp:76-80
    const R m = H<1, 1>();
    return cv() + T * rho * rho * m * m / p_rho();
  }
  // LTEHelmholtzEOS.hpp:99-110: viscosity (plane 4) or thermal conductivity
  // (plane 5), bilinear in the cell's coordinates
  template <int PROP>
  __device__ R transport() const {
    const R *p = cell.H + PROP * cell.plane;
    return p[0] * (1 - xi) * (1 - eta) + p[1] * xi * (1 - eta) + p[cell.stride] * (1 - xi) * eta +
           p[cell.stride + 1] * xi * eta;
  }
  __device__ R c2() const {
    const R m = H<1, 1>();
    return p_rho() - rho * rho * m * m / H<0, 2>();
  }
};

// Evaluation at (rho, T) in a known cell (i, j)
template <typename R = double>
__device__ __forceinline__ Eval<R> at_cell(const View &v, R rho, R T, int i, int j) {
  Eval<R> r;
  const int off = j * v.nx + i;
  const R dr = R(v.rho[i + 1] - v.rho[i]), dt = R(v.T[j + 1] - v.T[j]);
  r.cell.H = planes_of<R>(v) + off;
  r.cell.stride = v.nx;
  r.cell.plane = v.nx * v.ny;
  r.cell.wr = dr;
  r.cell.wt = dt;
  r.xi = (rho - R(v.rho[i])) / dr;
  r.eta = (T - R(v.T[j])) / dt;
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
template <typename R = double>
__device__ __forceinline__ Eval<R> at_packed(const View &v, R rho, R T, int cell) {
  Eval<R> r = at_cell<R>(v, rho, T, cell < 0 ? 0 : (cell & 0xFFFF), cell < 0 ? 0 : (cell >> 16));
  if (cell < 0) {
    r.valid = false;
    r.xi = r.eta = R(__builtin_nan(""));
  }
  return r;
}

// LTEHelmholtzEOS.hpp:106-130; an invalid query reads cell (0,0) and is flagged
template <typename R = double>
__device__ __forceinline__ Eval<R> at(const View &v, R rho, R T) {
  const int i = find_cell(v.rho, v.nx, double(rho));
  const int j = find_cell(v.T, v.ny, double(T));
  Eval<R> r = at_cell<R>(v, rho, T, i < 0 ? 0 : i, j < 0 ? 0 : j);
  r.valid = i >= 0 && j >= 0;
  if (!r.valid) r.xi = r.eta = R(__builtin_nan(""));  // every derived quantity is NaN
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
    const Eval<double> s = at<double>(v, rho, T);
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
template <Q P, typename R>
__device__ __forceinline__ R value(const Eval<R> &s) {
  if constexpr (P == Q::pressure) return s.pressure();
  else return s.gibbs();
}
template <Q P, typename R>
__device__ __forceinline__ R density_secant(const Eval<R> &a, const Eval<R> &b, R eta) {
  const Cell<R> &c = a.cell;
  const R r1 = a.rho, r2 = b.rho;
  const R h1 = evaluate<1, 0>(c, a.xi, eta);
  const R h2 = evaluate<1, 0>(c, b.xi, eta);
  const R mean_hr = R(0.5) * (h1 + h2);
  const R div_hr = divided<0, 1, 0>(c, a.xi, eta, b.xi, eta);
  if constexpr (P == Q::pressure) {
    return R(0.5) * (r1 * r1 + r2 * r2) * div_hr + (r1 + r2) * mean_hr;
  } else {
    const R div_h = divided<0>(c, a.xi, eta, b.xi, eta);
    return div_h + R(0.5) * (r1 + r2) * div_hr + mean_hr;
  }
}
template <Q P, typename R>
__device__ __forceinline__ R beta_secant(const Eval<R> &a, const Eval<R> &b, R xi, R rho) {
  const Cell<R> &c = a.cell;
  const R mean_t = R(0.5) * (a.T + b.T);
  const R h1 = evaluate<1, 0>(c, xi, a.eta);
  const R h2 = evaluate<1, 0>(c, xi, b.eta);
  const R mean_hr = R(0.5) * (h1 + h2);
  const R div_hr = divided<1, 1, 0>(c, xi, a.eta, xi, b.eta);
  const R hr_beta = mean_hr - mean_t * div_hr;
  if constexpr (P == Q::pressure) {
    return rho * rho * hr_beta;
  } else {
    const R g1 = evaluate<0, 0>(c, xi, a.eta);
    const R g2 = evaluate<0, 0>(c, xi, b.eta);
    const R mean_h = R(0.5) * (g1 + g2);
    const R div_h = divided<1>(c, xi, a.eta, xi, b.eta);
    return mean_h - mean_t * div_h + rho * hr_beta;
  }
}
template <typename R>
struct Gradient {
  R density = 0, beta = 0;
};
// Below these differences the discrete gradient takes the derivative form
// (LTEHelmholtzEOS.hpp:246-300): the reference's 2.22e-15 + 1.49e-8 x (about
// 10 eps and sqrt(eps) of FP64), and the same multiples of FP32's epsilon
// for the mixed-precision form.
template <typename R>
struct SecantTol {
  static constexpr double abs = 2.22e-15, rel = 1.49e-8;
};
template <>
struct SecantTol<float> {
  static constexpr float abs = 1.19e-6f, rel = 3.45e-4f;
};
template <Q P, typename R>
__device__ __forceinline__ Gradient<R> discrete_gradient(const View &v, const Eval<R> &left, const Eval<R> &right) {
  const bool reverse = left.rho > right.rho || (left.rho == right.rho && left.T > right.T);
  const Eval<R> &a = reverse ? right : left;
  const Eval<R> &b = reverse ? left : right;
  const R b1 = 1 / a.T, b2 = 1 / b.T;
  const R dr = b.rho - a.rho, db = b2 - b1;
  const R rmax = a.rho > b.rho ? a.rho : b.rho;
  const R bmax = b1 > b2 ? b1 : b2;
  const bool dd = fabs(dr) >= R(SecantTol<R>::abs) + R(SecantTol<R>::rel) * rmax;
  const bool bd = fabs(db) >= R(SecantTol<R>::abs) + R(SecantTol<R>::rel) * bmax;
  const R f11 = value<P>(a), f22 = value<P>(b);
  Gradient<R> g;
  const bool same_cell = a.cell.H == b.cell.H;
  if (same_cell) {
    if (dd) {
      const R s1 = density_secant<P>(a, b, a.eta);
      const R s2 = density_secant<P>(a, b, b.eta);
      g.density = R(0.5) * (s1 * b1 + s2 * b2);
    }
    if (bd) {
      const R s1 = beta_secant<P>(a, b, a.xi, a.rho);
      const R s2 = beta_secant<P>(a, b, b.xi, b.rho);
      g.beta = R(0.5) * (s1 + s2);
    }
  } else if (dd || bd) {
    // the cross corners (rho_a, T_b), (rho_b, T_a) lie in the cells (i_a, j_b)
    // and (i_b, j_a): the searches of at() would find exactly these
    const bool ok = a.valid && b.valid;
    const R f12 = value<P>(ok ? at_cell<R>(v, a.rho, b.T, a.ci, b.cj) : at<R>(v, a.rho, b.T));
    const R f21 = value<P>(ok ? at_cell<R>(v, b.rho, a.T, b.ci, a.cj) : at<R>(v, b.rho, a.T));
    if (dd) g.density = R(0.5) * ((f21 * b1 - f11 * b1) + (f22 * b2 - f12 * b2)) / dr;
    if (bd) g.beta = R(0.5) * ((f12 * b2 - f11 * b1) + (f22 * b2 - f21 * b1)) / db;
  }
  if (!dd) {
    if constexpr (P == Q::pressure) g.density = R(0.5) * (a.p_rho() * b1 + b.p_rho() * b2);
    else g.density = R(0.5) * (a.g_rho() * b1 + b.g_rho() * b2);
  }
  if (!bd) {
    R d1, d2;
    if constexpr (P == Q::pressure) {
      d1 = a.p_T();
      d2 = b.p_T();
    } else {
      d1 = a.g_T();
      d2 = b.g_T();
    }
    g.beta = R(0.5) * ((f11 - a.T * d1) + (f22 - b.T * d2));
  }
  return g;
}

// LTEHelmholtzEOS.hpp:386-397.  In FP64 the two discrete gradients are
// taken in separate passes of a loop that is not unrolled: written one after
// the other they share their cell evaluations, which keeps both gradients'
// table values live at once (about 590 registers, spilled to scratch); apart
// they fit the register file (256 VGPRs + 238 AGPRs in the 3-D p = 3 RHS
// kernel, no scratch), and the RHS kernel runs 1.3x faster although the
// shared evaluations are repeated.  Same arithmetic, same results.
template <typename R>
__device__ __forceinline__ void thermodynamic_means(const View &v, const Eval<R> &l, const Eval<R> &r, R &rho_mean,
                                                    R &e_mean) {
  Gradient<R> p, g;
  if constexpr (sizeof(R) == sizeof(double)) {
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
      if (k == 0) p = discrete_gradient<Q::pressure>(v, l, r);
      else g = discrete_gradient<Q::gibbs>(v, l, r);
    }
  } else {
    p = discrete_gradient<Q::pressure>(v, l, r);
    g = discrete_gradient<Q::gibbs>(v, l, r);
  }
  rho_mean = p.density / g.density;
  e_mean = g.beta - p.beta / rho_mean;
}

// LTEECFlux.hpp:11-85 on evaluated endpoints (p and, for faces, c per
// endpoint); pd1, pd2: the pressures of the momentum flux's pressure term
// (p less a constant background in the mixed-precision form, which the
// divergence does not see; p itself in FP64)
template <int DIM, bool FACE, typename R>
__device__ __forceinline__ void lteec_flux(const View &v, const R *q1, const R *q2, const Eval<R> &s1,
                                           const Eval<R> &s2, R p1, R p2, R c1, R c2, const R *dir, R *F,
                                           R pd1, R pd2) {
  R rho_m, e_m;
  thermodynamic_means(v, s1, s2, rho_m, e_m);
  const R pm = R(0.5) * (pd1 + pd2);
  const R ir1 = R(1.0) / q1[0], ir2 = R(1.0) / q2[0];
  R vn = 0, ck = 0, pw = 0, v1n = 0, v2n = 0, n2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R u1 = q1[1 + d] * ir1, u2 = q2[1 + d] * ir2;
    vn += R(0.5) * (u1 + u2) * dir[d];
    ck += R(0.5) * u1 * u2;
    pw += R(0.5) * (p1 * u2 + p2 * u1) * dir[d];
    if (FACE) {
      v1n += u1 * dir[d];
      v2n += u2 * dir[d];
      n2 += dir[d] * dir[d];
    }
  }
  R diss = 0;
  if (FACE) {
    const R nm = sqrt(n2);
    const R sp1 = fabs(v1n) + c1 * nm, sp2 = fabs(v2n) + c2 * nm;
    diss = R(0.5) * (sp1 > sp2 ? sp1 : sp2);
  }
  const R mass = rho_m * vn;
  F[0] = mass - diss * (q2[0] - q1[0]);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const R vel = R(0.5) * (q1[1 + d] * ir1 + q2[1 + d] * ir2);
    F[1 + d] = mass * vel + pm * dir[d] - diss * (q2[1 + d] - q1[1 + d]);
  }
  F[DIM + 1] = mass * (e_m + ck) + pw - diss * (q2[DIM + 1] - q1[DIM + 1]);
}
template <int DIM, bool FACE, typename R>
__device__ __forceinline__ void lteec_flux(const View &v, const R *q1, const R *q2, const Eval<R> &s1,
                                           const Eval<R> &s2, R p1, R p2, R c1, R c2, const R *dir, R *F) {
  lteec_flux<DIM, FACE>(v, q1, q2, s1, s2, p1, p2, c1, c2, dir, F, p1, p2);
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

// After k_lte_nodes, for the pairs other than LTEEC: each node's pressure,
// sound speed and (when node_g is set) gamma = cp/cv at its recovered
// temperature.  A pass of its own: the evaluation needs about 160 registers,
// which inside the recovery would hold the Newton iterations to 3 waves per
// SIMD instead of 7.
template <int DIM>
__global__ void __launch_bounds__(256) k_lte_node_pcg(const double *U, long long Ntot, View v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Ntot) return;
  const Eval<double> ev = at_packed<double>(v, U[t], v.node_T[t], v.node_cell[t]);
  v.node_p[t] = ev.pressure();
  v.node_c[t] = sqrt(ev.c2());
  if (v.node_g) v.node_g[t] = ev.cp() / ev.cv();
}

// ---- the viscous path (LTEGasModel<LTEHelmholtzGasEOS, LTETransport>,
//      LTEGasModel.hpp:339; LTETransport.hpp:32-65 without SUTHERLAND)

// A node's quantities for the viscous flux: velocity, recovered temperature,
// tabulated viscosity and conductivity
template <int DIM, typename R = double>
struct ViscNode {
  R u[DIM];
  R T, mu, kappa;
};

// Entropy variables ((g - |u|^2/2)/T, u/T, -1/T) of a node whose temperature
// and cell were recovered (LTEHelmholtzGasEOS.hpp:144-157; the kinetic term
// is the kinetic energy density over rho, :40-50)
template <int DIM, typename R>
__device__ __forceinline__ void entropy_state(const View &v, const R *q, R T, int cell, R *E) {
  const Eval<R> s = at_packed<R>(v, q[0], T, cell);
  const R beta = 1 / s.T;
  R m2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
  const R kinetic = R(0.5) * m2 / q[0] / q[0];
  E[0] = beta * (s.gibbs() - kinetic);
#pragma unroll
  for (int d = 0; d < DIM; ++d) E[1 + d] = beta * (q[1 + d] / q[0]);
  E[DIM + 1] = -beta;
}

template <int DIM, typename R>
__device__ __forceinline__ ViscNode<DIM, R> visc_node(const View &v, const R *q, R T, int cell) {
  const Eval<R> s = at_packed<R>(v, q[0], T, cell);
  ViscNode<DIM, R> n;
#pragma unroll
  for (int d = 0; d < DIM; ++d) n.u[d] = q[1 + d] / q[0];
  n.T = T;
  n.mu = v.sutherland ? R(sutherland_viscosity(v, double(T))) : s.template transport<4>();
  n.kappa = s.template transport<5>();
  return n;
}

// The mixed-precision form's background pressure: the pressure of node 0 of
// a state (one lane writes it), as the ideal gas's k_pref takes it
template <int DIM>
__global__ void __launch_bounds__(64) k_pref_lte(const double *U, long long Ntot, View v, double *pref);

// A point's evaluated state from its conservative vector (LTEHelmholtzGasEOS.hpp:60-77)
template <int DIM>
__device__ __forceinline__ Eval<double> state_eval(const View &v, const double *q) {
  double m2 = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) m2 += q[1 + d] * q[1 + d];
  const double rho = q[0];
  const double e = (q[DIM + 1] - 0.5 * m2 / rho) / rho;
  const double T = temperature_from_energy(v, rho, e);
  return at<double>(v, rho, T);
}

template <int DIM>
__global__ void __launch_bounds__(64) k_pref_lte(const double *U, long long Ntot, View v, double *pref) {
  if (threadIdx.x != 0) return;
  double q[DIM + 2];
#pragma unroll
  for (int c = 0; c < DIM + 2; ++c) q[c] = U[c * Ntot];
  pref[0] = state_eval<DIM>(v, q).pressure();
}

}  // namespace lte
}  // namespace thx

#pragma clang fp contract(fast)
