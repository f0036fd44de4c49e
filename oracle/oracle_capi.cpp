// TEST INFRASTRUCTURE ONLY.  extern "C" shim over the CPU FP64 restatement
// (oracle_physics.hpp, oracle_operator.hpp) so that pytest (ctypes) and the
// bench's cpu_baseline leg can call it.  The product never loads this library.
#include <algorithm>
#include <cstdio>
#include <exception>
#include <string>

#include "oracle_operator.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

static thread_local std::string g_err;

extern "C" {

const char *oracle_last_error() { return g_err.c_str(); }

int oracle_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

double oracle_log_mean(double x, double y, double eps) { return ComputeLogMean(x, y, eps); }

void oracle_volume_flux(int id, int dim, double gamma, const double *q1, const double *q2, const double *m1,
                        const double *m2, double *F) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  VolumeFlux(id, g, q1, q2, m1, m2, F);
}

void oracle_face_flux(int id, int dim, double gamma, const double *q1, const double *q2, const double *n,
                      double *F) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  FaceFlux(id, g, q1, q2, n, F);
}

void oracle_physical_flux(int dim, double gamma, const double *q, double *F /* [MAXEQ*MAXDIM] */) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  double f[MAXEQ][MAXDIM] = {{0}};
  ComputeInviscidFluxKernel(g, q, f);
  for (int e = 0; e < MAXEQ; ++e)
    for (int d = 0; d < MAXDIM; ++d) F[e * MAXDIM + d] = f[e][d];
}

void oracle_entropy_state(int dim, double gamma, const double *q, double *E) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  g.entropy_state(q, E);
}

void oracle_grad_entropy_to_grad_prim(int dim, double gamma, const double *q, const double *dE, double *dP) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  g.grad_entropy_to_grad_prim(q, dE, dP);
}

void oracle_viscous_flux(int dim, double gamma, double Pr, double R, double mu, double mu_bulk, const double *q,
                         const double *dx, const double *dy, const double *dz, double *F /* [MAXEQ*MAXDIM] */) {
  Gas g(dim, gamma, Pr, R, mu, mu_bulk);
  double f[MAXEQ][MAXDIM] = {{0}};
  ComputeViscousFluxKernel(g, q, dx, dy, dz, f);
  for (int e = 0; e < MAXEQ; ++e)
    for (int d = 0; d < MAXDIM; ++d) F[e * MAXDIM + d] = f[e][d];
}

// Boundary conditions (bc_kernels.hpp, Flow.hpp); gas constants as
// oracle_viscous_flux; `face_id` selects the surface flux F*.
double oracle_slipwall_pstar(int dim, double gamma, const double *q, const double *n) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  return SlipWallPStarNormal(g, q, n);
}

void oracle_bc_inviscid(int face_id, int type, int dim, double gamma, const double *data, const double *q,
                        const double *n, double *F) {
  Gas g(dim, gamma, 0.72, 287.05, 0.02, 2.0 / 3.0);
  auto face = [&](const double *a, const double *b, const double *nn, double *out) { FaceFlux(face_id, g, a, b, nn, out); };
  BoundaryInviscidFlux(g, face, type, data, q, n, F);
}

void oracle_bc_viscous(int face_id, int type, int dim, double gamma, double Pr, double R, double mu, double mu_bulk,
                       const double *data, const double *q, const double *dx, const double *dy, const double *dz,
                       const double *n, double *F) {
  Gas g(dim, gamma, Pr, R, mu, mu_bulk);
  auto face = [&](const double *a, const double *b, const double *nn, double *out) { FaceFlux(face_id, g, a, b, nn, out); };
  const double *grads[3] = {dx, dy, dz};
  BoundaryViscousFlux(g, face, type, data, q, grads, n, F);
}

void oracle_bc_grad_jump(int type, int dim, double gamma, double R, const double *data, const double *S, double *J) {
  Gas g(dim, gamma, 0.72, R, 0.02, 2.0 / 3.0);
  BoundaryGradJump(g, type, data, S, J);
}

// x, w: Np; D, Dhat, Dhat2: Np*Np stored [i*Np + l]
void oracle_basis(int p, double *x, double *w, double *D, double *Dhat, double *Dhat2) {
  Basis b(p);
  for (int i = 0; i < b.Np; ++i) {
    x[i] = b.x[i];
    w[i] = b.w[i];
  }
  for (int k = 0; k < b.Np * b.Np; ++k) {
    D[k] = b.D[k];
    Dhat[k] = b.Dhat[k];
    Dhat2[k] = b.Dhat2[k];
  }
}

void *oracle_case_create(const theseus_case_desc *desc) {
  try {
    return new Case(*desc);
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void oracle_case_destroy(void *h) { delete static_cast<Case *>(h); }

// dudt = L(U); alpha (may be NULL) receives the per-element blending coefficient.
int oracle_rhs(void *h, const double *U, double *dudt, double *alpha) {
  try {
    const Case &c = *static_cast<Case *>(h);
    std::vector<double> a;
    RHS(c, U, dudt, &a);
    if (alpha && !a.empty())
      for (int e = 0; e < c.ne; ++e) alpha[e] = a[e];
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int oracle_blending(void *h, const double *U, double *alpha, double *eta) {
  const Case &c = *static_cast<Case *>(h);
  BlendingCoefficient(c, U, alpha, eta);
  return 0;
}

int oracle_estimate_stability(void *h, const double *U, double *out3) {
  try {
    EstimateStability(*static_cast<Case *>(h), U, out3);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// out9 = {mass, ke, en, max_press, min_press, max_temp, min_temp, max_dens, min_dens}
int oracle_integral_measures(void *h, const double *U, double *out9) {
  try {
    const Measures m = IntegralMeasures(*static_cast<Case *>(h), U);
    std::memcpy(out9, &m, sizeof m);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int oracle_entropy_production(void *h, const double *U, const double *dU, double *out2) {
  try {
    EntropyProduction(*static_cast<Case *>(h), U, dU, out2);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int oracle_ssprk3_step(void *h, double *U, double dt) {
  try {
    SSPRK3Step(*static_cast<Case *>(h), U, dt);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// ---- tabulated LTE gas (oracle_lte.hpp) --------------------------------
void *oracle_lte_synthetic(int nx, int ny, double rho_lo, double rho_hi, double T_lo, double T_hi) {
  try {
    return new lte::Table(lte::synthetic_table(nx, ny, rho_lo, rho_hi, T_lo, T_hi));
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void *oracle_lte_from_arrays(int nx, int ny, const double *rho, const double *T, const double *e,
                             const double *planes, const double *inv) {
  lte::Table *t = new lte::Table();
  t->nx = nx;
  t->ny = ny;
  t->rho.assign(rho, rho + nx);
  t->T.assign(T, T + ny);
  t->e.assign(e, e + ny);
  t->planes.assign(planes, planes + 6 * (size_t)nx * ny);
  t->inv.assign(inv, inv + (size_t)nx * ny);
  return t;
}

void oracle_lte_free(void *t) { delete static_cast<lte::Table *>(t); }

void oracle_lte_arrays(void *h, double *rho, double *T, double *e, double *planes, double *inv) {
  const lte::Table &t = *static_cast<lte::Table *>(h);
  std::copy(t.rho.begin(), t.rho.end(), rho);
  std::copy(t.T.begin(), t.T.end(), T);
  std::copy(t.e.begin(), t.e.end(), e);
  std::copy(t.planes.begin(), t.planes.end(), planes);
  std::copy(t.inv.begin(), t.inv.end(), inv);
}

// out: pressure, energy, entropy, gibbs, cv, c^2, p_rho, p_T, g_rho, g_T, H, H_rho, H_T
int oracle_lte_eval(void *h, double rho, double T, double *out) {
  const lte::Eval s = lte::at(*static_cast<lte::Table *>(h), rho, T);
  if (!s.valid) return -1;
  const double v[13] = {s.pressure(), s.energy(), s.entropy(), s.gibbs(), s.cv(), s.sound_speed_squared(),
                        s.pressure_density_derivative(), s.pressure_temperature_derivative(),
                        s.gibbs_density_derivative(), s.gibbs_temperature_derivative(), s.H<0, 0>(), s.H<1, 0>(),
                        s.H<0, 1>()};
  for (int k = 0; k < 13; ++k) out[k] = v[k];
  return 0;
}

int oracle_lte_recover(void *h, double rho, double e, double *T, int *iterations) {
  const lte::Recovery r = lte::temperature_from_energy(*static_cast<lte::Table *>(h), rho, e);
  *T = r.T;
  if (iterations) *iterations = r.iterations;
  return r.status;
}

int oracle_lte_means(void *h, double r1, double t1, double r2, double t2, double *out2) {
  const lte::Table &t = *static_cast<lte::Table *>(h);
  const lte::Eval a = lte::at(t, r1, t1), b = lte::at(t, r2, t2);
  if (!a.valid || !b.valid) return -1;
  const lte::Means m = lte::thermodynamic_means(t, a, b);
  out2[0] = m.density;
  out2[1] = m.energy;
  return 0;
}

int oracle_lte_flux(void *h, int face, int dim, const double *q1, const double *q2, const double *d1,
                    const double *d2, double *F) {
  try {
    lte::lteec_flux(*static_cast<lte::Table *>(h), face != 0, dim, q1, q2, d1, d2, F);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// The tabulated gas at one conservative state q (oracle_lte.hpp TableGas):
// out = T, p, mu, lambda, cp, cv, the entropy variables [dim+2], and the
// primitive gradient [dim+2] that the entropy-variable gradient dS converts to.
int oracle_lte_gas_point(void *h, int dim, double mu_bulk, const double *q, const double *dS, double *out) {
  try {
    const lte::TableGas g{static_cast<lte::Table *>(h), dim, mu_bulk};
    const lte::Eval s = g.state(q);
    const int neq = dim + 2;
    out[0] = s.T;
    out[1] = s.pressure();
    out[2] = g.viscosity(q);
    out[3] = g.thermal_conductivity(q);
    out[4] = s.cp();
    out[5] = s.cv();
    g.entropy_state(q, out + 6);
    g.grad_entropy_to_grad_prim(q, dS, out + 6 + neq);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// NavierStokesFlux.hpp:160-290 with the tabulated gas; F as oracle_viscous_flux
int oracle_lte_viscous_flux(void *h, int dim, double mu_bulk, const double *q, const double *dx, const double *dy,
                            const double *dz, double *F) {
  try {
    const lte::TableGas g{static_cast<lte::Table *>(h), dim, mu_bulk};
    double f[MAXEQ][MAXDIM] = {{0}};
    ComputeViscousFluxKernel(g, q, dx, dy, dz, f);
    for (int e = 0; e < MAXEQ; ++e)
      for (int d = 0; d < MAXDIM; ++d) F[e * MAXDIM + d] = f[e][d];
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// The table's viscosity by Sutherland's law (the reference's SUTHERLAND build).
void oracle_lte_set_sutherland(void *h, double mu0, double T0, double Ts) {
  lte::Table &t = *static_cast<lte::Table *>(h);
  t.sutherland = true;
  t.mu0 = mu0;
  t.T0 = T0;
  t.Ts = Ts;
}

// The case's gas becomes the table (kept alive by the caller) with the LTEEC flux.
void oracle_case_set_lte(void *c, void *t) { static_cast<Case *>(c)->lte = static_cast<lte::Table *>(t); }

}  // extern "C"
