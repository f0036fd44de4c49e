// TEST INFRASTRUCTURE ONLY.  CPU FP64 restatement of the reference's DGSEM
// RHS operator (chess-uiuc/Theseus): geometry cache on affine elements, the
// Euler and BR1 Navier-Stokes operators, Persson-Peraire blending with
// subcell FV, the stability estimate and the SSP-RK3 step.  The product
// never links or calls this code.
//
// Restated files: include/dgsem_cache_utilities.hpp, DGSEMIntegrator.hpp,
// EulerOperator_impl.hpp, NSOperator_impl.hpp, RHSOperator_impl.hpp,
// StabilityEstimate.hpp, src/ModalBasis.cpp, and MFEM's RK3SSPSolver (Shu-Osher
// form, selected at src/Simulation.cpp:265-268).
//
// Differences from the reference that are intentional and bounded:
//  * GLL nodes come from a long-double Newton solve, not MFEM's
//    IntegrationRules (both are correctly rounded to within an ulp).
//  * The face list is visited element by element (low-side element first),
//    each interior face with the low-xi element as the "minus" side; MFEM's
//    face numbering and Elem1/Elem2 choice may differ.  All face fluxes pass
//    the normal-reversal test (F*(qR,qL,-n) = -F*(qL,qR,n)) at 1e-12, so the
//    two orders agree to rounding.
//  * Affine elements: the subcell FV normals are the adjugate rows, which the
//    reference's ComputeSubcellMetrics (dgsem_cache_utilities.hpp:641-763)
//    reproduces up to the rounding of sum_m D(l,m) (zero by construction).
#pragma once

#include <array>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle_lte.hpp"

#include "../include/theseus_rhs.h"
#include "oracle_physics.hpp"
#include "oracle_bc.hpp"

namespace oracle {

// ---------------------------------------------------------------------------
// 1-D GLL basis on [0,1] and the derivative operators
// ---------------------------------------------------------------------------
struct Basis {
  int p = 0, Np = 0;
  std::vector<double> x, w;          // GLL nodes/weights on [0,1] (weights sum to 1)
  std::vector<double> D, Dhat, Dhat2;  // D[i*Np+l] = d l_l/dx at x_i (reference's D_T(i,l) before transpose)
  std::vector<double> Vinv1;         // not used by the operator; kept for tests

  explicit Basis(int p_ = 1) { build(p_); }

  void build(int p_) {
    p = p_;
    Np = p + 1;
    x.assign(Np, 0.0);
    w.assign(Np, 0.0);
    // Legendre-Gauss-Lobatto nodes on [-1,1]: endpoints and roots of P'_p.
    const int N = p;
    std::vector<long double> xr(Np), wr(Np);
    for (int i = 0; i < Np; ++i) {
      long double z = -std::cos((long double)M_PI * i / N);
      if (i > 0 && i < N) {
        for (int it = 0; it < 200; ++it) {
          // P_N, P_N', P_N'' via recursion
          long double P0 = 1, P1 = z;
          for (int k = 1; k < N; ++k) {
            long double P2 = ((2 * k + 1) * z * P1 - k * P0) / (k + 1);
            P0 = P1;
            P1 = P2;
          }
          // P1 = P_N(z), P0 = P_{N-1}(z)
          const long double dP = N * (z * P1 - P0) / (z * z - 1);
          const long double d2P = (2 * z * dP - N * (N + 1) * P1) / (1 - z * z);
          const long double dz = dP / d2P;
          z -= dz;
          if (std::fabs((double)dz) < 1e-22) break;
        }
      }
      xr[i] = z;
      long double P0 = 1, P1 = z;
      for (int k = 1; k < N; ++k) {
        long double P2 = ((2 * k + 1) * z * P1 - k * P0) / (k + 1);
        P0 = P1;
        P1 = P2;
      }
      if (N == 0) P1 = 1;
      wr[i] = 2.0L / (N * (N + 1) * P1 * P1);
    }
    for (int i = 0; i < Np; ++i) {
      x[i] = (double)((1.0L + xr[i]) / 2.0L);
      w[i] = (double)(wr[i] / 2.0L);
    }
    x[0] = 0.0;
    x[N] = 1.0;
    for (int i = 0; i < Np / 2; ++i) {  // exact symmetry
      x[N - i] = 1.0 - x[i];
      w[N - i] = w[i];
    }
    if (Np % 2 == 1) x[N / 2] = 0.5;

    // dgsem_cache_utilities.hpp:124-181
    std::vector<double> wBary(Np, 1.0);
    for (int i = 1; i < Np; i++)
      for (int j = 0; j < i; j++) {
        wBary[j] *= (x[j] - x[i]);
        wBary[i] *= (x[i] - x[j]);
      }
    for (int i = 0; i < Np; ++i) wBary[i] = 1.0 / wBary[i];
    std::vector<double> DT(Np * Np, 0.0);  // DT[i*Np+iL] = D_T(i,iL)
    for (int iL = 0; iL < Np; iL++)
      for (int i = 0; i < Np; i++)
        if (iL != i) {
          DT[i * Np + iL] = wBary[iL] / wBary[i] / (x[i] - x[iL]);
          DT[i * Np + i] -= DT[i * Np + iL];
        }
    D = DT;
    Dhat = DT;
    Dhat[0] += 1.0 / w[0];
    Dhat[(Np - 1) * Np + Np - 1] -= 1.0 / w[Np - 1];
    Dhat2.assign(Np * Np, 0.0);
    for (int k = 0; k < Np * Np; ++k) Dhat2[k] = DT[k] * 2.0;
    Dhat2[0] += 1.0 / w[0];
    Dhat2[(Np - 1) * Np + Np - 1] -= 1.0 / w[Np - 1];
  }
  // Coefficient names as the reference's device arrays are indexed:
  //   D_d[l + Np*i] = D(i,l), Dhat_d[l + Np*i] = Dhat(i,l), Dhat2_d[m + Np*i] = Dhat2(i,m)
  double Dd(int i, int l) const { return D[i * Np + l]; }
  double Dhatd(int i, int l) const { return Dhat[i * Np + l]; }
  double Dhat2d(int i, int m) const { return Dhat2[i * Np + m]; }
};

// StabilityEstimate.hpp:61-91
inline double ReferenceAdvectionSpectralScale(int order) {
  static const double calibrated[] = {0.0,           2.19706891727, 5.41995189335, 9.64849524786, 14.7297419321,
                                      20.5984751391, 27.2141622680, 34.5459214016, 42.5691944324, 51.2637780299,
                                      60.6126232210, 70.6010566044, 81.2162499718};
  if (order > 0 && order < 13) return 1.05 * calibrated[order];
  const double points = std::max(1, order + 1);
  return 0.65 * points * points;
}
inline double ReferenceBR1DiffusionSpectralScale(int order) {
  static const double calibrated[] = {0.0,           4.0,           25.5969429388, 82.9000427145, 203.470468982,
                                      426.230394541, 799.733889878, 1383.0975821,  2244.38177282, 3461.2550528,
                                      5121.29136549, 7321.77632194, 10169.7148876};
  if (order > 0 && order < 13) return 1.25 * calibrated[order];
  const double points = std::max(1, order + 1);
  return 0.5 * points * points * points * points;
}

// MFEM Poly_1D::CalcLegendre: P_n(2x-1) by the three-term recursion.
inline void CalcLegendre01(int p, double x, double *u) {
  u[0] = 1.0;
  if (p == 0) return;
  const double z = 2.0 * x - 1.0;
  u[1] = z;
  for (int n = 1; n < p; n++) u[n + 1] = ((2 * n + 1) * z * u[n] - n * u[n - 1]) / (n + 1);
}

// ---------------------------------------------------------------------------
// Case: geometry cache + operator settings, built from the C-ABI descriptor
// ---------------------------------------------------------------------------
struct Case {
  int dim, p, Np, ndof, neq, ne, nfp;
  int flow, vol_flux, face_flux, blending;
  double alpha_max, alpha_min, sharpness_fac, modalThreshold;
  Gas gas;
  const lte::Table *lte = nullptr;  // tabulated LTE gas (its flux pairing in vol_flux, face_flux) when set
  Basis basis;
  std::vector<double> J, adj;  // per element
  std::vector<int> nbr;        // [ne][2*dim]
  std::vector<int> bc_type;
  std::vector<double> bc_data;  // [num_bc][BC_STRIDE]
  // Persson-Peraire cache (dgsem_cache_utilities.hpp:867-943)
  std::vector<double> modal, keep_M1, keep_M2;
  double adv_scale, diff_scale, surf_scale;
  // Interior faces of direction d (by low-side element) split into colours
  // so that no two faces of one colour share an element: faces of one colour
  // can be processed in parallel.  The colour is the parity of the low-side
  // element's position along its chain of d-neighbours; a closed chain of odd
  // length puts its last face in a third colour.
  std::vector<std::vector<int>> face_cols[3];

  void build_face_colours() {
    for (int dd = 0; dd < 3; ++dd) face_cols[dd].assign(3, {});
    for (int dd = 0; dd < dim; ++dd) {
      std::vector<int> pos(ne, -1);
      for (int e = 0; e < ne; ++e) {
        if (pos[e] >= 0) continue;
        int start = e, guard = 0;
        while (true) {  // walk back to the start of the chain (or once around a ring)
          const int prev = nbr[(size_t)start * 2 * dim + 2 * dd];
          if (prev < 0 || prev == e || guard++ > ne) break;
          start = prev;
        }
        std::vector<int> chain;
        int cur = start;
        while (cur >= 0 && pos[cur] < 0) {
          pos[cur] = (int)chain.size();
          chain.push_back(cur);
          cur = nbr[(size_t)cur * 2 * dim + 2 * dd + 1];
        }
        const bool ring = cur == start;
        const int len = (int)chain.size();
        for (int k = 0; k < len; ++k) {
          const int el = chain[k];
          if (nbr[(size_t)el * 2 * dim + 2 * dd + 1] < 0) continue;
          int col = k & 1;
          if (ring && (len & 1) && k == len - 1) col = 2;
          if (ring && len == 1) col = 0;
          face_cols[dd][col].push_back(el);
        }
      }
    }
  }

  explicit Case(const theseus_case_desc &d) : basis(d.order) {
    // the oracle works in the aligned frame only (the engine's orientation
    // pass is checked against it by permuting the states)
    if (d.elem_orient)
      for (int e = 0; e < d.num_elements; ++e)
        if (d.elem_orient[e] != 0) throw std::invalid_argument("the oracle takes aligned meshes (elem_orient all zero)");
    dim = d.dim;
    p = d.order;
    Np = p + 1;
    ndof = 1;
    nfp = 1;
    for (int k = 0; k < dim; ++k) ndof *= Np;
    for (int k = 0; k < dim - 1; ++k) nfp *= Np;
    neq = dim + 2;
    ne = d.num_elements;
    flow = d.flow_model;
    vol_flux = d.volume_flux;
    face_flux = d.surface_flux;
    blending = d.blending;
    alpha_max = d.alpha_max;
    alpha_min = d.alpha_min;
    sharpness_fac = 9.21024;                                            // RHSOperator.hpp:33
    modalThreshold = 0.5 * std::pow(10.0, -1.8 * std::pow(p, 0.25));  // RHSOperator.hpp:53
    gas = Gas(dim, d.gamma, d.Pr, d.R_gas, d.mu, d.mu_bulk);
    J.assign(d.elem_jac, d.elem_jac + ne);
    adj.assign(d.elem_adj, d.elem_adj + (size_t)ne * dim * dim);
    nbr.assign(d.face_neighbor, d.face_neighbor + (size_t)ne * 2 * dim);
    if (d.num_bc > 0) bc_type.assign(d.bc_type, d.bc_type + d.num_bc);
    bc_data.assign((size_t)std::max(d.num_bc, 0) * BC_STRIDE, 0.0);
    if (d.num_bc > 0 && d.bc_data) bc_data.assign(d.bc_data, d.bc_data + (size_t)d.num_bc * BC_STRIDE);
    adv_scale = ReferenceAdvectionSpectralScale(p);
    diff_scale = ReferenceBR1DiffusionSpectralScale(p);
    surf_scale = adv_scale / (1.0 / basis.w[0]);
    build_persson();
    build_face_colours();
  }

  // Modal transform: V(i,k) = prod_j L_{deg(k,j)}(x_j(i)) (ModalBasis.cpp:47-140),
  // inverted by Gauss-Jordan with partial pivoting in long double.
  void build_persson() {
    const int n = ndof;
    std::vector<std::array<int, 3>> degs(n);
    {
      int m = 0;
      if (dim == 2) {
        for (int i = 0; i < Np; i++)
          for (int j = 0; j < Np; j++) degs[m++] = {i, j, 0};
      } else {
        for (int i = 0; i < Np; i++)
          for (int j = 0; j < Np; j++)
            for (int k = 0; k < Np; k++) degs[m++] = {i, j, k};
      }
    }
    std::vector<long double> V((size_t)n * n), I((size_t)n * n, 0.0L);
    std::vector<double> L(Np);
    for (int node = 0; node < n; ++node) {
      int idx[3] = {node % Np, (node / Np) % Np, node / (Np * Np)};
      for (int j = 0; j < dim; ++j) {
        CalcLegendre01(p, basis.x[idx[j]], L.data());
        for (int k = 0; k < n; ++k) {
          const double v = L[degs[k][j]];
          V[(size_t)node * n + k] = (j == 0) ? v : (double)(V[(size_t)node * n + k] * v);
        }
      }
    }
    for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0L;
    for (int c = 0; c < n; ++c) {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (std::fabs((double)V[(size_t)r * n + c]) > std::fabs((double)V[(size_t)piv * n + c])) piv = r;
      if (piv != c)
        for (int k = 0; k < n; ++k) {
          std::swap(V[(size_t)c * n + k], V[(size_t)piv * n + k]);
          std::swap(I[(size_t)c * n + k], I[(size_t)piv * n + k]);
        }
      const long double inv = 1.0L / V[(size_t)c * n + c];
      for (int k = 0; k < n; ++k) {
        V[(size_t)c * n + k] *= inv;
        I[(size_t)c * n + k] *= inv;
      }
      for (int r = 0; r < n; ++r)
        if (r != c) {
          const long double f = V[(size_t)r * n + c];
          if (f == 0.0L) continue;
          for (int k = 0; k < n; ++k) {
            V[(size_t)r * n + k] -= f * V[(size_t)c * n + k];
            I[(size_t)r * n + k] -= f * I[(size_t)c * n + k];
          }
        }
    }
    // modal[q*n + m] = Vinv(m, q) (transposed store, dgsem_cache_utilities.hpp:896-906)
    modal.assign((size_t)n * n, 0.0);
    for (int q = 0; q < n; ++q)
      for (int m = 0; m < n; ++m) modal[(size_t)q * n + m] = (double)I[(size_t)m * n + q];
    keep_M1.assign(n, 0.0);
    keep_M2.assign(n, 0.0);
    for (int m = 0; m < n; ++m) {
      bool k1 = true, k2 = true;
      for (int d = 0; d < dim; ++d) {
        if (degs[m][d] > p - 2) k2 = false;
        if (degs[m][d] > p - 1) k1 = false;
      }
      keep_M1[m] = k1 ? 1.0 : 0.0;
      keep_M2[m] = k2 ? 1.0 : 0.0;
    }
  }

  size_t ntot() const { return (size_t)ne * ndof; }
  const double *adj_row(int e, int dir) const { return &adj[((size_t)e * dim + dir) * dim]; }
  int node_of(int d, int side, int t) const {
    const int id = side ? Np - 1 : 0;
    if (dim == 2) return d == 0 ? id + Np * t : t + Np * id;
    if (d == 0) return id + Np * t;
    if (d == 1) return (t % Np) + Np * id + Np * Np * (t / Np);
    return t + Np * Np * id;
  }
  void gather(const double *U, int e, int node, double *q) const {
    for (int k = 0; k < neq; ++k) q[k] = U[k * ntot() + (size_t)e * ndof + node];
  }
  void line(int point, int &i, int &j, int &k) const {
    i = point % Np;
    j = (dim > 1) ? (point / Np) % Np : 0;
    k = (dim > 2) ? point / (Np * Np) : 0;
  }
};

// Flux and sound speed of the case's gas: the ideal gas (ABI flux ids) or
// the tabulated LTE gas with the pairings the reference's factory builds for
// it (SimFactory_impl.hpp:140-245): LTEEC (LTEECFlux.hpp:93-107), Pirozzoli,
// or the central volume flux with an HLL or LLF face, the last three on the
// table's pressure, sound speed and cp/cv (lte::TableGas).
inline void CaseVolumeFlux(const Case &c, const double *q1, const double *q2, const double *m1, const double *m2,
                           double *F) {
  if (!c.lte) VolumeFlux(c.vol_flux, c.gas, q1, q2, m1, m2, F);
  else if (c.vol_flux == THESEUS_VOLUME_LTEEC) lte::lteec_flux(*c.lte, false, c.dim, q1, q2, m1, m2, F);
  else {
    const lte::TableGas g{c.lte, c.dim, c.gas.mu_bulk};
    if (c.vol_flux == THESEUS_VOLUME_PIROZZOLI) PirozzoliVolume(g, q1, q2, m1, m2, F);
    else CentralVolume(g, q1, q2, m1, m2, F);
  }
}
inline void CaseFaceFlux(const Case &c, const double *q1, const double *q2, const double *n, double *F) {
  if (!c.lte) FaceFlux(c.face_flux, c.gas, q1, q2, n, F);
  else if (c.face_flux == THESEUS_SURFACE_LTEEC) lte::lteec_flux(*c.lte, true, c.dim, q1, q2, n, n, F);
  else {
    const lte::TableGas g{c.lte, c.dim, c.gas.mu_bulk};
    if (c.face_flux == THESEUS_SURFACE_PIROZZOLI) PirozzoliFace(g, q1, q2, n, F);
    else if (c.face_flux == THESEUS_SURFACE_HLL) HLLFace(g, q1, q2, n, F);
    else LLFFace(g, q1, q2, n, F);
  }
}
inline double CaseSoundSpeed(const Case &c, const double *q) {
  if (c.lte) return std::sqrt(lte::evaluate_state(*c.lte, q, c.dim).sound_speed_squared());
  return c.gas.sound_speed(q);
}

// ---------------------------------------------------------------------------
// Volume kernels (E-vector element layout el[q*ndof + node])
// ---------------------------------------------------------------------------

// DGSEMIntegrator.hpp:124-218 (affine metric: every node carries the element adjugate)
inline void VolumePoint(const Case &c, const double *el_u, int e, int point, double *el_du) {
  const int Np = c.Np, dim = c.dim, neq = c.neq, dof = c.ndof;
  int i, j, k;
  c.line(point, i, j, k);
  double sl[MAXEQ], su[MAXEQ], flux[MAXEQ], rate[MAXEQ] = {0, 0, 0, 0, 0};
  const int idx[3] = {i, j, k};
  const int stride[3] = {1, Np, Np * Np};
  for (int dir = 0; dir < dim; ++dir) {
    const int me = idx[dir];
    const int base = point - me * stride[dir];
    const double *met = c.adj_row(e, dir);
    for (int m = 0; m < Np; ++m) {
      if (m == me) continue;
      const int lo = m < me ? m : me;
      const int hi = m < me ? me : m;
      const int lp = base + lo * stride[dir], up = base + hi * stride[dir];
      for (int q = 0; q < neq; ++q) {
        sl[q] = el_u[q * dof + lp];
        su[q] = el_u[q * dof + up];
      }
      CaseVolumeFlux(c, sl, su, met, met, flux);
      const double coef = c.basis.Dhat2d(me, m);
      for (int q = 0; q < neq; ++q) rate[q] += coef * flux[q];
    }
  }
  const double scale = -1.0 / c.J[e];
  for (int q = 0; q < neq; ++q) el_du[point + q * dof] = scale * rate[q];
}

// DGSEMIntegrator.hpp:338-483 (subcell normals = affine adjugate rows)
inline void FVFluxes(const Case &c, const double *el_u, int e, double *el_du) {
  const int dim = c.dim, Np = c.Np, neq = c.neq, npe = c.ndof;
  const double *qW = c.basis.w.data();
  const double J = c.J[e];
  double fnum[MAXEQ], dsub[MAXEQ], s1[MAXEQ], s2[MAXEQ];
  for (int i = 0; i < npe * neq; i++) el_du[i] = 0.0;
  const int Nz = dim > 2 ? Np : 1, Ny = dim > 1 ? Np : 1;
  auto gath = [&](int id, double *s) {
    for (int q = 0; q < neq; ++q) s[q] = el_u[q * npe + id];
  };
  // xi lines: first direction assigns, later directions add
  for (int dir = 0; dir < dim; ++dir) {
    const int stride = dir == 0 ? 1 : (dir == 1 ? Np : Np * Np);
    const double *nor = c.adj_row(e, dir);
    const int nA = dir == 0 ? Ny : Np;       // outer index ranges over the other directions
    const int nB = dir == 2 ? Ny : Nz;
    for (int b = 0; b < nB; ++b)
      for (int a = 0; a < nA; ++a) {
        int id1;
        if (dir == 0) id1 = b * Ny * Np + a * Np;
        else if (dir == 1) id1 = b * Ny * Np + a;
        else id1 = b * Np + a;
        for (int q = 0; q < neq; ++q) dsub[q] = 0.0;
        gath(id1, s1);
        for (int ii = 0; ii < Np - 1; ++ii) {
          const int id2 = id1 + stride;
          gath(id2, s2);
          CaseFaceFlux(c, s1, s2, nor, fnum);
          for (int q = 0; q < neq; ++q) dsub[q] -= fnum[q];
          for (int q = 0; q < neq; ++q) dsub[q] /= (J * qW[ii]);
          for (int q = 0; q < neq; ++q) {
            if (dir == 0) el_du[id1 + q * npe] = dsub[q];
            else el_du[id1 + q * npe] += dsub[q];
          }
          for (int q = 0; q < neq; ++q) {
            dsub[q] = fnum[q];
            s1[q] = s2[q];
          }
          id1 = id2;
        }
        for (int q = 0; q < neq; ++q) dsub[q] /= (J * qW[Np - 1]);
        for (int q = 0; q < neq; ++q) {
          if (dir == 0) el_du[id1 + q * npe] = dsub[q];
          else el_du[id1 + q * npe] += dsub[q];
        }
      }
  }
}

// RHSOperator_impl.hpp:355-527 (indicator rho*p, modal energy ratio, sigmoid)
inline void BlendingCoefficient(const Case &c, const double *U, double *alpha_out, double *eta_out) {
  const int n = c.ndof;
  constexpr int block_size = 256;
#pragma omp parallel
  {
  std::vector<double> ind(n);
  double q[MAXEQ];
#pragma omp for schedule(static)
  for (int e = 0; e < c.ne; ++e) {
    for (int node = 0; node < n; ++node) {
      c.gather(U, e, node, q);
      ind[node] = c.gas.pressure(q) * c.gas.density(q);
    }
    double mm_s[block_size], m1_s[block_size], m2_s[block_size];
    for (int t = 0; t < block_size; ++t) mm_s[t] = m1_s[t] = m2_s[t] = 0.0;
    for (int m = 0; m < n; ++m) {
      double mode = 0.0;
      for (int qq = 0; qq < n; ++qq) mode += c.modal[(size_t)qq * n + m] * ind[qq];
      const int t = m % block_size;
      const double mode2 = mode * mode;
      mm_s[t] += mode2;
      m1_s[t] += c.keep_M1[m] * mode2;
      m2_s[t] += c.keep_M2[m] * mode2;
    }
    for (int stride = block_size / 2; stride > 0; stride /= 2)
      for (int t = 0; t < stride; ++t) {
        mm_s[t] += mm_s[t + stride];
        m1_s[t] += m1_s[t + stride];
        m2_s[t] += m2_s[t + stride];
      }
    const double mm = mm_s[0], m1m1 = m1_s[0], m2m2 = m2_s[0];
    const double eps = 1.0e-30;
    double val = 0.0;
    if (mm > eps) {
      val = 1.0 - m1m1 / mm;
      if (m1m1 > eps) val = std::max(val, 1.0 - m2m2 / m1m1);
      else val = 1.0;
    }
    const double eta = std::min(std::max(val, 0.0), 1.0);
    if (eta_out) eta_out[e] = eta;
    double a = 1.0 / (1.0 + std::exp(-c.sharpness_fac * (eta - c.modalThreshold) / c.modalThreshold));
    if (a < c.alpha_min) a = 0.0;
    else if (a > (1.0 - c.alpha_min)) a = 1.0;
    alpha_out[e] = std::min(a, c.alpha_max);
  }
  }
}

// ---------------------------------------------------------------------------
// Face traversal helpers
// ---------------------------------------------------------------------------
// Interior faces, visited direction by direction and colour by colour (see
// Case::face_cols); each colour runs in parallel.  Every dof therefore
// receives its face contributions in a fixed order (x faces, then y, then z)
// whatever the thread count.
template <typename F>
inline void ForEachInteriorFace(const Case &c, F &&f) {
  for (int d = 0; d < c.dim; ++d) {
    for (int col = 0; col < 3; ++col) {
      const std::vector<int> &list = c.face_cols[d][col];
      const long n = (long)list.size();
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n; ++k) {
        const int e = list[k];
        f(e, c.nbr[(size_t)e * 2 * c.dim + 2 * d + 1], d);
      }
    }
  }
}
template <typename F>
inline void ForEachInteriorFaceSerial(const Case &c, F &&f) {
  for (int d = 0; d < c.dim; ++d)
    for (int e = 0; e < c.ne; ++e) {
      const int r = c.nbr[(size_t)e * 2 * c.dim + 2 * d + 1];
      if (r >= 0) f(e, r, d);
    }
}
template <typename F>
inline void ForEachBoundaryFace(const Case &c, F &&f) {
  for (int e = 0; e < c.ne; ++e)
    for (int d = 0; d < c.dim; ++d)
      for (int s = 0; s < 2; ++s) {
        const int code = c.nbr[(size_t)e * 2 * c.dim + 2 * d + s];
        if (code < 0) f(e, d, s, -(code + 1));
      }
}

// ---------------------------------------------------------------------------
// Euler operator (EulerOperator_impl.hpp:28-421) with blending.
// ---------------------------------------------------------------------------
inline void VolumeWithBlending(const Case &c, const double *U, const double *alpha, std::vector<double> &dUe) {
  const size_t N = c.ntot();
  const int dof = c.ndof, neq = c.neq;
  dUe.assign((size_t)c.ne * dof * neq, 0.0);
#pragma omp parallel
  {
  std::vector<double> el_u((size_t)dof * neq), du_fv((size_t)dof * neq);
#pragma omp for schedule(static)
  for (int e = 0; e < c.ne; ++e) {
    for (int q = 0; q < neq; ++q)
      for (int n = 0; n < dof; ++n) el_u[q * dof + n] = U[q * N + (size_t)e * dof + n];
    double *du = &dUe[(size_t)e * dof * neq];
    for (int pt = 0; pt < dof; ++pt) VolumePoint(c, el_u.data(), e, pt, du);
    if (alpha && alpha[e] > 1e-16) {  // EulerOperator_impl.hpp:134-152
      const double a_fv = alpha[e], a_dg = 1.0 - a_fv;
      FVFluxes(c, el_u.data(), e, du_fv.data());
      for (int v = 0; v < dof * neq; ++v) du[v] = a_dg * du[v] + a_fv * du_fv[v];
    }
  }
  }
}

inline void ScatterE(const Case &c, const std::vector<double> &dUe, double *pdudt) {
  const size_t N = c.ntot();
  const int dof = c.ndof, neq = c.neq;
  for (int e = 0; e < c.ne; ++e)
    for (int q = 0; q < neq; ++q)
      for (int n = 0; n < dof; ++n) pdudt[q * N + (size_t)e * dof + n] += dUe[(size_t)e * dof * neq + q * dof + n];
}

inline void EulerRHS(const Case &c, const double *U, double *dudt, const double *alpha) {
  const size_t N = c.ntot();
  const int neq = c.neq, dim = c.dim;
  const double w0 = c.basis.w[0];
  std::vector<double> dUe;
  VolumeWithBlending(c, U, alpha, dUe);
  for (size_t i = 0; i < N * neq; ++i) dudt[i] = 0.0;  // RHSOperator_impl.hpp:763-768
  ScatterE(c, dUe, dudt);
  // interior faces (EulerOperator_impl.hpp:211-290)
  std::vector<double> fd(N * neq, 0.0);
  double qm[MAXEQ], F[MAXEQ];
  ForEachInteriorFace(c, [&](int L, int R, int d) {
    const double *n = c.adj_row(L, d);
    const double wm = 1.0 / (w0 * c.J[L]), wp = 1.0 / (w0 * c.J[R]);
    double qm[MAXEQ], qp[MAXEQ], F[MAXEQ];
    for (int t = 0; t < c.nfp; ++t) {
      const int nl = c.node_of(d, 1, t), nr = c.node_of(d, 0, t);
      c.gather(U, L, nl, qm);
      c.gather(U, R, nr, qp);
      CaseFaceFlux(c, qm, qp, n, F);
      for (int q = 0; q < neq; ++q) {
        fd[q * N + (size_t)L * c.ndof + nl] += -wm * F[q];
        fd[q * N + (size_t)R * c.ndof + nr] += wp * F[q];
      }
    }
  });
  for (size_t i = 0; i < N * neq; ++i) dudt[i] += fd[i];
  // boundary faces (EulerOperator_impl.hpp:292-392)
  bool any_bnd = false;
  std::fill(fd.begin(), fd.end(), 0.0);
  // (ApplyBoundaryConditionInviscid, bc_kernels.hpp:219-287)
  auto face = [&](const double *a, const double *b, const double *n, double *out) { CaseFaceFlux(c, a, b, n, out); };
  ForEachBoundaryFace(c, [&](int e, int d, int s, int bc) {
    any_bnd = true;
    const int type = c.bc_type.at(bc);
    const double *data = &c.bc_data[(size_t)bc * BC_STRIDE];
    double n[3];
    const double *a = c.adj_row(e, d);
    for (int k = 0; k < dim; ++k) n[k] = s ? a[k] : -a[k];
    const double scale = -1.0 / (w0 * c.J[e]);
    for (int t = 0; t < c.nfp; ++t) {
      const int nd = c.node_of(d, s, t);
      c.gather(U, e, nd, qm);
      BoundaryInviscidFlux(c.gas, face, type, data, qm, n, F);
      for (int q = 0; q < neq; ++q) fd[q * N + (size_t)e * c.ndof + nd] += scale * F[q];
    }
  });
  if (any_bnd)
    for (size_t i = 0; i < N * neq; ++i) dudt[i] += fd[i];
}

// ---------------------------------------------------------------------------
// BR1 Navier-Stokes operator (NSOperator_impl.hpp:10-1105).
// ---------------------------------------------------------------------------
template <class G>
inline void NSRHS(const Case &c, const G &gas, const double *U, double *dudt, const double *alpha) {
  const size_t N = c.ntot();
  const int neq = c.neq, dim = c.dim, dof = c.ndof, Np = c.Np;
  const double w0 = c.basis.w[0];
  auto face = [&](const double *a, const double *b, const double *n, double *out) { CaseFaceFlux(c, a, b, n, out); };
  auto outward = [&](int e, int d, int s, double *n) {
    const double *a = c.adj_row(e, d);
    for (int k = 0; k < dim; ++k) n[k] = s ? a[k] : -a[k];
  };

  // ComputeEntropyState (NSOperator_impl.hpp:10-78)
  std::vector<double> S(N * neq);
#pragma omp parallel for schedule(static)
  for (long pt = 0; pt < (long)N; ++pt) {
    double q[MAXEQ], s[MAXEQ];
    for (int k = 0; k < neq; ++k) q[k] = U[k * N + pt];
    gas.entropy_state(q, s);
    for (int k = 0; k < neq; ++k) S[k * N + pt] = s[k];
  }
  // GradOperator (NSOperator_impl.hpp:194-577)
  std::vector<std::vector<double>> grad(dim, std::vector<double>(N * neq, 0.0));
#pragma omp parallel for schedule(static)
  for (int e = 0; e < c.ne; ++e)
    for (int pt = 0; pt < dof; ++pt) {  // AssembleGradVolumePointKernel (DGSEMIntegrator.hpp:764-850)
      int i, j, k;
      c.line(pt, i, j, k);
      double dxi[MAXEQ] = {0}, deta[MAXEQ] = {0}, dzeta[MAXEQ] = {0};
      const size_t eb = (size_t)e * dof;
      for (int l = 0; l < Np; ++l) {
        const int sample = k * Np * Np + j * Np + l;
        const double coef = c.basis.Dd(i, l);
        for (int qq = 0; qq < neq; ++qq) dxi[qq] += S[qq * N + eb + sample] * coef;
      }
      if (dim > 1)
        for (int l = 0; l < Np; ++l) {
          const int sample = k * Np * Np + l * Np + i;
          const double coef = c.basis.Dd(j, l);
          for (int qq = 0; qq < neq; ++qq) deta[qq] += S[qq * N + eb + sample] * coef;
        }
      if (dim > 2)
        for (int l = 0; l < Np; ++l) {
          const int sample = l * Np * Np + j * Np + i;
          const double coef = c.basis.Dd(k, l);
          for (int qq = 0; qq < neq; ++qq) dzeta[qq] += S[qq * N + eb + sample] * coef;
        }
      const double invJ = 1.0 / c.J[e];
      const double *a = &c.adj[(size_t)e * dim * dim];
      for (int qq = 0; qq < neq; ++qq) {
        const size_t o = qq * N + eb + pt;
        if (dim == 2) {
          grad[0][o] = invJ * (dxi[qq] * a[0] + deta[qq] * a[2]);
          grad[1][o] = invJ * (dxi[qq] * a[1] + deta[qq] * a[3]);
        } else {
          grad[0][o] = invJ * (dxi[qq] * a[0] + deta[qq] * a[3] + dzeta[qq] * a[6]);
          grad[1][o] = invJ * (dxi[qq] * a[1] + deta[qq] * a[4] + dzeta[qq] * a[7]);
          grad[2][o] = invJ * (dxi[qq] * a[2] + deta[qq] * a[5] + dzeta[qq] * a[8]);
        }
      }
    }
  {  // GradOperator_InteriorFaces: jump lift, then p_grad += duInt per direction
    std::vector<std::vector<double>> gi(dim, std::vector<double>(N * neq, 0.0));
    ForEachInteriorFace(c, [&](int L, int R, int d) {
      const double *n = c.adj_row(L, d);
      const double wm = 1.0 / (w0 * c.J[L]), wp = 1.0 / (w0 * c.J[R]);
      for (int t = 0; t < c.nfp; ++t) {
        const int nl = c.node_of(d, 1, t), nr = c.node_of(d, 0, t);
        for (int qq = 0; qq < neq; ++qq) {
          const size_t ol = qq * N + (size_t)L * dof + nl, orr = qq * N + (size_t)R * dof + nr;
          const double jump = 0.5 * (S[orr] - S[ol]);
          for (int dd = 0; dd < dim; ++dd) {
            const double f = jump * n[dd];
            gi[dd][ol] += wm * f;
            gi[dd][orr] += wp * f;
          }
        }
      }
    });
    for (int dd = 0; dd < dim; ++dd)
#pragma omp parallel for schedule(static)
      for (long v = 0; v < (long)(N * neq); ++v) grad[dd][v] += gi[dd][v];
  }
  {  // GradOperator_BoundaryFaces (NSOperator_impl.hpp:286-403): += w (S_bc - S) n
    std::vector<std::vector<double>> gb(dim, std::vector<double>(N * neq, 0.0));
    bool any = false;
    ForEachBoundaryFace(c, [&](int e, int d, int s, int bc) {
      any = true;
      const int type = c.bc_type.at(bc);
      const double *data = &c.bc_data[(size_t)bc * BC_STRIDE];
      double n[3], Sv[MAXEQ], jump[MAXEQ];
      outward(e, d, s, n);
      const double wt = 1.0 / (w0 * c.J[e]);
      for (int t = 0; t < c.nfp; ++t) {
        const int nd = c.node_of(d, s, t);
        for (int qq = 0; qq < neq; ++qq) Sv[qq] = S[qq * N + (size_t)e * dof + nd];
        BoundaryGradJump(c.gas, type, data, Sv, jump);
        for (int dd = 0; dd < dim; ++dd)
          for (int qq = 0; qq < neq; ++qq) gb[dd][qq * N + (size_t)e * dof + nd] += wt * (jump[qq] * n[dd]);
      }
    });
    if (any)
      for (int dd = 0; dd < dim; ++dd)
        for (size_t v = 0; v < N * neq; ++v) grad[dd][v] += gb[dd][v];
  }
  // ComputeGradPrimFromGradEntropy (NSOperator_impl.hpp:80-159)
  for (int dd = 0; dd < dim; ++dd)
#pragma omp parallel for schedule(static)
    for (long pt = 0; pt < (long)N; ++pt) {
      double q[MAXEQ], dS[MAXEQ], dP[MAXEQ];
      for (int k = 0; k < neq; ++k) {
        q[k] = U[k * N + pt];
        dS[k] = grad[dd][k * N + pt];
      }
      gas.grad_entropy_to_grad_prim(q, dS, dP);
      for (int k = 0; k < neq; ++k) grad[dd][k * N + pt] = dP[k];
    }
  // MultCNS_Volume (NSOperator_impl.hpp:844-1073)
  std::vector<double> dUe;
  VolumeWithBlending(c, U, alpha, dUe);
#pragma omp parallel
  {
    std::vector<double> vflux((size_t)dim * neq * dof);
#pragma omp for schedule(static)
    for (int e = 0; e < c.ne; ++e) {
      const size_t eb = (size_t)e * dof;
      for (int pt = 0; pt < dof; ++pt) {  // ComputeViscousVolumeFluxPointKernel (DGSEMIntegrator.hpp:486-524)
        double q[MAXEQ], dq[3][MAXEQ] = {{0}}, pf[MAXEQ][MAXDIM] = {{0}};
        for (int k = 0; k < neq; ++k) {
          q[k] = U[k * N + eb + pt];
          for (int dd = 0; dd < dim; ++dd) dq[dd][k] = grad[dd][k * N + eb + pt];
        }
        ComputeViscousFluxKernel(gas, q, dq[0], dq[1], dq[2], pf);
        for (int dir = 0; dir < dim; ++dir) {
          const double *row = c.adj_row(e, dir);
          for (int qq = 0; qq < neq; ++qq) {
            double f = 0.0;
            for (int ph = 0; ph < dim; ++ph) f += row[ph] * pf[qq][ph];
            vflux[(size_t)(dir * neq + qq) * dof + pt] = f;
          }
        }
      }
      double *du = &dUe[(size_t)e * dof * neq];
      for (int pt = 0; pt < dof; ++pt) {  // AssembleViscousVolumePointKernel (DGSEMIntegrator.hpp:526-622)
        int i, j, k;
        c.line(pt, i, j, k);
        double dv[MAXEQ] = {0, 0, 0, 0, 0};
        for (int l = 0; l < Np; ++l) {
          const int sample = k * Np * Np + j * Np + l;
          const double coef = c.basis.Dhatd(i, l);
          for (int qq = 0; qq < neq; ++qq) dv[qq] += coef * vflux[(size_t)(0 * neq + qq) * dof + sample];
        }
        if (dim > 1)
          for (int l = 0; l < Np; ++l) {
            const int sample = k * Np * Np + l * Np + i;
            const double coef = c.basis.Dhatd(j, l);
            for (int qq = 0; qq < neq; ++qq) dv[qq] += coef * vflux[(size_t)(1 * neq + qq) * dof + sample];
          }
        if (dim > 2)
          for (int l = 0; l < Np; ++l) {
            const int sample = l * Np * Np + j * Np + i;
            const double coef = c.basis.Dhatd(k, l);
            for (int qq = 0; qq < neq; ++qq) dv[qq] += coef * vflux[(size_t)(2 * neq + qq) * dof + sample];
          }
        const double sc = 1.0 / c.J[e];
        for (int qq = 0; qq < neq; ++qq) du[pt + qq * dof] += sc * dv[qq];
      }
    }
  }
  for (size_t v = 0; v < N * neq; ++v) dudt[v] = 0.0;
  ScatterE(c, dUe, dudt);
  // MultCNS_InteriorFaces (NSOperator_impl.hpp:579-713; DGSEMIntegrator.hpp:261-316)
  std::vector<double> fd(N * neq, 0.0);
  ForEachInteriorFace(c, [&](int L, int R, int d) {
    const double *n = c.adj_row(L, d);
    const double wm = 1.0 / (w0 * c.J[L]), wp = 1.0 / (w0 * c.J[R]);
    for (int t = 0; t < c.nfp; ++t) {
      const int nl = c.node_of(d, 1, t), nr = c.node_of(d, 0, t);
      double qm[MAXEQ], qp[MAXEQ], pf[MAXEQ], vm[MAXEQ][MAXDIM], vp[MAXEQ][MAXDIM];
      double gm[3][MAXEQ] = {{0}}, gp[3][MAXEQ] = {{0}};
      for (int k = 0; k < neq; ++k) {
        const size_t ol = k * N + (size_t)L * dof + nl, orr = k * N + (size_t)R * dof + nr;
        qm[k] = U[ol];
        qp[k] = U[orr];
        for (int dd = 0; dd < dim; ++dd) {
          gm[dd][k] = grad[dd][ol];
          gp[dd][k] = grad[dd][orr];
        }
      }
      CaseFaceFlux(c, qm, qp, n, pf);
      ComputeViscousFluxKernel(gas, qm, gm[0], gm[1], gm[2], vm);
      ComputeViscousFluxKernel(gas, qp, gp[0], gp[1], gp[2], vp);
      for (int k = 0; k < neq; ++k)
        for (int dd = 0; dd < dim; ++dd) {
          const double avg = 0.5 * (vm[k][dd] + vp[k][dd]);
          pf[k] -= n[dd] * avg;
        }
      for (int k = 0; k < neq; ++k) {
        fd[k * N + (size_t)L * dof + nl] += -wm * pf[k];
        fd[k * N + (size_t)R * dof + nr] += wp * pf[k];
      }
    }
  });
#pragma omp parallel for schedule(static)
  for (long v = 0; v < (long)(N * neq); ++v) dudt[v] += fd[v];
  // MultCNS_BoundaryFaces (NSOperator_impl.hpp:717-840): -w * ApplyViscousBoundaryCondition
  bool any = false;
  std::fill(fd.begin(), fd.end(), 0.0);
  ForEachBoundaryFace(c, [&](int e, int d, int s, int bc) {
    any = true;
    const int type = c.bc_type.at(bc);
    const double *data = &c.bc_data[(size_t)bc * BC_STRIDE];
    double n[3];
    outward(e, d, s, n);
    const double scale = -1.0 / (w0 * c.J[e]);
    for (int t = 0; t < c.nfp; ++t) {
      const int nd = c.node_of(d, s, t);
      double q[MAXEQ], gq[3][MAXEQ] = {{0}}, F[MAXEQ];
      for (int k = 0; k < neq; ++k) {
        const size_t o = k * N + (size_t)e * dof + nd;
        q[k] = U[o];
        for (int dd = 0; dd < dim; ++dd) gq[dd][k] = grad[dd][o];
      }
      const double *grads[3] = {gq[0], gq[1], gq[2]};
      BoundaryViscousFlux(c.gas, face, type, data, q, grads, n, F);
      for (int k = 0; k < neq; ++k) fd[k * N + (size_t)e * dof + nd] += scale * F[k];
    }
  });
  if (any)
    for (size_t v = 0; v < N * neq; ++v) dudt[v] += fd[v];
}

inline void RHS(const Case &c, const double *U, double *dudt, std::vector<double> *alpha_out = nullptr) {
  std::vector<double> alpha;
  const double *ap = nullptr;
  if (c.blending) {
    alpha.assign(c.ne, 0.0);
    BlendingCoefficient(c, U, alpha.data(), nullptr);
    ap = alpha.data();
  }
  if (c.flow == THESEUS_FLOW_NAVIER_STOKES) {
    if (c.lte) {
      for (int nb : c.nbr)
        if (nb < 0) throw std::runtime_error("the tabulated gas with Navier-Stokes needs a periodic mesh");
      NSRHS(c, lte::TableGas{c.lte, c.dim, c.gas.mu_bulk}, U, dudt, ap);
    } else {
      NSRHS(c, c.gas, U, dudt, ap);
    }
  }
  else EulerRHS(c, U, dudt, ap);
  if (alpha_out) *alpha_out = alpha;
}

// RHSOperator_impl.hpp:10-270
inline void EstimateStability(const Case &c, const double *U, double out[3]) {
  const size_t N = c.ntot();
  const int dim = c.dim, neq = c.neq;
  const bool viscous = c.flow == THESEUS_FLOW_NAVIER_STOKES;
  double adv = 0.0, surf = 0.0, diff = 0.0;
  double q[MAXEQ];
  for (int e = 0; e < c.ne; ++e)
    for (int node = 0; node < c.ndof; ++node) {
      c.gather(U, e, node, q);
      const double inverse_jacobian = 1.0 / c.J[e];
      double directional_sum = 0.0, metric_square_sum = 0.0;
      for (int rd = 0; rd < dim; ++rd) {
        double vdm = 0.0, mn2 = 0.0;
        const double *row = c.adj_row(e, rd);
        for (int pd = 0; pd < dim; ++pd) {
          const double value = row[pd];
          vdm += c.gas.velocity(q, pd) * value;
          mn2 += value * value;
        }
        directional_sum += (std::abs(vdm) + CaseSoundSpeed(c, q) * std::sqrt(mn2)) * inverse_jacobian;
        metric_square_sum += mn2 * inverse_jacobian * inverse_jacobian;
      }
      adv = std::max(adv, c.adv_scale * directional_sum);
      if (viscous) {
        // StabilityEstimate.hpp:23-38
        const double density = q[0];
        double shear, conductivity, gamma, cp;
        if (c.lte) {  // LTETransport.hpp:32-65; gamma = cp/cv of the table (LTEHelmholtzEOS.hpp:71-91)
          const lte::Eval s = lte::evaluate_state(*c.lte, q, dim);
          shear = lte::viscosity(*c.lte, s);
          conductivity = lte::transport(*c.lte, s, 5);
          cp = s.cp();
          gamma = cp / s.cv();
        } else {
          shear = c.gas.viscosity(q);
          conductivity = c.gas.thermal_conductivity(q);
          gamma = c.gas.gamma;
          cp = c.gas.cp;
        }
        const double longitudinal = (2.0 - c.gas.bulk_viscosity(q)) * shear;
        const double momentum_diffusivity = std::max(shear, longitudinal) / density;
        const double thermal_diffusivity = conductivity * gamma / (density * cp);
        const double eff = std::max(momentum_diffusivity, thermal_diffusivity);
        diff = std::max(diff, c.diff_scale * eff * metric_square_sum);
      }
    }
  const double w0 = c.basis.w[0];
  double qm[MAXEQ], qp[MAXEQ];
  ForEachInteriorFaceSerial(c, [&](int L, int R, int d) {
    const double *n = c.adj_row(L, d);
    const double wm = 1.0 / (w0 * c.J[L]), wp = 1.0 / (w0 * c.J[R]);
    for (int t = 0; t < c.nfp; ++t) {
      c.gather(U, L, c.node_of(d, 1, t), qm);
      c.gather(U, R, c.node_of(d, 0, t), qp);
      double n2 = 0.0, vm = 0.0, vp = 0.0;
      for (int k = 0; k < dim; ++k) {
        n2 += n[k] * n[k];
        vm += c.gas.velocity(qm, k) * n[k];
        vp += c.gas.velocity(qp, k) * n[k];
      }
      const double nm = std::sqrt(n2);
      const double ws = std::max(std::abs(vm) + CaseSoundSpeed(c, qm) * nm, std::abs(vp) + CaseSoundSpeed(c, qp) * nm);
      surf = std::max(surf, ws * std::max(std::abs(wm), std::abs(wp)) * c.surf_scale);
    }
  });
  // boundary points (RHSOperator_impl.hpp:176-267): a prescribed or inflow
  // state adds its own wave speed and, when viscous, its diffusivity with the
  // element's metric sum (dgsem_cache_utilities.hpp:430-446)
  ForEachBoundaryFace(c, [&](int e, int d, int s, int bc) {
    const double *a = c.adj_row(e, d);
    const double wt = 1.0 / (w0 * c.J[e]);
    const int type = c.bc_type.at(bc);
    const double *ext = &c.bc_data[(size_t)bc * BC_STRIDE];
    const bool has_ext = type == BC_PRESCRIBED_STATE || type == BC_SUPERSONIC_INFLOW;
    double metric_sum = 0.0;
    const double ij = 1.0 / c.J[e];
    for (int rd = 0; rd < dim; ++rd) {
      double r2 = 0.0;
      for (int pd = 0; pd < dim; ++pd) r2 += c.adj_row(e, rd)[pd] * c.adj_row(e, rd)[pd];
      metric_sum += r2 * ij * ij;
    }
    for (int t = 0; t < c.nfp; ++t) {
      c.gather(U, e, c.node_of(d, s, t), qm);
      double n2 = 0.0, vn = 0.0;
      for (int k = 0; k < dim; ++k) {
        const double nk = s ? a[k] : -a[k];
        n2 += nk * nk;
        vn += c.gas.velocity(qm, k) * nk;
      }
      double ws = std::abs(vn) + CaseSoundSpeed(c, qm) * std::sqrt(n2);
      if (has_ext) {
        if (viscous) {
          const double density = ext[0];
          const double shear = c.gas.viscosity(ext);
          const double longitudinal = (2.0 - c.gas.bulk_viscosity(ext)) * shear;
          const double eff = std::max(std::max(shear, longitudinal) / density,
                                      c.gas.thermal_conductivity(ext) * c.gas.gamma / (density * c.gas.cp));
          diff = std::max(diff, c.diff_scale * eff * metric_sum);
        }
        double ve = 0.0;
        for (int k = 0; k < dim; ++k) ve += c.gas.velocity(ext, k) * (s ? a[k] : -a[k]);
        ws = std::max(ws, std::abs(ve) + CaseSoundSpeed(c, ext) * std::sqrt(n2));
      }
      surf = std::max(surf, ws * std::abs(wt) * c.surf_scale);
    }
  });
  (void)neq;
  out[0] = adv;
  out[1] = surf;
  out[2] = diff;
}

// IntegralMeasures (dgsem_cache.hpp:16-26) of one state, as
// RHSOperator::ComputeIntegralMeasures computes it (RHSOperator_impl.hpp:530-717):
// node quadrature with weights J * w_i w_j w_k (dgsem_cache_utilities.hpp:485-491;
// no axisymmetric multiplier), per-element sums in node order then summed over
// elements in order; maxima start at 0 and minima at 1e32.  Ideal gas:
// ke = 0.5 |m|^2 / rho, p = (gamma - 1)(rhoE - ke), T = p / (rho R)
// (EOS.hpp:90-157); tabulated gas: T recovered from e, p from the table
// (LTEEOS.hpp:74-180).
struct Measures {
  double mass, ke, en, max_press, min_press, max_temp, min_temp, max_dens, min_dens;
};

// Entropy production of a right-hand side (the engine's FP64 check of a
// mixed-precision RHS; no counterpart in the reference): sum over elements
// and nodes of w_i J_e S(U_i) . dU_i and of its absolute values, S the gas's
// entropy variables (EOS.hpp:206-234; the tabulated gas's,
// LTEHelmholtzGasEOS.hpp:144-157), w the GLL weights on [0,1]^dim.
inline void EntropyProduction(const Case &c, const double *U, const double *dU, double out[2]) {
  const lte::TableGas tab{c.lte, c.dim, c.gas.mu_bulk};
  const size_t N = c.ntot();
  const int Np = c.Np, dim = c.dim, neq = c.neq;
  double prod = 0.0, mag = 0.0;
  double q[MAXEQ], S[MAXEQ];
  for (int e = 0; e < c.ne; ++e)
    for (int node = 0; node < c.ndof; ++node) {
      c.gather(U, e, node, q);
      if (c.lte) tab.entropy_state(q, S);
      else c.gas.entropy_state(q, S);
      double sd = 0.0;
      for (int k = 0; k < neq; ++k) sd += S[k] * dU[k * N + (size_t)e * c.ndof + node];
      const int i = node % Np, j = dim > 1 ? (node / Np) % Np : 0, k = dim > 2 ? node / (Np * Np) : 0;
      double w = c.basis.w[i];
      if (dim > 1) w *= c.basis.w[j];
      if (dim > 2) w *= c.basis.w[k];
      prod += w * c.J[e] * sd;
      mag += w * c.J[e] * std::abs(sd);
    }
  out[0] = prod;
  out[1] = mag;
}

inline Measures IntegralMeasures(const Case &c, const double *U) {
  Measures m{0.0, 0.0, 0.0, 0.0, 1e32, 0.0, 1e32, 0.0, 1e32};
  const int Np = c.Np, dim = c.dim;
  double q[MAXEQ];
  for (int e = 0; e < c.ne; ++e) {
    double mass = 0.0, ke_i = 0.0, en = 0.0;
    double min_d = 1e32, max_d = 0.0, min_t = 1e32, max_t = 0.0, min_p = 1e32, max_p = 0.0;
    for (int node = 0; node < c.ndof; ++node) {
      c.gather(U, e, node, q);
      const int i = node % Np, j = dim > 1 ? (node / Np) % Np : 0, k = dim > 2 ? node / (Np * Np) : 0;
      double w = c.basis.w[i];
      if (dim > 1) w *= c.basis.w[j];
      if (dim > 2) w *= c.basis.w[k];
      const double rho = q[0];
      double m2 = 0.0;
      for (int d = 0; d < dim; ++d) m2 += q[1 + d] * q[1 + d];
      const double ke = 0.5 * m2 / rho;
      const double rhoE = q[dim + 1];
      double press, temper;
      if (c.lte) {
        const lte::Eval s = lte::evaluate_state(*c.lte, q, dim);
        press = s.pressure();
        temper = s.T;
      } else {
        press = c.gas.gammaM1 * (rhoE - ke);
        temper = press / (rho * c.gas.R_gas);
      }
      const double measure = c.J[e] * w;
      mass += rho * measure;
      ke_i += ke * measure;
      en += rhoE * measure;
      min_t = std::min(min_t, temper);
      max_t = std::max(max_t, temper);
      min_d = std::min(min_d, rho);
      max_d = std::max(max_d, rho);
      min_p = std::min(min_p, press);
      max_p = std::max(max_p, press);
    }
    m.mass += mass;
    m.ke += ke_i;
    m.en += en;
    m.min_press = std::min(m.min_press, min_p);
    m.max_press = std::max(m.max_press, max_p);
    m.min_temp = std::min(m.min_temp, min_t);
    m.max_temp = std::max(m.max_temp, max_t);
    m.min_dens = std::min(m.min_dens, min_d);
    m.max_dens = std::max(m.max_dens, max_d);
  }
  return m;
}

// mfem::RK3SSPSolver::Step (Shu-Osher form).
inline void SSPRK3Step(const Case &c, double *x, double dt) {
  const size_t M = c.ntot() * c.neq;
  std::vector<double> k(M), y(M);
  RHS(c, x, k.data());
  for (size_t i = 0; i < M; ++i) y[i] = x[i] + dt * k[i];
  RHS(c, y.data(), k.data());
  for (size_t i = 0; i < M; ++i) y[i] += dt * k[i];
  for (size_t i = 0; i < M; ++i) y[i] = 3. / 4 * x[i] + 1. / 4 * y[i];
  RHS(c, y.data(), k.data());
  for (size_t i = 0; i < M; ++i) y[i] += dt * k[i];
  for (size_t i = 0; i < M; ++i) x[i] = 1. / 3 * x[i] + 2. / 3 * y[i];
}

}  // namespace oracle
