// gfx950 DGSEM right-hand-side kernels.
//
// One element per workgroup (several small 2-D/low-order elements per
// 64-lane wave), one lane per GLL node.  Each workgroup
//   1. loads its element's conservative state with coalesced loads from the
//      byNODES L-vector (for each equation the element's nodes are
//      contiguous) into LDS, and derives velocity and pressure once per node;
//   2. evaluates the split-form volume term sum_m 2 D(i,m) F#(u_lo,u_hi)
//      along the three GLL lines through its node (partners read from LDS);
//   3. evaluates the Riemann flux at every point of its own faces, reading
//      the neighbour's trace straight from the L-vector (L2-resident), so
//      no face-trace arrays and no scatter/add passes exist;
//   4. optionally evaluates the Persson-Peraire indicator and the subcell
//      FV residual and blends;
//   5. writes dU (or the fused SSP-RK3 stage combination) once per node.
// For Navier-Stokes (BR1) a first kernel computes the entropy-variable
// gradient with its face lift, converts it to primitive gradients and
// stores only the physical viscous flux (symmetric stress + energy flux,
// 9 values/node in 3-D); the second kernel is the kernel above plus the
// viscous volume divergence and the central viscous face term.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_physics.hpp"

namespace thx {

constexpr int MAXNP = 8;

template <int DIM, int P>
struct Shape {
  static constexpr int Np = P + 1;
  static constexpr int NDOF = DIM == 2 ? Np * Np : Np * Np * Np;
  static constexpr int NFP = DIM == 2 ? Np : Np * Np;
  static constexpr int NEQ = DIM + 2;
  static constexpr int EPB = NDOF >= 64 ? 1 : 64 / NDOF;  // elements per workgroup
  static constexpr int TPB = ((EPB * NDOF + 63) / 64) * 64;
  static constexpr int NPOW2 = NDOF <= 8 ? 8 : NDOF <= 16 ? 16 : NDOF <= 32 ? 32 : NDOF <= 64 ? 64 : NDOF <= 128 ? 128 : 256;
};

struct BasisArgs {
  double D[MAXNP * MAXNP];      // D(i,l) at [i*Np + l]
  double Dhat[MAXNP * MAXNP];   // D + (1/w0) e0 e0^T - (1/wN) eN eN^T
  double Dhat2[MAXNP * MAXNP];  // 2D with the same corner corrections
  double w[MAXNP];              // GLL weights on [0,1]
};

struct RHSArgs {
  const double *U;    // stage state (L-vector)
  double *out;        // mode 0: dU; mode 1: U + dt dU; mode 2: cx X + cy (U + dt dU)
  const double *X;    // RK base state for mode 2
  double cx, cy, dt;
  int mode;
  const double *Fv;   // NS: viscous flux components (VisComp layout, SoA over nodes)
  double *Fv_out;     // NS pass 1 output
  double *alpha;      // blending coefficient per element (written when blending)
  long long Ntot;     // ne * NDOF
  int ne;
  int blending;
  double alpha_max, alpha_min, sharp, thresh;
  const double *modal;  // [q*NDOF + m] = Vinv(m, q)
  const double *keep1;
  const double *keep2;
  const double *J;      // [ne]
  const double *adj;    // [ne][DIM][DIM]
  const int *nbr;       // [ne][2*DIM]
  GasC<double> gas;
  BasisArgs basis;
};

// Node index of in-face point t on local face (d, side)
template <int DIM, int Np>
__device__ __forceinline__ int face_node(int d, int side, int t) {
  const int id = side ? Np - 1 : 0;
  if constexpr (DIM == 2) {
    return d == 0 ? id + Np * t : t + Np * id;
  } else {
    if (d == 0) return id + Np * t;
    if (d == 1) return (t % Np) + Np * id + Np * Np * (t / Np);
    return t + Np * Np * id;
  }
}

// In-face index of node (i,j,k) on a face normal to d
template <int DIM, int Np>
__device__ __forceinline__ int face_index(int d, int i, int j, int k) {
  if constexpr (DIM == 2) {
    return d == 0 ? j : i;
  } else {
    return d == 0 ? j + Np * k : (d == 1 ? i + Np * k : i + Np * j);
  }
}

// XCD-aware block order: the 8 XCDs each receive a contiguous range of
// elements so that face neighbours share an L2.
__device__ __forceinline__ int logical_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  if ((nb & 7) == 0) return (b & 7) * (nb >> 3) + (b >> 3);
  return b;
}

// Persson-Peraire indicator for one element (RHSOperator_impl.hpp:398-527):
// rho*p per node, modal energy ratios, sigmoid.  Reductions follow the
// reference's power-of-two tree so eta is reproduced bit for bit (the mode
// sums use non-contracted mul/add like the host reference build).
template <int DIM, int P>
__device__ __forceinline__ double blending_alpha(const RHSArgs &a, double (*sred)[Shape<DIM, P>::NPOW2],
                                                 const double *ind, int node, bool lane_ok, bool valid) {
  using S = Shape<DIM, P>;
  constexpr int ND = S::NDOF, NP2 = S::NPOW2;
  double mode = 0.0;
  if (valid) {
    for (int q = 0; q < ND; ++q) mode = __dadd_rn(mode, __dmul_rn(a.modal[q * ND + node], ind[q]));
  }
  const double mode2 = __dmul_rn(mode, mode);
  if (lane_ok) {
    sred[0][node] = valid ? mode2 : 0.0;
    sred[1][node] = valid ? __dmul_rn(a.keep1[node], mode2) : 0.0;
    sred[2][node] = valid ? __dmul_rn(a.keep2[node], mode2) : 0.0;
    for (int t = node + ND; t < NP2; t += ND) {  // zero padding up to the power of two
      sred[0][t] = 0.0;
      sred[1][t] = 0.0;
      sred[2][t] = 0.0;
    }
  }
  __syncthreads();
  for (int stride = NP2 / 2; stride > 0; stride /= 2) {
    if (lane_ok && node < stride) {
      sred[0][node] = __dadd_rn(sred[0][node], sred[0][node + stride]);
      sred[1][node] = __dadd_rn(sred[1][node], sred[1][node + stride]);
      sred[2][node] = __dadd_rn(sred[2][node], sred[2][node + stride]);
    }
    __syncthreads();
  }
  const double mm = sred[0][0], m1m1 = sred[1][0], m2m2 = sred[2][0];
  const double eps = 1.0e-30;
  double val = 0.0;
  if (mm > eps) {
    val = 1.0 - m1m1 / mm;
    if (m1m1 > eps) val = fmax(val, 1.0 - m2m2 / m1m1);
    else val = 1.0;
  }
  const double eta = fmin(fmax(val, 0.0), 1.0);
  double al = 1.0 / (1.0 + exp(-a.sharp * (eta - a.thresh) / a.thresh));
  if (al < a.alpha_min) al = 0.0;
  else if (al > (1.0 - a.alpha_min)) al = 1.0;
  return fmin(al, a.alpha_max);
}

// ---------------------------------------------------------------------------
// Fused volume + face kernel (Euler; Navier-Stokes pass 2 when VISC)
// ---------------------------------------------------------------------------
template <int DIM, int P, int VOL, int SURF, bool VISC>
__global__ void __launch_bounds__((Shape<DIM, P>::TPB)) k_rhs(RHSArgs a) {
  using S = Shape<DIM, P>;
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB;
  constexpr int NV = VisComp<DIM>::N;
  constexpr int NFACE = 2 * DIM * NF;
  constexpr int NI = (Np - 1) * (ND / Np);  // subcell interfaces per direction

  __shared__ double sq[EPB][NQ][ND];
  __shared__ double sv[EPB][DIM][ND];
  __shared__ double sp[EPB][ND];
  __shared__ double sface[EPB][2 * DIM][NF][NQ];
  __shared__ double sD2[Np * Np];
  __shared__ double sDh[Np * Np];
  __shared__ double sFv[VISC ? EPB : 1][VISC ? NV : 1][ND];
  __shared__ double sred[EPB][3][S::NPOW2];
  __shared__ double sfv[EPB][NI][NQ];
  __shared__ double sind[EPB][ND];
  __shared__ int sblend;

  const GasC<double> &g = a.gas;
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const bool lane_ok = el < EPB;
  const int elc = lane_ok ? el : 0;
  const int e = logical_block() * EPB + el;
  const bool valid = lane_ok && e < a.ne;
  const long long Ntot = a.Ntot;

  for (int k = tid; k < Np * Np; k += S::TPB) {
    sD2[k] = a.basis.Dhat2[k];
    sDh[k] = a.basis.Dhat[k];
  }
  if (tid == 0) sblend = 0;

  Pt<double, DIM> me;
  const size_t ebase = (size_t)(valid ? e : 0) * ND + node;
  if (valid) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      me.q[q] = a.U[q * Ntot + ebase];
      sq[el][q][node] = me.q[q];
    }
    derive(g, me);
#pragma unroll
    for (int d = 0; d < DIM; ++d) sv[el][d][node] = me.v[d];
    sp[el][node] = me.p;
    if constexpr (VISC) {
#pragma unroll
      for (int c = 0; c < NV; ++c) sFv[el][c][node] = a.Fv[c * Ntot + ebase];
    }
  }
  __syncthreads();

  const int i = node % Np;
  const int j = DIM > 1 ? (node / Np) % Np : 0;
  const int k = DIM > 2 ? node / (Np * Np) : 0;
  const int idx[3] = {i, j, k};
  constexpr int stride[3] = {1, Np, Np * Np};
  const double *adj_e = a.adj + (size_t)(valid ? e : 0) * DIM * DIM;
  const double Je = valid ? a.J[e] : 1.0;

  // ---- volume: split-form flux differencing (DGSEMIntegrator.hpp:124-218)
  double rate[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) rate[q] = 0.0;
  if (valid) {
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
      double met[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) met[d] = adj_e[dir * DIM + d];
      const int mi = idx[dir];
      const int base = node - mi * stride[dir];
#pragma unroll
      for (int o = 1; o < Np; ++o) {
        int m = mi + o;
        if (m >= Np) m -= Np;
        const int pn = base + m * stride[dir];
        Pt<double, DIM> pt;
#pragma unroll
        for (int q = 0; q < NQ; ++q) pt.q[q] = sq[el][q][pn];
#pragma unroll
        for (int d = 0; d < DIM; ++d) pt.v[d] = sv[el][d][pn];
        pt.p = sp[el][pn];
        const bool lower = m < mi;
        Pt<double, DIM> lo, hi;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          lo.q[q] = lower ? pt.q[q] : me.q[q];
          hi.q[q] = lower ? me.q[q] : pt.q[q];
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          lo.v[d] = lower ? pt.v[d] : me.v[d];
          hi.v[d] = lower ? me.v[d] : pt.v[d];
        }
        lo.p = lower ? pt.p : me.p;
        hi.p = lower ? me.p : pt.p;
        double F[NQ];
        volume_flux<VOL, double, DIM>(g, lo, hi, met, F);
        const double coef = sD2[mi * Np + m];
#pragma unroll
        for (int q = 0; q < NQ; ++q) rate[q] += coef * F[q];
      }
    }
  }
  double du[NQ];
  {
    const double scale = -1.0 / Je;
#pragma unroll
    for (int q = 0; q < NQ; ++q) du[q] = scale * rate[q];
  }

  // ---- blending: indicator and subcell FV (RHSOperator_impl.hpp:355-437,
  //      EulerOperator_impl.hpp:120-157, DGSEMIntegrator.hpp:338-483)
  if (a.blending) {
    if (valid) sind[el][node] = me.p * me.q[0];  // indicator field pressure*density
    __syncthreads();
    const double alpha = blending_alpha<DIM, P>(a, sred[elc], sind[elc], node, lane_ok, valid);
    if (valid && node == 0) a.alpha[e] = alpha;
    if (valid && alpha > 1e-16) sblend = 1;
    __syncthreads();
    if (sblend) {
      double dfv[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) dfv[q] = 0.0;
#pragma unroll
      for (int dir = 0; dir < DIM; ++dir) {
        if (valid) {
          double nor[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) nor[d] = adj_e[dir * DIM + d];
          for (int ii = node; ii < NI; ii += ND) {
            const int L = ii / (Np - 1), pos = ii - L * (Np - 1);
            const int ln = face_node<DIM, Np>(dir, 0, L) + pos * stride[dir];
            const int rn = ln + stride[dir];
            Pt<double, DIM> sl, sr;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
              sl.q[q] = sq[el][q][ln];
              sr.q[q] = sq[el][q][rn];
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              sl.v[d] = sv[el][d][ln];
              sr.v[d] = sv[el][d][rn];
            }
            sl.p = sp[el][ln];
            sr.p = sp[el][rn];
            double F[NQ];
            surface_flux<SURF, double, DIM>(g, sl, sr, nor, F);
#pragma unroll
            for (int q = 0; q < NQ; ++q) sfv[el][ii][q] = F[q];
          }
        }
        __syncthreads();
        if (valid) {
          const int L = face_index<DIM, Np>(dir, i, j, k);
          const int pos = idx[dir];
          const double den = Je * a.basis.w[pos];
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const double fm = pos > 0 ? sfv[el][L * (Np - 1) + pos - 1][q] : 0.0;
            const double fp = pos < Np - 1 ? sfv[el][L * (Np - 1) + pos][q] : 0.0;
            const double val = (fm - fp) / den;
            dfv[q] = dir == 0 ? val : dfv[q] + val;
          }
        }
        __syncthreads();
      }
      if (valid && alpha > 1e-16) {
        const double a_dg = 1.0 - alpha;
#pragma unroll
        for (int q = 0; q < NQ; ++q) du[q] = a_dg * du[q] + alpha * dfv[q];
      }
    }
  }

  // ---- faces (EulerOperator_impl.hpp:211-392; NSOperator_impl.hpp:579-713)
  const double w0 = a.basis.w[0];
  if (valid) {
    const double wown = 1.0 / (w0 * Je);
    for (int fi = node; fi < NFACE; fi += ND) {
      const int f = fi / NF;
      const int t = fi - f * NF;
      const int d = f >> 1, s = f & 1;
      const int on = face_node<DIM, Np>(d, s, t);
      Pt<double, DIM> own;
#pragma unroll
      for (int q = 0; q < NQ; ++q) own.q[q] = sq[el][q][on];
#pragma unroll
      for (int dd = 0; dd < DIM; ++dd) own.v[dd] = sv[el][dd][on];
      own.p = sp[el][on];
      const int nb = a.nbr[(size_t)e * 2 * DIM + f];
      double F[NQ], contrib_scale;
      if (nb >= 0) {
        const int nn = face_node<DIM, Np>(d, 1 - s, t);
        const size_t nbase = (size_t)nb * ND + nn;
        Pt<double, DIM> oth;
#pragma unroll
        for (int q = 0; q < NQ; ++q) oth.q[q] = a.U[q * Ntot + nbase];
        derive(g, oth);
        double n[DIM];
        const double *adj_L = s ? adj_e : a.adj + (size_t)nb * DIM * DIM;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) n[dd] = adj_L[d * DIM + dd];
        if (s) surface_flux<SURF, double, DIM>(g, own, oth, n, F);
        else surface_flux<SURF, double, DIM>(g, oth, own, n, F);
        if constexpr (VISC) {
          // F* - n . 0.5 (Fv_L + Fv_R)   (DGSEMIntegrator.hpp:261-316)
          double vo[NV], vn[NV];
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            vo[c] = sFv[el][c][on];
            vn[c] = a.Fv[c * Ntot + nbase];
          }
          const double *vL = s ? vo : vn;
          const double *vR = s ? vn : vo;
#pragma unroll
          for (int q = 1; q < NQ; ++q)
#pragma unroll
            for (int dd = 0; dd < DIM; ++dd) {
              const int c = (q == NQ - 1) ? VisComp<DIM>::en(dd) : VisComp<DIM>::tau(q - 1, dd);
              const double avg = 0.5 * (vL[c] + vR[c]);
              F[q] -= n[dd] * avg;
            }
        }
        contrib_scale = s ? -wown : wown;
      } else {
        // boundary: supersonic outflow F*(q,q,n_out) (bc_kernels.hpp:236-238)
        double n[DIM];
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) n[dd] = s ? adj_e[d * DIM + dd] : -adj_e[d * DIM + dd];
        surface_flux<SURF, double, DIM>(g, own, own, n, F);
        contrib_scale = -wown;
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q) sface[el][f][t][q] = contrib_scale * F[q];
    }
  }

  // ---- viscous volume divergence (DGSEMIntegrator.hpp:486-622)
  if constexpr (VISC) {
    if (valid) {
      double dv[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) dv[q] = 0.0;
#pragma unroll
      for (int dir = 0; dir < DIM; ++dir) {
        double row[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) row[d] = adj_e[dir * DIM + d];
        const int mi = idx[dir];
        const int base = node - mi * stride[dir];
#pragma unroll
        for (int l = 0; l < Np; ++l) {
          const int sn = base + l * stride[dir];
          const double coef = sDh[mi * Np + l];
#pragma unroll
          for (int q = 1; q < NQ; ++q) {
            double fr = 0.0;
#pragma unroll
            for (int ph = 0; ph < DIM; ++ph) {
              const int c = (q == NQ - 1) ? VisComp<DIM>::en(ph) : VisComp<DIM>::tau(q - 1, ph);
              fr += row[ph] * sFv[el][c][sn];
            }
            dv[q] += coef * fr;
          }
        }
      }
      const double sc = 1.0 / Je;
#pragma unroll
      for (int q = 1; q < NQ; ++q) du[q] += sc * dv[q];
    }
  }
  __syncthreads();

  // ---- gather face contributions and write
  if (valid) {
    double fsum[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) fsum[q] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int t = face_index<DIM, Np>(d, i, j, k);
      if (idx[d] == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) fsum[q] += sface[el][2 * d][t][q];
      }
      if (idx[d] == Np - 1) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) fsum[q] += sface[el][2 * d + 1][t][q];
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const double r = du[q] + fsum[q];
      double o;
      if (a.mode == 0) o = r;
      else if (a.mode == 1) o = me.q[q] + a.dt * r;
      else {
        const double tmp = me.q[q] + a.dt * r;
        o = a.cx * a.X[q * Ntot + ebase] + a.cy * tmp;
      }
      a.out[q * Ntot + ebase] = o;
    }
  }
}

// ---------------------------------------------------------------------------
// Navier-Stokes pass 1 (BR1): entropy variables, their gradient with the
// interior-face lift, conversion to primitive gradients and the physical
// viscous flux (NSOperator_impl.hpp:10-577, DGSEMIntegrator.hpp:764-1044,
// EOS.hpp:206-263, NavierStokesFlux.hpp:389-519).
// ---------------------------------------------------------------------------
template <int DIM, int P>
__global__ void __launch_bounds__((Shape<DIM, P>::TPB)) k_ns_grad(RHSArgs a) {
  using S = Shape<DIM, P>;
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB;
  constexpr int NV = VisComp<DIM>::N;
  constexpr int NFACE = 2 * DIM * NF;

  __shared__ double sS[EPB][NQ][ND];
  __shared__ double sjump[EPB][2 * DIM][NF][NQ];
  __shared__ double sn[EPB][2 * DIM][DIM];
  __shared__ double sD[Np * Np];

  const GasC<double> &g = a.gas;
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const bool lane_ok = el < EPB;
  const int e = logical_block() * EPB + el;
  const bool valid = lane_ok && e < a.ne;
  const long long Ntot = a.Ntot;

  for (int k = tid; k < Np * Np; k += S::TPB) sD[k] = a.basis.D[k];

  double q[NQ];
  const size_t ebase = (size_t)(valid ? e : 0) * ND + node;
  if (valid) {
#pragma unroll
    for (int c = 0; c < NQ; ++c) q[c] = a.U[c * Ntot + ebase];
    double Sv[NQ];
    entropy_state<double, DIM>(g, q, Sv);
#pragma unroll
    for (int c = 0; c < NQ; ++c) sS[el][c][node] = Sv[c];
  }
  const double *adj_e = a.adj + (size_t)(valid ? e : 0) * DIM * DIM;
  const double Je = valid ? a.J[e] : 1.0;
  const double w0 = a.basis.w[0];
  __syncthreads();

  // face lift data: canonical jump 0.5*(S_R - S_L) and the face normal of the low-side element
  if (valid) {
    for (int fi = node; fi < NFACE; fi += ND) {
      const int f = fi / NF;
      const int t = fi - f * NF;
      const int d = f >> 1, s = f & 1;
      const int on = face_node<DIM, Np>(d, s, t);
      const int nb = a.nbr[(size_t)e * 2 * DIM + f];
      // (interior faces only; Navier-Stokes boundaries are rejected at create time)
      const int nn = face_node<DIM, Np>(d, 1 - s, t);
      double qn[NQ], Sn[NQ];
#pragma unroll
      for (int c = 0; c < NQ; ++c) qn[c] = a.U[c * Ntot + (size_t)nb * ND + nn];
      entropy_state<double, DIM>(g, qn, Sn);
#pragma unroll
      for (int c = 0; c < NQ; ++c) {
        const double so = sS[el][c][on];
        sjump[el][f][t][c] = s ? 0.5 * (Sn[c] - so) : 0.5 * (so - Sn[c]);
      }
      if (t == 0) {
        const double *adj_L = s ? adj_e : a.adj + (size_t)nb * DIM * DIM;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) sn[el][f][dd] = adj_L[d * DIM + dd];
      }
    }
  }
  __syncthreads();
  if (!valid) return;

  const int i = node % Np;
  const int j = DIM > 1 ? (node / Np) % Np : 0;
  const int k = DIM > 2 ? node / (Np * Np) : 0;
  const int idx[3] = {i, j, k};
  constexpr int stride[3] = {1, Np, Np * Np};

  // volume gradient (AssembleGradVolumePointKernel)
  double dref[DIM][NQ];
#pragma unroll
  for (int dir = 0; dir < DIM; ++dir) {
#pragma unroll
    for (int c = 0; c < NQ; ++c) dref[dir][c] = 0.0;
    const int mi = idx[dir];
    const int base = node - mi * stride[dir];
#pragma unroll
    for (int l = 0; l < Np; ++l) {
      const double coef = sD[mi * Np + l];
      const int sn_ = base + l * stride[dir];
#pragma unroll
      for (int c = 0; c < NQ; ++c) dref[dir][c] += sS[el][c][sn_] * coef;
    }
  }
  const double invJ = 1.0 / Je;
  double grad[DIM][NQ];
#pragma unroll
  for (int dd = 0; dd < DIM; ++dd)
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
      double acc = dref[0][c] * adj_e[0 * DIM + dd];
#pragma unroll
      for (int r = 1; r < DIM; ++r) acc += dref[r][c] * adj_e[r * DIM + dd];
      grad[dd][c] = invJ * acc;
    }
  // interior-face lift (AssembleGradInteriorFacePointKernel): + w * (jump * n_d)
  {
    const double wown = 1.0 / (w0 * Je);
    double lift[DIM][NQ];
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd)
#pragma unroll
      for (int c = 0; c < NQ; ++c) lift[dd][c] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int t = face_index<DIM, Np>(d, i, j, k);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (idx[d] == (s ? Np - 1 : 0)) {
          const int f = 2 * d + s;
#pragma unroll
          for (int dd = 0; dd < DIM; ++dd) {
            const double nd = sn[el][f][dd];
#pragma unroll
            for (int c = 0; c < NQ; ++c) lift[dd][c] += wown * (sjump[el][f][t][c] * nd);
          }
        }
      }
    }
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd)
#pragma unroll
      for (int c = 0; c < NQ; ++c) grad[dd][c] += lift[dd][c];
  }
  // primitive gradients and viscous flux
  double dP[DIM][NQ];
#pragma unroll
  for (int dd = 0; dd < DIM; ++dd) grad_entropy_to_grad_prim<double, DIM>(g, q, grad[dd], dP[dd]);
  double vf[NV];
  viscous_flux<double, DIM>(g, q, dP, vf);
#pragma unroll
  for (int c = 0; c < NV; ++c) a.Fv_out[c * Ntot + ebase] = vf[c];
}

// ---------------------------------------------------------------------------
// Stability estimate (RHSOperator_impl.hpp:10-270): per-node and per-face
// rates reduced on the device; non-negative doubles are max-reduced through
// their bit patterns with 64-bit integer atomics.
// ---------------------------------------------------------------------------
struct StabArgs {
  const double *U;
  unsigned long long *rates;  // [3] advective, surface, diffusive (bit patterns)
  long long Ntot;
  int ne;
  int viscous;
  double adv_scale, diff_scale, surf_scale;
  const double *J;
  const double *adj;
  const int *nbr;
  GasC<double> gas;
  double w0;
};

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int DIM, int P>
__global__ void __launch_bounds__((Shape<DIM, P>::TPB)) k_stability(StabArgs a) {
  using S = Shape<DIM, P>;
  constexpr int Np = S::Np, ND = S::NDOF, NF = S::NFP, NQ = S::NEQ, EPB = S::EPB;
  const GasC<double> &g = a.gas;
  const int tid = threadIdx.x;
  const int el = tid / ND;
  const int node = tid - el * ND;
  const int e = blockIdx.x * EPB + el;
  const bool valid = el < EPB && e < a.ne;
  double adv = 0.0, surf = 0.0, diff = 0.0;
  if (valid) {
    const long long Ntot = a.Ntot;
    Pt<double, DIM> s;
    const size_t ebase = (size_t)e * ND + node;
#pragma unroll
    for (int c = 0; c < NQ; ++c) s.q[c] = a.U[c * Ntot + ebase];
    derive(g, s);
    const double c_s = sound_speed(g, s);
    const double invJ = 1.0 / a.J[e];
    const double *adj_e = a.adj + (size_t)e * DIM * DIM;
    double dsum = 0.0, msum = 0.0;
#pragma unroll
    for (int rd = 0; rd < DIM; ++rd) {
      double vdm = 0.0, mn2 = 0.0;
#pragma unroll
      for (int pd = 0; pd < DIM; ++pd) {
        const double val = adj_e[rd * DIM + pd];
        vdm += s.v[pd] * val;
        mn2 += val * val;
      }
      dsum += (fabs(vdm) + c_s * sqrt(mn2)) * invJ;
      msum += mn2 * invJ * invJ;
    }
    adv = a.adv_scale * dsum;
    if (a.viscous) {
      const double rho = s.q[0];
      const double shear = g.mu;
      const double longitudinal = (2.0 - g.mu_bulk) * shear;
      const double md = fmax(shear, longitudinal) / rho;
      const double kappa = g.mu * g.cp * g.PrInverse;
      const double td = kappa * g.gamma / (rho * g.cp);
      diff = a.diff_scale * fmax(md, td) * msum;
    }
    // face rates: each node on a face takes the max over the faces it touches
    const int i = node % Np;
    const int j = DIM > 1 ? (node / Np) % Np : 0;
    const int k = DIM > 2 ? node / (Np * Np) : 0;
    const int idx[3] = {i, j, k};
    const double wown = 1.0 / (a.w0 * a.J[e]);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        if (idx[d] != (sd ? Np - 1 : 0)) continue;
        const int nb = a.nbr[(size_t)e * 2 * DIM + 2 * d + sd];
        double n[DIM];
        const double *adj_n = (nb >= 0 && !sd) ? a.adj + (size_t)nb * DIM * DIM : adj_e;
        double n2 = 0.0, vn = 0.0;
#pragma unroll
        for (int dd = 0; dd < DIM; ++dd) {
          n[dd] = (nb < 0 && !sd) ? -adj_e[d * DIM + dd] : adj_n[d * DIM + dd];
          n2 += n[dd] * n[dd];
          vn += s.v[dd] * n[dd];
        }
        const double nm = sqrt(n2);
        double ws = fabs(vn) + c_s * nm;
        double wmax = fabs(wown);
        if (nb >= 0) {
          const int t = face_index<DIM, Np>(d, i, j, k);
          const int nn = face_node<DIM, Np>(d, 1 - sd, t);
          Pt<double, DIM> o;
#pragma unroll
          for (int c = 0; c < NQ; ++c) o.q[c] = a.U[c * Ntot + (size_t)nb * ND + nn];
          derive(g, o);
          double vo = 0.0;
#pragma unroll
          for (int dd = 0; dd < DIM; ++dd) vo += o.v[dd] * n[dd];
          ws = fmax(ws, fabs(vo) + sound_speed(g, o) * nm);
          wmax = fmax(wmax, fabs(1.0 / (a.w0 * a.J[nb])));
        }
        surf = fmax(surf, ws * wmax * a.surf_scale);
      }
  }
  adv = wave_max(adv);
  surf = wave_max(surf);
  diff = wave_max(diff);
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&a.rates[0], (unsigned long long)__double_as_longlong(adv));
    atomicMax(&a.rates[1], (unsigned long long)__double_as_longlong(surf));
    atomicMax(&a.rates[2], (unsigned long long)__double_as_longlong(diff));
  }
}

}  // namespace thx
