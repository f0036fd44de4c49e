// Boundary conditions on the device: the exterior state of the inviscid
// boundary flux (bc_kernels.hpp:219-287), the slip-wall pressure
// (Flow.hpp:55-170), the viscous wall and prescribed-state terms
// (bc_kernels.hpp:154-216, 289-375) and the BR1 gradient's boundary jump
// (bc_kernels.hpp:32-135).  Boundary points are a small fraction of the face
// points, so these run in FP64 from the FP64 trace wherever the result
// feeds a pressure (both precisions), and in the kernels' compute type for
// the viscous terms.
#pragma once

#include "device_physics.hpp"

namespace thx {

// BCType (bc_cache_utilities.hpp:11-23) and the payload stride per boundary index
enum : int {
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

// Exterior state of the inviscid boundary flux F*(q, q_ext, n_out): on entry
// q holds the interior trace, on return the exterior state.  Outflow and the
// wall types keep the trace (the walls replace the flux afterwards);
// inflow and prescribed state take the boundary's state; symmetry mirrors
// the momentum about the unit normal, m - 2 (m.n^) n^.
template <int DIM>
__device__ __forceinline__ void bc_exterior(int type, const double *data, const double *n, double *q) {
  if (type == BC_SUPERSONIC_INFLOW || type == BC_PRESCRIBED_STATE) {
#pragma unroll
    for (int k = 0; k < DIM + 2; ++k) q[k] = data[k];
  } else if (type == BC_SYMMETRY) {
    double n2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) n2 += n[d] * n[d];
    const double fac = 1.0 / sqrt(n2);
    double un[DIM], nv = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) un[d] = n[d] * fac;
#pragma unroll
    for (int d = 0; d < DIM; ++d) nv += q[1 + d] * un[d];
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[1 + d] = -2.0 * nv * un[d] + q[1 + d];
  }
}

// Wall pressure p* of the 1-D Riemann problem against a wall
// (slipwall_pstar_normal, Flow.hpp:156-170): the state is seen in the wall
// frame (normal component first, tangents as the reference picks them,
// Flow.hpp:77-154); a shock when the flow moves into the wall (v > 0),
// otherwise a rarefaction, floored at 1e-4 of the velocity ratio term.
template <int DIM>
__device__ __forceinline__ double slipwall_pstar(const GasC<double> &g, const double *q, const double *n) {
  double m2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) m2 += n[d] * n[d];
  if (m2 == 0.0) return 0.0;
  const double inv = 1.0 / sqrt(m2);
  const double nx = n[0] * inv, ny = n[1] * inv, nz = DIM > 2 ? n[2] * inv : 0.0;
  double r[DIM];
  r[0] = q[1] * nx + q[2] * ny;
  if constexpr (DIM == 3) r[0] += q[3] * nz;
  if constexpr (DIM == 2) {
    r[1] = q[1] * -ny + q[2] * nx;
  } else {
    double tx, ty, tz;
    const double ax = fabs(nx), ay = fabs(ny), az = fabs(nz);
    if (ax <= ay && ax <= az) {
      tx = 0.0;
      ty = nz;
      tz = -ny;
    } else if (ay <= ax && ay <= az) {
      tx = -nz;
      ty = 0.0;
      tz = nx;
    } else {
      tx = ny;
      ty = -nx;
      tz = 0.0;
    }
    r[1] = q[1] * tx + q[2] * ty + q[3] * tz;
    r[2] = q[1] * (ny * tz - nz * ty) + q[2] * (nz * tx - nx * tz) + q[3] * (nx * ty - ny * tx);
  }
  const double rho = q[0];
  double rm2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) rm2 += r[d] * r[d];
  const double p = g.gammaM1 * (q[DIM + 1] - 0.5 * rm2 / rho);
  const double c = sqrt(g.gamma * p / rho);
  const double v = r[0] / rho;
  const double gP1 = g.gamma + 1.0, gM1 = g.gammaM1;
  const double gP1i = 1.0 / gP1;
  if (v > 0.0) return p + 0.25 * v * gP1 * rho * (v + sqrt(v * v + 8.0 * gP1i * p * (gM1 * gP1i + 1.0) / rho));
  return p * pow(fmax(1.0 + 0.5 * gM1 * v / c, 0.0001), 2.0 * g.gamma * g.gammaM1Inverse);
}

// Does the boundary replace the Riemann flux by (0, p* n, 0)?  Slip walls
// always; the no-slip walls in the Navier-Stokes operator (their viscous
// part is added by the gradient pass).  In the Euler operator the no-slip
// types take the reference's default branch, a zero flux.
__device__ __forceinline__ bool bc_wall_pressure(int type, bool visc) {
  return type == BC_SLIP_WALL || (visc && (type == BC_NO_SLIP_ADIABATIC || type == BC_NO_SLIP_ISOTHERMAL));
}
__device__ __forceinline__ bool bc_zero_flux(int type, bool visc) {
  return !visc && (type == BC_NO_SLIP_ADIABATIC || type == BC_NO_SLIP_ISOTHERMAL);
}

// BR1 boundary jump S_bc - S in entropy variables (ComputeBdrFaceGradFlux,
// bc_kernels.hpp:32-135); zero for the types without one.
template <typename R, int DIM>
__device__ __forceinline__ void bc_grad_jump(const GasC<R> &g, int type, const double *data, const R *S, R *jump) {
#pragma unroll
  for (int k = 0; k < DIM + 2; ++k) jump[k] = R(0);
  if (type == BC_PRESCRIBED_STATE) {
    R qb[DIM + 2], E[DIM + 2];
#pragma unroll
    for (int k = 0; k < DIM + 2; ++k) qb[k] = R(data[k]);
    entropy_state<R, DIM>(g, qb, E);
#pragma unroll
    for (int k = 0; k < DIM + 2; ++k) jump[k] = E[k] - S[k];
  } else if (type == BC_NO_SLIP_ADIABATIC || type == BC_NO_SLIP_ISOTHERMAL) {
    // wall velocity at the node's own beta (adiabatic) or at 1/(R T_wall) (isothermal)
    const R beta = type == BC_NO_SLIP_ADIABATIC ? -S[DIM + 1] : R(1.0 / (double(g.R_gas) * data[DIM]));
#pragma unroll
    for (int d = 0; d < DIM; ++d) jump[1 + d] = R(data[d]) * beta - S[1 + d];
    jump[DIM + 1] = -beta - S[DIM + 1];
  }
}

// Viscous share of the boundary flux, as the right-hand side adds it:
// -(the viscous part of fluxN), with vn[c] = n_out . Fv of momentum c
// (c < DIM, the traction) and of the energy (c = DIM) at the node, u the
// node's velocity.  Prescribed state: the full n.Fv (bc_kernels.hpp:314-323);
// adiabatic wall: the traction and q_wall |n| + v_wall . traction
// (:154-184); isothermal wall: the traction and n.Fv_E - u.traction +
// v_wall . traction (:186-216); zero otherwise.
template <typename R, int DIM>
__device__ __forceinline__ void bc_viscous_share(int type, const double *data, const R *vn, const R *u, const R *n,
                                                 R *out) {
#pragma unroll
  for (int c = 0; c <= DIM; ++c) out[c] = R(0);
  if (type == BC_PRESCRIBED_STATE) {
#pragma unroll
    for (int c = 0; c <= DIM; ++c) out[c] = vn[c];
  } else if (type == BC_NO_SLIP_ADIABATIC) {
    R m2 = R(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) m2 += n[d] * n[d];
    R e = R(data[DIM]) * dsqrt(m2);
#pragma unroll
    for (int d = 0; d < DIM; ++d) e += R(data[d]) * vn[d];
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = vn[d];
    out[DIM] = e;
  } else if (type == BC_NO_SLIP_ISOTHERMAL) {
    R e = vn[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) e -= u[d] * vn[d];
#pragma unroll
    for (int d = 0; d < DIM; ++d) e += R(data[d]) * vn[d];
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = vn[d];
    out[DIM] = e;
  }
}

}  // namespace thx
