// Kernel registry: every compiled (dim, order, volume flux, surface flux)
// combination registers its launchers at load time; theseus_rhs_create
// looks the combination up and fails with THESEUS_ERR_UNSUPPORTED when it
// was not compiled.
#pragma once

#include <hip/hip_runtime.h>

#include "rhs_kernels.hpp"
#include "rhs_lte.hpp"

namespace thx {

using RHSLaunch = void (*)(const RHSArgs &, int nblocks, hipStream_t);
using StabLaunch = void (*)(const StabArgs &, int nblocks, hipStream_t);
using MeasLaunch = void (*)(const MeasArgs &, int nblocks, hipStream_t);

// [precision][bnd]: precision 0 = FP64 arithmetic, 1 = FP32 flux
// arithmetic (mixed precision); bnd 1 = the variant with the boundary
// conditions (meshes with boundary faces), 0 = periodic meshes
struct FluxKernels {
  RHSLaunch euler[2][2] = {};  // fused volume + face
  RHSLaunch ns[2][2] = {};     // Navier-Stokes pass 2
};

struct ShapeKernels {
  RHSLaunch ns_grad[2][2] = {};  // Navier-Stokes pass 1
  StabLaunch stability = nullptr;
  MeasLaunch measures = nullptr;
  int tpb = 0, epb = 0;
};

// Tile shapes of the Euler kernel the serving pool's autotuner chooses from:
// [tile][precision][bnd], tile 0 = the default elements per workgroup, tile 1 = twice that.
struct TileKernels {
  RHSLaunch euler[2][2][2] = {};
  int epb[2] = {0, 0};
  int tpb[2] = {0, 0};
};

// 3-D Euler over orders 2, 3 and 4 in one launch (k_rhs_orders):
// [precision][bnd]; a[i] are the arguments of order i + 2 (its basis is
// basis[i], device memory).
using OrdersLaunch = void (*)(const RHSArgs a[3], const BasisArgs *basis, const int off[2], const int nb[3],
                              hipStream_t);
struct OrdersKernels {
  OrdersLaunch euler[2][2] = {};
};

// The tabulated LTE gas.  [flux]: the pairing (LteFlux, rhs_lte.hpp);
// [precision]: 0 FP64, 1 the mixed-precision form (FP32 flux arithmetic).
// Each instantiation file registers the fields it sets (the rest null).
struct LteKernels {
  RHSLaunch rhs[4][2] = {};        // Euler
  RHSLaunch nodes = nullptr;       // temperature recovery per node (k_lte_nodes, FP64), before rhs
  RHSLaunch ns_grad[2] = {};       // Navier-Stokes pass 1 of the tabulated gas (after nodes)
  RHSLaunch ns[4][2] = {};         // Navier-Stokes pass 2 (after ns_grad)
  RHSLaunch pref = nullptr;        // mixed precision: the background pressure of a state
  int tpb = 0, epb = 0;
};

void register_flux(int dim, int p, int vol, int surf, FluxKernels k);
void register_lte(int dim, int p, LteKernels k);
const LteKernels *find_lte(int dim, int p);
void register_tiles(int dim, int p, int vol, int surf, TileKernels k);
void register_orders(int vol, int surf, OrdersKernels k);
const OrdersKernels *find_orders(int vol, int surf);
const TileKernels *find_tiles(int dim, int p, int vol, int surf);
void register_shape(int dim, int p, ShapeKernels k);
const FluxKernels *find_flux(int dim, int p, int vol, int surf);
const ShapeKernels *find_shape(int dim, int p);

template <typename R, int DIM, int P, int VOL, int SURF, bool VISC, int EPBX = 0, bool BND = false>
void launch_rhs(const RHSArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_rhs<R, DIM, P, VOL, SURF, VISC, EPBX, BND>), dim3(nb), dim3(Shape<DIM, P, EPBX>::TPB), 0, s,
                     a);
}
template <typename R, int DIM, int P, bool BND = false>
void launch_ns_grad(const RHSArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_ns_grad<R, DIM, P, 0, BND>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}
template <int DIM, int P>
void launch_stability(const StabArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_stability<DIM, P>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}

template <int DIM, int P>
void launch_measures(const MeasArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_measures<DIM, P>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}

template <int DIM, int VOL, int SURF, int P>
struct FluxRegistrar {
  FluxRegistrar() {
    FluxKernels k;
    k.euler[0][0] = &launch_rhs<double, DIM, P, VOL, SURF, false>;
    k.ns[0][0] = &launch_rhs<double, DIM, P, VOL, SURF, true>;
    k.euler[1][0] = &launch_rhs<float, DIM, P, VOL, SURF, false>;
    k.ns[1][0] = &launch_rhs<float, DIM, P, VOL, SURF, true>;
    k.euler[0][1] = &launch_rhs<double, DIM, P, VOL, SURF, false, 0, true>;
    k.ns[0][1] = &launch_rhs<double, DIM, P, VOL, SURF, true, 0, true>;
    k.euler[1][1] = &launch_rhs<float, DIM, P, VOL, SURF, false, 0, true>;
    k.ns[1][1] = &launch_rhs<float, DIM, P, VOL, SURF, true, 0, true>;
    register_flux(DIM, P, VOL, SURF, k);
  }
};

template <int DIM, int VOL, int SURF, int P>
struct TileRegistrar {
  TileRegistrar() {
    constexpr int E0 = Shape<DIM, P>::EPB, E1 = 2 * Shape<DIM, P>::EPB;
    TileKernels k;
    k.euler[0][0][0] = &launch_rhs<double, DIM, P, VOL, SURF, false>;
    k.euler[0][1][0] = &launch_rhs<float, DIM, P, VOL, SURF, false>;
    k.euler[1][0][0] = &launch_rhs<double, DIM, P, VOL, SURF, false, E1>;
    k.euler[1][1][0] = &launch_rhs<float, DIM, P, VOL, SURF, false, E1>;
    k.euler[0][0][1] = &launch_rhs<double, DIM, P, VOL, SURF, false, 0, true>;
    k.euler[0][1][1] = &launch_rhs<float, DIM, P, VOL, SURF, false, 0, true>;
    k.euler[1][0][1] = &launch_rhs<double, DIM, P, VOL, SURF, false, E1, true>;
    k.euler[1][1][1] = &launch_rhs<float, DIM, P, VOL, SURF, false, E1, true>;
    k.epb[0] = E0;
    k.epb[1] = E1;
    k.tpb[0] = Shape<DIM, P>::TPB;
    k.tpb[1] = Shape<DIM, P, E1>::TPB;
    register_tiles(DIM, P, VOL, SURF, k);
  }
};

template <typename R, int VOL, int SURF, bool BND>
void launch_rhs_orders(const RHSArgs a[3], const BasisArgs *basis, const int off[2], const int nb[3], hipStream_t s) {
  const RHSCore a2 = a[0], a3 = a[1], a4 = a[2];
  hipLaunchKernelGGL((k_rhs_orders<R, VOL, SURF, BND>), dim3(off[1] + nb[2]), dim3(128), 0, s, a2, a3, a4, basis,
                     off[0], off[1], nb[0], nb[1], nb[2]);
}

template <int VOL, int SURF>
struct OrdersRegistrar {
  OrdersRegistrar() {
    OrdersKernels k;
    k.euler[0][0] = &launch_rhs_orders<double, VOL, SURF, false>;
    k.euler[1][0] = &launch_rhs_orders<float, VOL, SURF, false>;
    k.euler[0][1] = &launch_rhs_orders<double, VOL, SURF, true>;
    k.euler[1][1] = &launch_rhs_orders<float, VOL, SURF, true>;
    register_orders(VOL, SURF, k);
  }
};

template <typename R, int DIM, int P, bool VISC = false, int FX = LTE_FLUX_LTEEC>
void launch_rhs_lte(const RHSArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_rhs_lte<R, DIM, P, VISC, FX>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}
template <typename R, int DIM, int P>
void launch_ns_grad_lte(const RHSArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_ns_grad<R, DIM, P, 0, false, true>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}
template <int DIM>
void launch_pref_lte(const RHSArgs &a, int, hipStream_t s) {
  hipLaunchKernelGGL((lte::k_pref_lte<DIM>), dim3(1), dim3(64), 0, s, a.U, a.Ntot, a.lte, const_cast<double *>(a.pref));
}

template <int DIM>
void launch_lte_nodes(const RHSArgs &a, int, hipStream_t s) {
  const dim3 grid((unsigned)((a.Ntot + 255) / 256));
  hipLaunchKernelGGL((lte::k_lte_nodes<DIM>), grid, dim3(256), 0, s, a.U, a.Ntot, a.lte);
  if (a.lte.node_p) hipLaunchKernelGGL((lte::k_lte_node_pcg<DIM>), grid, dim3(256), 0, s, a.U, a.Ntot, a.lte);
}

template <int DIM, int P>
struct LteRegistrar {
  LteRegistrar() {
    LteKernels k;
    k.rhs[LTE_FLUX_LTEEC][0] = &launch_rhs_lte<double, DIM, P>;
    k.rhs[LTE_FLUX_LTEEC][1] = &launch_rhs_lte<float, DIM, P>;
    k.nodes = &launch_lte_nodes<DIM>;
    k.ns_grad[0] = &launch_ns_grad_lte<double, DIM, P>;
    k.ns_grad[1] = &launch_ns_grad_lte<float, DIM, P>;
    k.ns[LTE_FLUX_LTEEC][0] = &launch_rhs_lte<double, DIM, P, true>;
    k.ns[LTE_FLUX_LTEEC][1] = &launch_rhs_lte<float, DIM, P, true>;
    k.pref = &launch_pref_lte<DIM>;
    k.tpb = Shape<DIM, P>::TPB;
    k.epb = Shape<DIM, P>::EPB;
    register_lte(DIM, P, k);
  }
};

// The Pirozzoli, HLL and LLF pairings of the tabulated gas
template <int DIM, int P, int FX>
struct LteFluxRegistrar {
  LteFluxRegistrar() {
    LteKernels k;
    k.rhs[FX][0] = &launch_rhs_lte<double, DIM, P, false, FX>;
    k.rhs[FX][1] = &launch_rhs_lte<float, DIM, P, false, FX>;
    k.ns[FX][0] = &launch_rhs_lte<double, DIM, P, true, FX>;
    k.ns[FX][1] = &launch_rhs_lte<float, DIM, P, true, FX>;
    register_lte(DIM, P, k);
  }
};

template <int DIM, int P>
struct ShapeRegistrar {
  ShapeRegistrar() {
    ShapeKernels k;
    k.ns_grad[0][0] = &launch_ns_grad<double, DIM, P>;
    k.ns_grad[1][0] = &launch_ns_grad<float, DIM, P>;
    k.ns_grad[0][1] = &launch_ns_grad<double, DIM, P, true>;
    k.ns_grad[1][1] = &launch_ns_grad<float, DIM, P, true>;
    k.stability = &launch_stability<DIM, P>;
    k.measures = &launch_measures<DIM, P>;
    k.tpb = Shape<DIM, P>::TPB;
    k.epb = Shape<DIM, P>::EPB;
    register_shape(DIM, P, k);
  }
};

}  // namespace thx
