// Kernel registry: every compiled (dim, order, volume flux, surface flux)
// combination registers its launchers at load time; theseus_rhs_create
// looks the combination up and fails with THESEUS_ERR_UNSUPPORTED when it
// was not compiled.
#pragma once

#include <hip/hip_runtime.h>

#include "rhs_kernels.hpp"

namespace thx {

using RHSLaunch = void (*)(const RHSArgs &, int nblocks, hipStream_t);
using StabLaunch = void (*)(const StabArgs &, int nblocks, hipStream_t);

struct FluxKernels {
  RHSLaunch euler = nullptr;  // fused volume + face
  RHSLaunch ns = nullptr;     // Navier-Stokes pass 2
};

struct ShapeKernels {
  RHSLaunch ns_grad = nullptr;  // Navier-Stokes pass 1
  StabLaunch stability = nullptr;
  int tpb = 0, epb = 0;
};

void register_flux(int dim, int p, int vol, int surf, FluxKernels k);
void register_shape(int dim, int p, ShapeKernels k);
const FluxKernels *find_flux(int dim, int p, int vol, int surf);
const ShapeKernels *find_shape(int dim, int p);

template <int DIM, int P, int VOL, int SURF, bool VISC>
void launch_rhs(const RHSArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_rhs<DIM, P, VOL, SURF, VISC>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}
template <int DIM, int P>
void launch_ns_grad(const RHSArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_ns_grad<DIM, P>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}
template <int DIM, int P>
void launch_stability(const StabArgs &a, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_stability<DIM, P>), dim3(nb), dim3(Shape<DIM, P>::TPB), 0, s, a);
}

template <int DIM, int VOL, int SURF, int P>
struct FluxRegistrar {
  FluxRegistrar() {
    FluxKernels k;
    k.euler = &launch_rhs<DIM, P, VOL, SURF, false>;
    k.ns = &launch_rhs<DIM, P, VOL, SURF, true>;
    register_flux(DIM, P, VOL, SURF, k);
  }
};

template <int DIM, int P>
struct ShapeRegistrar {
  ShapeRegistrar() {
    ShapeKernels k;
    k.ns_grad = &launch_ns_grad<DIM, P>;
    k.stability = &launch_stability<DIM, P>;
    k.tpb = Shape<DIM, P>::TPB;
    k.epb = Shape<DIM, P>::EPB;
    register_shape(DIM, P, k);
  }
};

}  // namespace thx
