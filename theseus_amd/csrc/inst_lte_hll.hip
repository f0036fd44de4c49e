// Kernel instantiations: the tabulated LTE gas with the central volume flux and HLL face
#include "registry.hpp"

namespace thx {
namespace {
LteFluxRegistrar<2, 2, LTE_FLUX_HLL> l22;
LteFluxRegistrar<2, 3, LTE_FLUX_HLL> l23;
LteFluxRegistrar<2, 4, LTE_FLUX_HLL> l24;
LteFluxRegistrar<3, 1, LTE_FLUX_HLL> l31;
LteFluxRegistrar<3, 2, LTE_FLUX_HLL> l32;
LteFluxRegistrar<3, 3, LTE_FLUX_HLL> l33;
LteFluxRegistrar<3, 4, LTE_FLUX_HLL> l34;
}  // namespace
}  // namespace thx
