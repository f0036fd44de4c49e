// Kernel instantiations: the tabulated LTE gas with the Pirozzoli flux (volume and face)
#include "registry.hpp"

namespace thx {
namespace {
LteFluxRegistrar<2, 2, LTE_FLUX_PIROZZOLI> l22;
LteFluxRegistrar<2, 3, LTE_FLUX_PIROZZOLI> l23;
LteFluxRegistrar<2, 4, LTE_FLUX_PIROZZOLI> l24;
LteFluxRegistrar<3, 1, LTE_FLUX_PIROZZOLI> l31;
LteFluxRegistrar<3, 2, LTE_FLUX_PIROZZOLI> l32;
LteFluxRegistrar<3, 3, LTE_FLUX_PIROZZOLI> l33;
LteFluxRegistrar<3, 4, LTE_FLUX_PIROZZOLI> l34;
}  // namespace
}  // namespace thx
