// Kernel instantiations: the tabulated LTE gas with the central volume flux and LLF face
#include "registry.hpp"

namespace thx {
namespace {
LteFluxRegistrar<2, 2, LTE_FLUX_LLF> l22;
LteFluxRegistrar<2, 3, LTE_FLUX_LLF> l23;
LteFluxRegistrar<2, 4, LTE_FLUX_LLF> l24;
LteFluxRegistrar<3, 1, LTE_FLUX_LLF> l31;
LteFluxRegistrar<3, 2, LTE_FLUX_LLF> l32;
LteFluxRegistrar<3, 3, LTE_FLUX_LLF> l33;
LteFluxRegistrar<3, 4, LTE_FLUX_LLF> l34;
}  // namespace
}  // namespace thx
