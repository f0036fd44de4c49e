// Kernel instantiations: the tabulated LTE gas with the LTEEC flux, and the
// kernels every pairing shares (temperature recovery, BR1 gradient, p_ref).
#include "registry.hpp"

namespace thx {
namespace {
LteRegistrar<2, 2> l22;
LteRegistrar<2, 3> l23;
LteRegistrar<2, 4> l24;
LteRegistrar<3, 1> l31;
LteRegistrar<3, 2> l32;
LteRegistrar<3, 3> l33;
LteRegistrar<3, 4> l34;
}  // namespace
}  // namespace thx
