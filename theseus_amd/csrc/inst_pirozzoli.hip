// Kernel instantiations: reference 'pirozzoli': KEP volume + KEP/Roe-dissipation face.
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 3, 2, 1> r1;
  FluxRegistrar<DIM, 3, 2, 2> r2;
  FluxRegistrar<DIM, 3, 2, 3> r3;
  FluxRegistrar<DIM, 3, 2, 4> r4;
  FluxRegistrar<DIM, 3, 2, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
