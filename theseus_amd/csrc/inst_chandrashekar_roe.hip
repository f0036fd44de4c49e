// Kernel instantiations: Chandrashekar volume + Roe face (2-D isentropic vortex workload).
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 1, 3, 1> r1;
  FluxRegistrar<DIM, 1, 3, 2> r2;
  FluxRegistrar<DIM, 1, 3, 3> r3;
  FluxRegistrar<DIM, 1, 3, 4> r4;
  FluxRegistrar<DIM, 1, 3, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
