// Kernel instantiations: Ranocha volume + Lax-Friedrichs face (3-D Taylor-Green workload).
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 2, 5, 1> r1;
  FluxRegistrar<DIM, 2, 5, 2> r2;
  FluxRegistrar<DIM, 2, 5, 3> r3;
  FluxRegistrar<DIM, 2, 5, 4> r4;
  FluxRegistrar<DIM, 2, 5, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
