// Kernel instantiations: reference 'chandrashekar': EC volume + EC/scalar-dissipation face (SimFactory_impl.hpp:97-99).
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 1, 0, 1> r1;
  FluxRegistrar<DIM, 1, 0, 2> r2;
  FluxRegistrar<DIM, 1, 0, 3> r3;
  FluxRegistrar<DIM, 1, 0, 4> r4;
  FluxRegistrar<DIM, 1, 0, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
